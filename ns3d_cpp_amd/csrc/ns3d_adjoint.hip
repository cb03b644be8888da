// ns3d_adjoint.hip -- tangent-linear (TLM) and discrete adjoint of the FP64
// RHS (component-wise or characteristic reconstruction; periodic or
// wall / symmetry / outflow / inflow sides), and the dissipation objective.
//
// Notation: W = (rho, u, v, w, p) with T = p / (rho R); U = conserved.
// R(U) = R_inv + R_vis as in range_kutta.cpp:11-99.  With periodic ghosts
// every face of the reference's face range is a periodic face f (right face
// of cell f, left face of cell f+1), so the adjoint of the flux difference is
//     lambda_F(f) = (L(f+1) - L(f)) / h
// and the central differences satisfy D^T = -D.  All adjoint arrays hold
// interior cells; periodic images are reached by index wrap.  A non-periodic
// axis adds the first cell's left face (lambda = 0 outside the box), the
// boundary fill's ghosts (the transpose of the fill maps their adjoints back
// to the cells they copy) and the closures of choose_scheme, whose
// transpose is applied row by row (dT_closure).
#include "ns3d_ad.hpp"
#include "ns3d_internal.hpp"

namespace ns3d {

__device__ __forceinline__ int awrap(int q, int ng, int n)
{
    // any q in [ng - n, ng + 2n)
    return q < ng ? q + n : (q >= ng + n ? q - n : q);
}

// central difference of a periodic interior array at (i,j,k) along `axis`
template <int VIS>
__device__ __forceinline__ double dwrap(const DevParams &P, const double *a, int i, int j, int k, int axis, double h)
{
    constexpr int R = VisReach<VIS>::R;
    double win[2 * R + 1];
#pragma unroll
    for (int o = -R; o <= R; ++o) {
        int ii = i, jj = j, kk = k;
        if (axis == 0) ii = awrap(i + o, P.ng, P.nx);
        else if (axis == 1) jj = awrap(j + o, P.ng, P.ny);
        else kk = awrap(k + o, P.ng, P.nz);
        win[o + R] = a[(kk * P.sy + jj) * P.sx + ii];
    }
    return diff_window<VIS>(win + R, h);
}

// derivative along `axis` at (i,j,k): the periodic central difference
// (index wrap), or on a non-periodic axis the closures of choose_scheme
// (ViscousFlux.cpp:43-118), which read interior values only
template <int VIS>
__device__ __forceinline__ double dgen(const DevParams &P, const double *a, int i, int j, int k, int axis, double h)
{
    if (P.periodic[axis]) return dwrap<VIS>(P, a, i, j, k, axis, h);
    const int q = axis == 0 ? i : axis == 1 ? j : k;
    const int n = axis == 0 ? P.nx : axis == 1 ? P.ny : P.nz;
    const long st = axis == 0 ? 1 : axis == 1 ? P.sx : (long)P.sx * P.sy;
    return diff_central_bc<VIS>(a + (k * P.sy + j) * P.sx + i, st, h, q - P.ng, P.ng + n - 1 - q);
}

// (D^T a)(i,j,k) along a non-periodic `axis`: over the interior cells r
// within reach, the weight of this cell in r's closure row times a(r).  The
// weight is r's own row formula applied to a unit vector (exact: each row
// is linear in its window).
template <int VIS>
__device__ __forceinline__ double dT_closure(const DevParams &P, const double *a, int i, int j, int k, int axis,
                                             double h)
{
    constexpr int R = VisReach<VIS>::R;
    const int n = axis == 0 ? P.nx : axis == 1 ? P.ny : P.nz;
    const int q = (axis == 0 ? i : axis == 1 ? j : k) - P.ng;
    const long st = axis == 0 ? 1 : axis == 1 ? P.sx : (long)P.sx * P.sy;
    const double *c = a + (k * P.sy + j) * P.sx + i;
    double s = 0.0;
#pragma unroll
    for (int o = -R; o <= R; ++o) {
        const int r = q - o;   // the row's cell; this cell sits at its offset o
        if (r < 0 || r >= n) continue;
        double e[2 * R + 1];
#pragma unroll
        for (int t = 0; t < 2 * R + 1; ++t) e[t] = t == R + o ? 1.0 : 0.0;
        s += diff_window_bc<VIS>(e + R, h, r, n - 1 - r) * c[-o * st];
    }
    return s;
}

// index of the cell (i,j,k) with its coordinate along DIR replaced by q
template <int DIR>
__device__ __forceinline__ int along(const DevParams &P, int i, int j, int k, int q)
{
    return DIR == 0 ? (k * P.sy + j) * P.sx + q : DIR == 1 ? (k * P.sy + q) * P.sx + i : (q * P.sy + j) * P.sx + i;
}

__device__ __forceinline__ double dmu_dT(const DevParams &P, double T, double mu)
{
    if (P.mu_law == 1) return 0.0;
    if (P.mu_law == 2) return mu * P.mu_power / T;
    return mu * (1.5 / T - 1.0 / (T + P.sutherland_s));
}

// Periodic ghost fill (face regions, cross directions over interior cells)
// of up to 12 adjoint arrays, for the central-difference stencils.
struct ArraySet {
    double *p[12];
    int n;
};

__global__ void ka_fill(DevParams P, ArraySet A)
{
    const int axis = blockIdx.y;
    const int ng = P.ng;
    const int nd = axis == 0 ? P.nx : axis == 1 ? P.ny : P.nz;
    const int na = axis == 0 ? P.ny : P.nx;
    const int nb = axis == 2 ? P.ny : P.nz;
    const int total = 2 * ng * na * nb;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int g = t % (2 * ng);
    const int rest = t / (2 * ng);
    const int a = ng + rest % na;
    const int b = ng + rest / na;
    int q, qs;
    if (g < ng) { q = g; qs = g + nd; }
    else { q = ng + nd + (g - ng); qs = ng + (g - ng); }
    int dst, src;
    if (axis == 0) { dst = (b * P.sy + a) * P.sx + q; src = (b * P.sy + a) * P.sx + qs; }
    else if (axis == 1) { dst = (b * P.sy + q) * P.sx + a; src = (b * P.sy + qs) * P.sx + a; }
    else { dst = (q * P.sy + b) * P.sx + a; src = (qs * P.sy + b) * P.sx + a; }
    for (int m = 0; m < A.n; ++m) A.p[m][dst] = A.p[m][src];
}

static void fill_set(hipStream_t s, const DevParams &P, const ArraySet &A)
{
    int m = P.ny * P.nz;
    if (P.nx * P.nz > m) m = P.nx * P.nz;
    if (P.nx * P.ny > m) m = P.nx * P.ny;
    const int total = 2 * P.ng * m;
    hipLaunchKernelGGL(ka_fill, dim3((total + 255) / 256, 3), dim3(256), 0, s, P, A);
}


// -------------------------------------------------------------- utilities

// TLM: dW = C2P'(U) dU at every interior cell
__global__ void kt_c2p(DevParams P, Fields F, AdjArrays A)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double rho = F.W[0][id], u = F.W[1][id], v = F.W[2][id], w = F.W[3][id];
    const double d0 = A.X[0][id], d1 = A.X[1][id], d2 = A.X[2][id], d3 = A.X[3][id], d4 = A.X[4][id];
    A.Y[0][id] = d0;
    A.Y[1][id] = (d1 - u * d0) / rho;
    A.Y[2][id] = (d2 - v * d0) / rho;
    A.Y[3][id] = (d3 - w * d0) / rho;
    A.Y[4][id] = (P.gamma - 1.0) * (d4 - (u * d1 + v * d2 + w * d3) + 0.5 * (u * u + v * v + w * w) * d0);
}

// linearised ghost fill (boundary.cpp:58-207) of the non-periodic side
// `face` for the tangent dW in Y: wall and symmetry mirror it with the normal
// velocity negated, outflow repeats the last interior cell, inflow writes a
// constant state (tangent 0); the fill's density clamp has derivative 1 at
// the positive densities of a valid state
__global__ void kt_fill_bc(DevParams P, AdjArrays A, int face, int kind)
{
    const int axis = face >> 1;
    const int na = axis == 0 ? P.sy : P.sx;
    const int nb = axis == 2 ? P.sy : P.sz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.ng * na * nb) return;
    int dst, src, normal;
    bc_cell(P, face, kind, t, dst, src, normal);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double d = A.Y[m][src];
        A.Y[m][dst] = kind == BCK_INFLOW ? 0.0 : (m == normal && kind != BCK_OUTFLOW ? -d : d);
    }
}

// ------------------------------------------------------------------ TLM

// tangent of the face flux of the right face of every interior cell along
// DIR, from W (ghost-filled) and dW (interior, wrapped): dF[m] arrays.
// CHAR: the characteristic form (du_inviscous_character,
// InviscousFlux_new.cpp:165-437): the split fluxes projected on the left
// eigenvectors of the face's Roe state, reconstructed per characteristic
// field and projected back, with the Roe state's tangent included
template <int DIR, int RECON, int FVS, int CHAR>
__global__ __launch_bounds__(256) void kt_faces(DevParams P, Fields F, AdjArrays A)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int c = DIR == 0 ? i : DIR == 1 ? j : k;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const bool per = P.periodic[DIR];
    using D1 = Dual<1>;
    // the cell's right face; on a non-periodic axis the first cell also
    // forms its left face, from the ghost-filled W and dW
    const int nf = !per && c == P.ng ? 2 : 1;
    for (int t = 0; t < nf; ++t) {
        const int fc = c - t;   // the face's left cell
        D1 fp[S][5], fm[S][5], wl[5], wr[5];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int id = along<DIR>(P, i, j, k, per ? awrap(fc - MID + s, P.ng, n) : fc - MID + s);
            D1 x[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) { x[m].v = F.W[m][id]; x[m].d[0] = A.Y[m][id]; }
            split_t<FVS, DIR>(P, x[0], x[1], x[2], x[3], x[4], fp[s], fm[s]);
            if (CHAR && s == MID)
#pragma unroll
                for (int m = 0; m < 5; ++m) wl[m] = x[m];
            if (CHAR && s == MID + 1)
#pragma unroll
                for (int m = 0; m < 5; ++m) wr[m] = x[m];
        }
        const int id = along<DIR>(P, i, j, k, fc);
        if constexpr (CHAR) {
            D1 Lm[5][5], Rm[5][5], wc[5];
            char_matrices_t(P.gamma, wl, wr, DIR == 0 ? 1.0 : 0.0, DIR == 1 ? 1.0 : 0.0, DIR == 2 ? 1.0 : 0.0, Lm,
                            Rm);
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                D1 sp[S], sm[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    D1 ap = dconst<1>(0.0), am = dconst<1>(0.0);
#pragma unroll
                    for (int r = 0; r < 5; ++r) { ap = ap + Lm[e][r] * fp[s][r]; am = am + Lm[e][r] * fm[s][r]; }
                    sp[s] = ap;
                    sm[S - 1 - s] = am;
                }
                wc[e] = recon_t<RECON>(P, sp) + recon_t<RECON>(P, sm);
            }
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                D1 r = dconst<1>(0.0);
#pragma unroll
                for (int e = 0; e < 5; ++e) r = r + Rm[m][e] * wc[e];
                A.T[m][id] = r.d[0];
            }
        } else {
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                D1 sp[S], sm[S];
#pragma unroll
                for (int s = 0; s < S; ++s) { sp[s] = fp[s][m]; sm[S - 1 - s] = fm[s][m]; }
                D1 r = recon_t<RECON>(P, sp) + recon_t<RECON>(P, sm);
                A.T[m][id] = r.d[0];
            }
        }
    }
}

// dR_inv -= (dF(c) - dF(c-1)) / h along DIR (accumulated into Z)
template <int DIR>
__global__ void kt_diff(DevParams P, AdjArrays A, int first)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    // the left face: wrapped on a periodic axis, else the first cell's left
    // face kept at its ghost neighbour
    const int c = DIR == 0 ? i : DIR == 1 ? j : k;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const int il = along<DIR>(P, i, j, k, P.periodic[DIR] ? awrap(c - 1, P.ng, n) : c - 1);
    const double h = DIR == 0 ? P.dx : DIR == 1 ? P.dy : P.dz;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double dr = -(A.T[m][id] - A.T[m][il]) / h;
        A.Z[m][id] = first ? dr : A.Z[m][id] + dr;
    }
}

// tangent of the viscous fluxes (12 arrays into VT)
template <int VIS>
__global__ __launch_bounds__(256) void kt_vflux(DevParams P, Fields F, AdjArrays A)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    using D1 = Dual<1>;
    const double h[3] = {P.dx, P.dy, P.dz};
    D1 g[12];
    // tangent of T = p / (rho R) in the scratch A.T[0] (written by kt_tc)
    const double *vals[4] = {F.W[1], F.W[2], F.W[3], F.W[5]};
    const double *tans[4] = {A.Y[1], A.Y[2], A.Y[3], A.T[0]};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            g[3 * q + d].v = dgen<VIS>(P, vals[q], i, j, k, d, h[d]);
            g[3 * q + d].d[0] = dgen<VIS>(P, tans[q], i, j, k, d, h[d]);
        }
    D1 T;
    T.v = F.W[5][id];
    T.d[0] = A.T[0][id];
    D1 mu;
    if (P.mu_law == 1) mu = dconst<1>(P.mu);
    else if (P.mu_law == 2) mu = P.mu * pow_t(T, P.mu_power);
    else mu = P.mu * pow_t(T, 1.5) * (1.0 + P.sutherland_s) / (T + P.sutherland_s);
    D1 u, v, w;
    u.v = F.W[1][id]; u.d[0] = A.Y[1][id];
    v.v = F.W[2][id]; v.d[0] = A.Y[2][id];
    w.v = F.W[3][id]; w.d[0] = A.Y[3][id];
    const double kq = P.Cp / P.Pr;
    D1 divU = g[0] + g[4] + g[8];
    D1 txx = 2.0 * mu * g[0] - (2.0 / 3.0) * mu * divU;
    D1 tyy = 2.0 * mu * g[4] - (2.0 / 3.0) * mu * divU;
    D1 tzz = 2.0 * mu * g[8] - (2.0 / 3.0) * mu * divU;
    D1 txy = mu * (g[1] + g[3]);
    D1 txz = mu * (g[2] + g[6]);
    D1 tyz = mu * (g[5] + g[7]);
    D1 mk = mu * kq;
    D1 vf[12] = {txx, txy, txz, u * txx + v * txy + w * txz + mk * g[9],
                 txy, tyy, tyz, u * txy + v * tyy + w * tyz + mk * g[10],
                 txz, tyz, tzz, u * txz + v * tyz + w * tzz + mk * g[11]};
#pragma unroll
    for (int m = 0; m < 12; ++m) A.V[m][id] = vf[m].d[0];
}

// tangent of T into A.T[0]
__global__ void kt_tc(DevParams P, Fields F, AdjArrays A)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double rho = F.W[0][id], T = F.W[5][id];
    A.T[0][id] = A.Y[4][id] / (rho * P.Rgas) - T * A.Y[0][id] / rho;
}

template <int VIS>
__global__ void kt_vdiv(DevParams P, AdjArrays A)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double h[3] = {P.dx, P.dy, P.dz};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) s += dgen<VIS>(P, A.V[4 * d + m], i, j, k, d, h[d]);
        A.Z[m + 1][id] += s;
    }
}

// --------------------------------------------------------------- adjoint

// Inviscid adjoint along DIR with one flux component per lane: lanes 5g..5g+4
// of a wave march the same column (group g, 12 columns per wave), lane m
// keeping the S-cell windows of F+_m, F-_m and of their adjoints.  Each cell
// is owned by the segment that contains it; when a cell leaves the window,
// every lane applies the reverse rule of its component of the split and the
// 5 lanes of the group add their (rho,u,v,w,p) contributions through LDS.
constexpr int kAdjCols = 12;

template <int DIR, int RECON, int FVS>
__global__ __launch_bounds__(64) void ka_sweepc(DevParams P, Fields F, AdjArrays A, int seg_len)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    __shared__ double sh[kAdjCols + 1][5][5];
    const int lane = threadIdx.x;
    const int g = lane / 5;
    const int m = lane - 5 * g;          // lanes 60..63: g = 12, m = 0..3
    const int na = DIR == 0 ? P.ny : P.nx;
    const int a = blockIdx.x * kAdjCols + g;
    const bool valid = g < kAdjCols && a < na;
    const int ac = valid ? a : 0;        // inactive lanes read column 0
    const int b = blockIdx.y;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const int c0 = P.ng + blockIdx.z * seg_len;
    const int c1 = min(c0 + seg_len, P.ng + n);
    if (c0 >= c1) return;                // uniform over the block
    int base, st;
    if (DIR == 0) { base = ((P.ng + b) * P.sy + (P.ng + ac)) * P.sx; st = 1; }
    else if (DIR == 1) { base = ((P.ng + b) * P.sy) * P.sx + (P.ng + ac); st = P.sx; }
    else { base = (P.ng + b) * P.sx + (P.ng + ac); st = P.sx * P.sy; }
    const double inv_d = DIR == 0 ? P.idx : DIR == 1 ? P.idy : P.idz;
    auto cid = [&](int q) { return base + awrap(q, P.ng, n) * st; };
    const double *Xm = A.X[m];

    double fp[S], fm[S], fpb[S], fmb[S];
    const int fA = c0 - S + 1 + MID, fB = c1 - 1 + MID;
#pragma unroll
    for (int s = 1; s < S; ++s) {
        const int id = cid(fA - 1 - MID + s);
        split_comp<FVS, DIR>(P, F.W[0][id], F.W[1][id], F.W[2][id], F.W[3][id], F.W[4][id], m, fp[s], fm[s]);
        fpb[s] = 0.0;
        fmb[s] = 0.0;
    }
    double ll = Xm[cid(fA)];
    // loads of the next iteration are issued one face ahead
    double nw[5], nxv, qw[5];
    {
        const int id = cid(fA - MID + S - 1);
        const int iq = cid(fA - MID);
#pragma unroll
        for (int t = 0; t < 5; ++t) { nw[t] = F.W[t][id]; qw[t] = F.W[t][iq]; }
        nxv = Xm[cid(fA + 1)];
    }
#pragma unroll 1
    for (int f = fA; f <= fB; ++f) {
#pragma unroll
        for (int s = 0; s < S - 1; ++s) {
            fp[s] = fp[s + 1]; fm[s] = fm[s + 1];
            fpb[s] = fpb[s + 1]; fmb[s] = fmb[s + 1];
        }
        const double w0 = nw[0], w1 = nw[1], w2 = nw[2], w3 = nw[3], w4 = nw[4];
        const double lr = nxv;
        {
            const int id = cid(f + 1 - MID + S - 1);
#pragma unroll
            for (int t = 0; t < 5; ++t) nw[t] = F.W[t][id];
            nxv = Xm[cid(f + 2)];
        }
        split_comp<FVS, DIR>(P, w0, w1, w2, w3, w4, m, fp[S - 1], fm[S - 1]);
        fpb[S - 1] = 0.0;
        fmb[S - 1] = 0.0;
        const double gF = (lr - ll) * inv_d;
        ll = lr;
        {
            double sm[S], gb[S];
            recon_vjp<RECON>(P, fp, gF, gb);
#pragma unroll
            for (int s = 0; s < S; ++s) fpb[s] += gb[s];
#pragma unroll
            for (int s = 0; s < S; ++s) sm[S - 1 - s] = fm[s];
            recon_vjp<RECON>(P, sm, gF, gb);
#pragma unroll
            for (int s = 0; s < S; ++s) fmb[s] += gb[S - 1 - s];
        }
        const int q = f - MID;
        const double q0 = qw[0], q1 = qw[1], q2 = qw[2], q3 = qw[3], q4 = qw[4];
        {
            const int iq = cid(q + 1);
#pragma unroll
            for (int t = 0; t < 5; ++t) qw[t] = F.W[t][iq];
        }
        if (q >= c0 && q < c1) {
            const int id = base + q * st;
            double gv[5];
            split_comp_vjp<FVS, DIR>(P, q0, q1, q2, q3, q4, m, fpb[0], fmb[0], gv);
#pragma unroll
            for (int t = 0; t < 5; ++t) sh[g][m][t] = gv[t];
            __syncthreads();
            if (valid) {
                const double sum = (sh[g][0][m] + sh[g][1][m]) + (sh[g][2][m] + sh[g][3][m]) + sh[g][4][m];
                A.Y[m][id] += sum;
            }
            __syncthreads();
        }
    }
}

// Inviscid adjoint along DIR with whole periodic lines in LDS: a block holds
// TX = blockDim.x neighbouring lines (x-columns for DIR = 1, 2; a single
// x-row for DIR = 0) and one thread per cell, threadIdx.y along the line.
// Per flux component and side: the split values of the lines go to LDS;
// thread f forms lambda_F of its face (the face right of cell f) and the
// reverse rule of the reconstruction, leaving the contribution to every
// stencil slot in LDS; thread q then sums the S contributions that land on
// its cell.  After the 5 components the reverse rule of the full split at the
// cell adds into Y (store = 1 writes Y instead).  Loads and stores are
// coalesced over the TX lines, nothing is recomputed, and LDS holds
// (1 + S) * n * TX doubles.
template <int DIR, int RECON, int FVS>
__global__ __launch_bounds__(1024) void ka_sweep_line(DevParams P, Fields F, AdjArrays A, int store)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    extern __shared__ double lds[];
    const int TX = blockDim.x;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const int tx = threadIdx.x, q = threadIdx.y;
    const int L = n * TX;                // one LDS plane
    // two buffers of [n][TX] split values of one side + [S][n][TX] reverse
    // contributions, alternating between the ten (component, side) passes:
    // a pass writes the buffer the previous pass did not touch, so it needs
    // no barrier after its gather (two per pass instead of three)
    double *const buf0 = lds;
    double *const buf1 = lds + (size_t)(1 + S) * L;
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    int base, st, ncols;
    if (DIR == 0) {
        base = ((P.ng + by) * P.sy + (P.ng + bx)) * P.sx + P.ng;
        st = 1;
        ncols = 1;
    } else if (DIR == 1) {
        base = ((P.ng + by) * P.sy + P.ng) * P.sx + P.ng + bx * TX;
        st = P.sx;
        ncols = P.nx - bx * TX;
    } else {
        base = (P.ng * P.sy + (P.ng + by)) * P.sx + P.ng + bx * TX;
        st = P.sx * P.sy;
        ncols = P.nx - bx * TX;
    }
    const bool act = q < n && tx < ncols;
    const int qr = act ? q : 0, txr = act ? tx : 0;
    const int qn = qr + 1 == n ? 0 : qr + 1;
    const int id = base + txr + qr * st;
    const int idn = base + txr + qn * st;
    const int me = q * TX + tx;
    double w[5], fp[5], fm[5], lp[5], lm[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) w[t] = F.W[t][id];
    split_t<FVS, DIR>(P, w[0], w[1], w[2], w[3], w[4], fp, fm);
    const double inv_d = DIR == 0 ? P.idx : DIR == 1 ? P.idy : P.idz;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double gF = (A.X[m][idn] - A.X[m][id]) * inv_d;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            double *const spl = side == 0 ? buf0 : buf1;
            double *const con = spl + L;
            if (act) spl[me] = side == 0 ? fp[m] : fm[m];
            __syncthreads();
            if (act) {
                double sv[S], gb[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    int c = q - MID + s;
                    c = c < 0 ? c + n : (c >= n ? c - n : c);
                    // the minus side reads its stencil reversed
                    sv[side == 0 ? s : S - 1 - s] = spl[c * TX + tx];
                }
                recon_vjp<RECON>(P, sv, gF, gb);
#pragma unroll
                for (int s = 0; s < S; ++s) con[s * L + me] = gb[side == 0 ? s : S - 1 - s];
            }
            __syncthreads();
            if (act) {
                double a = 0.0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    int f = q + MID - s;   // face whose slot s is this cell
                    f = f < 0 ? f + n : (f >= n ? f - n : f);
                    a += con[s * L + f * TX + tx];
                }
                if (side == 0) lp[m] = a;
                else lm[m] = a;
            }
        }
    }
    if (!act) return;
    double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        double gv[5];
        split_comp_vjp<FVS, DIR>(P, w[0], w[1], w[2], w[3], w[4], m, lp[m], lm[m], gv);
#pragma unroll
        for (int t = 0; t < 5; ++t) g[t] += gv[t];
    }
    if (store) {
#pragma unroll
        for (int t = 0; t < 5; ++t) A.Y[t][id] = g[t];
    } else {
#pragma unroll
        for (int t = 0; t < 5; ++t) A.Y[t][id] += g[t];
    }
}


// Inviscid adjoint along DIR by faces, for the characteristic form
// (du_inviscous_character, InviscousFlux_new.cpp:165-437; CHAR = 1) and for
// non-periodic axes (either form).  ka_face reverses one face: the right
// face of every interior cell, and on a non-periodic axis also the first
// cell's left face.  lambda_F(f) = (lambda(f+1) - lambda(f)) / h, where a
// non-periodic axis has lambda = 0 outside the box.  The component-wise form
// reverses each component's two reconstructions; the characteristic form
// recomputes the face (Roe state, eigenvector matrices, the split fluxes
// projected on the left eigenvectors, the reconstruction of each field) and
// reverses lambda_wc = Rm^T lambda_F, each field's reconstructions and the
// projections.  Written per face, at its left cell's index: for every
// stencil slot s the adjoints of F+ and F- at that slot's cell (arrays 5 s + m
// and 5 (S + s) + m of B); CHAR also writes the adjoint through the Roe state
// of the two cells' primitives (arrays 10 S + t for the left cell, 10 S + 5 +
// t for the right), by forward mode over those ten inputs.  ka_gather sums
// them per cell in a fixed order.
template <int DIR, int RECON, int FVS, int CHAR>
__device__ __forceinline__ void face_adjoint(const DevParams &P, const Fields &F, const AdjArrays &A,
                                             const CharBuf &B, int i, int j, int k, int fc)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const bool per = P.periodic[DIR];
    int ids[S];
#pragma unroll
    for (int s = 0; s < S; ++s) ids[s] = along<DIR>(P, i, j, k, per ? awrap(fc - MID + s, P.ng, n) : fc - MID + s);
    const int id = ids[MID];
    const double inv_d = DIR == 0 ? P.idx : DIR == 1 ? P.idy : P.idz;
    // lambda of the face's two cells (0 outside a non-periodic box)
    const bool in_l = per || fc >= P.ng, in_r = per || fc + 1 < P.ng + n;
    const int idr = along<DIR>(P, i, j, k, per ? awrap(fc + 1, P.ng, n) : fc + 1);
    double lf[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) lf[m] = ((in_r ? A.X[m][idr] : 0.0) - (in_l ? A.X[m][id] : 0.0)) * inv_d;
    const size_t N = (size_t)B.n;
    if constexpr (!CHAR) {
        double fp[S][5], fm[S][5];
#pragma unroll
        for (int s = 0; s < S; ++s)
            split_t<FVS, DIR>(P, F.W[0][ids[s]], F.W[1][ids[s]], F.W[2][ids[s]], F.W[3][ids[s]], F.W[4][ids[s]],
                              fp[s], fm[s]);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sv[S], g[S];
#pragma unroll
            for (int s = 0; s < S; ++s) sv[s] = fp[s][m];
            recon_vjp<RECON>(P, sv, lf[m], g);
#pragma unroll
            for (int s = 0; s < S; ++s) B.base[(size_t)(5 * s + m) * N + id] = g[s];
#pragma unroll
            for (int s = 0; s < S; ++s) sv[S - 1 - s] = fm[s][m];
            recon_vjp<RECON>(P, sv, lf[m], g);
#pragma unroll
            for (int s = 0; s < S; ++s) B.base[(size_t)(5 * (S + s) + m) * N + id] = g[S - 1 - s];
        }
    } else {
        const double nv[3] = {DIR == 0 ? 1.0 : 0.0, DIR == 1 ? 1.0 : 0.0, DIR == 2 ? 1.0 : 0.0};
        double wl[5], wr[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) { wl[m] = F.W[m][id]; wr[m] = F.W[m][ids[MID + 1]]; }
        double Lm[5][5], Rm[5][5];
        char_matrices_t(P.gamma, wl, wr, nv[0], nv[1], nv[2], Lm, Rm);
        // the projected split fluxes: pp[e] in stencil order, pm[e] reversed
        double pp[5][S], pm[5][S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double fp[5], fm[5];
            split_t<FVS, DIR>(P, F.W[0][ids[s]], F.W[1][ids[s]], F.W[2][ids[s]], F.W[3][ids[s]], F.W[4][ids[s]], fp,
                              fm);
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                double ap = 0.0, am = 0.0;
#pragma unroll
                for (int r = 0; r < 5; ++r) { ap += Lm[e][r] * fp[r]; am += Lm[e][r] * fm[r]; }
                pp[e][s] = ap;
                pm[e][S - 1 - s] = am;
            }
        }
        double wc[5], lwc[5];
#pragma unroll
        for (int e = 0; e < 5; ++e) wc[e] = recon_t<RECON>(P, pp[e]) + recon_t<RECON>(P, pm[e]);
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            double a = 0.0;
#pragma unroll
            for (int m = 0; m < 5; ++m) a += Rm[m][e] * lf[m];
            lwc[e] = a;
        }
        // reverse rules of the reconstructions, in place: pp / pm now hold
        // the adjoints of the projected stencil values
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            double g[S];
            recon_vjp<RECON>(P, pp[e], lwc[e], g);
#pragma unroll
            for (int s = 0; s < S; ++s) pp[e][s] = g[s];
            recon_vjp<RECON>(P, pm[e], lwc[e], g);
#pragma unroll
            for (int s = 0; s < S; ++s) pm[e][s] = g[s];
        }
        // adjoint of Lm, and of F+ / F- at every slot
        double lL[5][5];
#pragma unroll
        for (int e = 0; e < 5; ++e)
#pragma unroll
            for (int r = 0; r < 5; ++r) lL[e][r] = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double fp[5], fm[5];
            split_t<FVS, DIR>(P, F.W[0][ids[s]], F.W[1][ids[s]], F.W[2][ids[s]], F.W[3][ids[s]], F.W[4][ids[s]], fp,
                              fm);
#pragma unroll
            for (int e = 0; e < 5; ++e)
#pragma unroll
                for (int r = 0; r < 5; ++r) lL[e][r] += pp[e][s] * fp[r] + pm[e][S - 1 - s] * fm[r];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                double bp = 0.0, bm = 0.0;
#pragma unroll
                for (int e = 0; e < 5; ++e) { bp += Lm[e][m] * pp[e][s]; bm += Lm[e][m] * pm[e][S - 1 - s]; }
                B.base[(size_t)(5 * s + m) * N + id] = bp;
                B.base[(size_t)(5 * (S + s) + m) * N + id] = bm;
            }
        }
        // the Roe state's eigenvector matrices, forward mode over the ten inputs
        for (int q = 0; q < 10; ++q) {
            Dual<1> dl[5], dr[5], dL[5][5], dR[5][5];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                dl[m] = dconst<1>(wl[m]);
                dr[m] = dconst<1>(wr[m]);
                if (q == m) dl[m].d[0] = 1.0;
                if (q == m + 5) dr[m].d[0] = 1.0;
            }
            char_matrices_t(P.gamma, dl, dr, nv[0], nv[1], nv[2], dL, dR);
            double a = 0.0;
#pragma unroll
            for (int e = 0; e < 5; ++e)
#pragma unroll
                for (int r = 0; r < 5; ++r) a += lL[e][r] * dL[e][r].d[0] + lf[e] * wc[r] * dR[e][r].d[0];
            B.base[(size_t)(10 * S + q) * N + id] = a;
        }
    }
}

template <int DIR, int RECON, int FVS, int CHAR>
__global__ __launch_bounds__(256) void ka_face(DevParams P, Fields F, AdjArrays A, CharBuf B)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int c = DIR == 0 ? i : DIR == 1 ? j : k;
    face_adjoint<DIR, RECON, FVS, CHAR>(P, F, A, B, i, j, k, c);
    if (!P.periodic[DIR] && c == P.ng) face_adjoint<DIR, RECON, FVS, CHAR>(P, F, A, B, i, j, k, c - 1);
}

// The adjoint of the primitives of cell x along DIR (an interior or, on a
// non-periodic axis, a ghost cell) from the face buffer: its slots of the
// faces whose stencils hold it (face f = x + MID - s at slot s), the reverse
// rule of the split at the cell, and the Roe-state terms of its two faces
// (left cell of face x, right cell of face x - 1).
template <int DIR, int RECON, int FVS, int CHAR>
__device__ __forceinline__ void cell_adjoint(const DevParams &P, const Fields &F, const CharBuf &B, int i, int j,
                                             int k, int x, double g[5])
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const bool per = P.periodic[DIR];
    const size_t N = (size_t)B.n;
    // faces of the box: every wrapped face, or ng - 1 .. ng + n - 1
    auto has_face = [&](int f) { return per || (f >= P.ng - 1 && f <= P.ng + n - 1); };
    auto face_id = [&](int f) { return along<DIR>(P, i, j, k, per ? awrap(f, P.ng, n) : f); };
    double lp[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, lm[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int f = x + MID - s;
        if (!has_face(f)) continue;
        const int fid = face_id(f);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            lp[m] += B.base[(size_t)(5 * s + m) * N + fid];
            lm[m] += B.base[(size_t)(5 * (S + s) + m) * N + fid];
        }
    }
    const int id = along<DIR>(P, i, j, k, x);
    const double w0 = F.W[0][id], w1 = F.W[1][id], w2 = F.W[2][id], w3 = F.W[3][id], w4 = F.W[4][id];
#pragma unroll
    for (int t = 0; t < 5; ++t) g[t] = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        double gv[5];
        split_comp_vjp<FVS, DIR>(P, w0, w1, w2, w3, w4, m, lp[m], lm[m], gv);
#pragma unroll
        for (int t = 0; t < 5; ++t) g[t] += gv[t];
    }
    if constexpr (CHAR) {
        if (has_face(x)) {
            const int fid = face_id(x);
#pragma unroll
            for (int t = 0; t < 5; ++t) g[t] += B.base[(size_t)(10 * S + t) * N + fid];
        }
        if (has_face(x - 1)) {
            const int fid = face_id(x - 1);
#pragma unroll
            for (int t = 0; t < 5; ++t) g[t] += B.base[(size_t)(10 * S + 5 + t) * N + fid];
        }
    }
}

// Per interior cell, after ka_face: the cell's own adjoint and, on a
// non-periodic axis, the adjoints of the ghost cells the boundary fill
// (boundary.cpp:58-207) copies from it: wall and symmetry mirror one ghost
// per cell of the first ng with the normal velocity negated, outflow
// repeats the end cell into all ng ghosts, inflow ghosts are constant.  Into
// Y (store: Y = instead of Y +=).
template <int DIR, int RECON, int FVS, int CHAR>
__global__ __launch_bounds__(256) void ka_gather(DevParams P, Fields F, AdjArrays A, CharBuf B, int store)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int q = DIR == 0 ? i : DIR == 1 ? j : k;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const int ng = P.ng;
    double g[5];
    cell_adjoint<DIR, RECON, FVS, CHAR>(P, F, B, i, j, k, q, g);
    if (!P.periodic[DIR]) {
        for (int side = 0; side < 2; ++side) {
            const int kind = P.bc[2 * DIR + side];
            int g0 = 0, g1 = 0;   // ghosts [g0, g1) that copy this cell
            if (kind == BCK_WALL || kind == BCK_SYMMETRY) {
                const int gq = side == 0 ? 2 * ng - 1 - q : 2 * (ng + n) - 1 - q;
                const bool ok = side == 0 ? gq >= 0 && gq < ng : gq >= ng + n && gq < 2 * ng + n;
                if (ok) { g0 = gq; g1 = gq + 1; }
            } else if (kind == BCK_OUTFLOW) {
                if (side == 0 && q == ng) { g0 = 0; g1 = ng; }
                if (side == 1 && q == ng + n - 1) { g0 = ng + n; g1 = 2 * ng + n; }
            }
            for (int x = g0; x < g1; ++x) {
                double gg[5];
                cell_adjoint<DIR, RECON, FVS, CHAR>(P, F, B, i, j, k, x, gg);
                if (kind != BCK_OUTFLOW) gg[1 + DIR] = -gg[1 + DIR];
#pragma unroll
                for (int t = 0; t < 5; ++t) g[t] += gg[t];
            }
        }
    }
    const int id = (k * P.sy + j) * P.sx + i;
    if (store) {
#pragma unroll
        for (int t = 0; t < 5; ++t) A.Y[t][id] = g[t];
    } else {
#pragma unroll
        for (int t = 0; t < 5; ++t) A.Y[t][id] += g[t];
    }
}

// Viscous adjoint, pointwise part.  Recomputes the gradients and viscous
// flux at the cell, forms lambda_VF = -D(L) (periodic), applies the reverse
// rule of compute_viscous_flux (ViscousFlux.cpp:282-339) into the gradient
// adjoint LG and LW, adds w_obj * d(Phi)/d(.) of the dissipation objective,
// and writes the block sum of (lambda_VF . VF + w_obj Phi) / mu_ref, the
// derivative with respect to 1/Re of lambda . R_vis + w_obj * sum Phi.
// BC: on non-periodic axes the gradients take the closures of
// choose_scheme and lambda_VF = D^T lambda the transposed closure rows.
template <int VIS, int BC>
__global__ __launch_bounds__(256) void ka_vflux(DevParams P, Fields F, AdjArrays A, double w_obj, double *partial)
{
    __shared__ double red[4];
    int bx, by, bz;
    xcd_block3_band(bx, by, bz);
    const int i = P.ng + bx * blockDim.x + threadIdx.x;
    const int j = P.ng + by * blockDim.y + threadIdx.y;
    const int k = P.ng + bz;
    double contrib = 0.0;
    if (i < P.ng + P.nx && j < P.ng + P.ny) {
        const int id = (k * P.sy + j) * P.sx + i;
        const int sxy = P.sx * P.sy;
        double g[12];
        const double *src[4] = {F.W[1], F.W[2], F.W[3], F.W[5]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double *c = src[q] + id;
            if constexpr (BC) {
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    g[3 * q + d] = dgen<VIS>(P, src[q], i, j, k, d, d == 0 ? P.dx : d == 1 ? P.dy : P.dz);
            } else {
                g[3 * q + 0] = diff_central<VIS>(c, 1, P.dx);
                g[3 * q + 1] = diff_central<VIS>(c, P.sx, P.dy);
                g[3 * q + 2] = diff_central<VIS>(c, sxy, P.dz);
            }
        }
        const double rho = F.W[0][id], u = F.W[1][id], v = F.W[2][id], w = F.W[3][id], T = F.W[5][id];
        const double mu = get_mu(P, T);
        const double kq = P.Cp / P.Pr;
        const double ux = g[0], uy = g[1], uz = g[2], vx = g[3], vy = g[4], vz = g[5];
        const double wx = g[6], wy = g[7], wz = g[8];
        const double divU = ux + vy + wz;
        // tau / mu and q / mu
        const double sxx = 2.0 * ux - (2.0 / 3.0) * divU, syy = 2.0 * vy - (2.0 / 3.0) * divU;
        const double szz = 2.0 * wz - (2.0 / 3.0) * divU;
        const double sxy_ = uy + vx, sxz = uz + wx, syz = vz + wy;
        const double txx = mu * sxx, tyy = mu * syy, tzz = mu * szz, txy = mu * sxy_, txz = mu * sxz, tyz = mu * syz;
        double lv[12];
        const double h[3] = {P.dx, P.dy, P.dz};
        const int strd[3] = {1, P.sx, sxy};
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                lv[4 * d + m] = BC && !P.periodic[d] ? dT_closure<VIS>(P, A.X[m + 1], i, j, k, d, h[d])
                                                     : -diff_central<VIS>(A.X[m + 1] + id, strd[d], h[d]);
        const double Ex = lv[3], Ey = lv[7], Ez = lv[11];
        const double btxx = lv[0] + u * Ex;
        const double btxy = lv[1] + lv[4] + v * Ex + u * Ey;
        const double btxz = lv[2] + lv[8] + w * Ex + u * Ez;
        const double btyy = lv[5] + v * Ey;
        const double btyz = lv[6] + lv[9] + w * Ey + v * Ez;
        const double btzz = lv[10] + w * Ez;
        double lg[12];
        const double bdiv = -(2.0 / 3.0) * mu * (btxx + btyy + btzz);
        lg[0] = 2.0 * mu * btxx + bdiv;
        lg[4] = 2.0 * mu * btyy + bdiv;
        lg[8] = 2.0 * mu * btzz + bdiv;
        lg[1] = mu * btxy; lg[3] = mu * btxy;
        lg[2] = mu * btxz; lg[6] = mu * btxz;
        lg[5] = mu * btyz; lg[7] = mu * btyz;
        lg[9] = mu * kq * Ex; lg[10] = mu * kq * Ey; lg[11] = mu * kq * Ez;
        double bmu = btxx * sxx + btyy * syy + btzz * szz + btxy * sxy_ + btxz * sxz + btyz * syz +
                     kq * (Ex * g[9] + Ey * g[10] + Ez * g[11]);
        // dissipation function Phi = tau_ij d_j u_i and its derivatives
        const double phi_mu = 2.0 * (ux * ux + vy * vy + wz * wz) - (2.0 / 3.0) * divU * divU +
                              sxy_ * sxy_ + sxz * sxz + syz * syz;
        if (w_obj != 0.0) {
            const double c4 = w_obj * mu;
            lg[0] += c4 * (4.0 * ux - (4.0 / 3.0) * divU);
            lg[4] += c4 * (4.0 * vy - (4.0 / 3.0) * divU);
            lg[8] += c4 * (4.0 * wz - (4.0 / 3.0) * divU);
            lg[1] += 2.0 * c4 * sxy_; lg[3] += 2.0 * c4 * sxy_;
            lg[2] += 2.0 * c4 * sxz; lg[6] += 2.0 * c4 * sxz;
            lg[5] += 2.0 * c4 * syz; lg[7] += 2.0 * c4 * syz;
            bmu += w_obj * phi_mu;
        }
        const double bT = bmu * dmu_dT(P, T, mu);
        const double bu = txx * Ex + txy * Ey + txz * Ez;
        const double bv = txy * Ex + tyy * Ey + tyz * Ez;
        const double bw = txz * Ex + tyz * Ey + tzz * Ez;
#pragma unroll
        for (int m = 0; m < 12; ++m) A.G[m][id] = lg[m];
        A.Y[0][id] += -bT * T / rho;
        A.Y[1][id] += bu;
        A.Y[2][id] += bv;
        A.Y[3][id] += bw;
        A.Y[4][id] += bT / (rho * P.Rgas);
        // lambda_VF . VF + w_obj Phi, divided by mu_ref (both are linear in it)
        double dot = 0.0;
        const double vf[12] = {txx, txy, txz, u * txx + v * txy + w * txz + mu * kq * g[9],
                               txy, tyy, tyz, u * txy + v * tyy + w * tyz + mu * kq * g[10],
                               txz, tyz, tzz, u * txz + v * tyz + w * tzz + mu * kq * g[11]};
#pragma unroll
        for (int m = 0; m < 12; ++m) dot += lv[m] * vf[m];
        contrib = (dot + w_obj * mu * phi_mu) / P.mu;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) contrib += __shfl_xor(contrib, o, 64);
    const int lane = (threadIdx.y * blockDim.x + threadIdx.x) & 63;
    const int wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    if (lane == 0) red[wid] = contrib;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// RK3 reverse combination applied to lambda_U = C2P'(U)^T LW:
//   acc = (init ? 0 : acc) + a * L_out
//   out = bx * L_out + bl * lambda_U + (add_acc ? acc : 0)
struct Combine {
    double a, bx, bl;
    int init, add_acc;
};

// The end of the gradient adjoint at one cell: LW = Y plus the gradient
// terms lq of u, v, w, T, then C2P'^T and the RK3 combination into C and Z.
// wv = rho, u, v, w, T; yv, xv, cv = Y, X, C at the cell (cv = 0 with init)
__device__ __forceinline__ void grad_finish(const DevParams &P, const AdjArrays &A, const Combine &cb, int id,
                                            const double lq[4], const double wv[5], const double yv[5],
                                            const double xv[5], const double cv[5])
{
    const double rho = wv[0], u = wv[1], v = wv[2], w = wv[3], T = wv[4];
    const double irho = 1.0 / rho;
    const double br = yv[0] - lq[3] * T * irho;
    const double bu = yv[1] + lq[0];
    const double bv = yv[2] + lq[1];
    const double bw = yv[3] + lq[2];
    const double bp = yv[4] + lq[3] * irho / P.Rgas;
    const double g1 = P.gamma - 1.0;
    double lu[5];
    lu[0] = br - (u * bu + v * bv + w * bw) * irho + g1 * 0.5 * (u * u + v * v + w * w) * bp;
    lu[1] = bu * irho - g1 * u * bp;
    lu[2] = bv * irho - g1 * v * bp;
    lu[3] = bw * irho - g1 * w * bp;
    lu[4] = g1 * bp;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double lo = xv[m];
        const double acc = cv[m] + cb.a * lo;
        A.C[m][id] = acc;
        A.Z[m][id] = cb.bx * lo + cb.bl * lu[m] + (cb.add_acc ? acc : 0.0);
    }
}

// Gradient adjoint LW += D^T LG (= -D LG on periodic arrays, the transposed
// closure rows on non-periodic axes with BC) for u, v, w, T, which completes
// LW at the cell, then C2P'^T and the RK3 combination.
template <int VIS, int BC>
__global__ __launch_bounds__(256) void ka_grad(DevParams P, Fields F, AdjArrays A, Combine cb)
{
    int bx, by, bz;
    xcd_block3_band(bx, by, bz);
    const int i = P.ng + bx * blockDim.x + threadIdx.x;
    const int j = P.ng + by * blockDim.y + threadIdx.y;
    const int k = P.ng + bz;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double h[3] = {P.dx, P.dy, P.dz};
    const int strd[3] = {1, P.sx, P.sx * P.sy};
    double lq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            s += BC && !P.periodic[d] ? dT_closure<VIS>(P, A.G[3 * q + d], i, j, k, d, h[d])
                                      : -diff_central<VIS>(A.G[3 * q + d] + id, strd[d], h[d]);
        lq[q] = s;
    }
    double wv[5], yv[5], xv[5], cv[5];
    const double *wsrc[5] = {F.W[0], F.W[1], F.W[2], F.W[3], F.W[5]};
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        wv[m] = wsrc[m][id];
        yv[m] = A.Y[m][id];
        xv[m] = A.X[m][id];
        cv[m] = cb.init ? 0.0 : A.C[m][id];
    }
    grad_finish(P, A, cb, id, lq, wv, yv, xv, cv);
}

// Gradient adjoint as a z march on a periodic box that 64 x 4 blocks tile
// exactly: each block marches a z segment of its 64 x 4 column; the four
// x-differentiated arrays of each plane are staged in LDS with an R-cell x
// halo, the four y-differentiated ones with an R-row y halo (both from the
// ghost layers, double-buffered, one barrier per plane), and the four
// z-differentiated ones slide through a register window.  Everything the
// next plane needs is loaded before this plane's arithmetic, so the loads
// overlap it: about 41 loads per cell instead of 108 (ka_grad is bound by
// its load instructions).  The same values in the same formulas as
// ka_grad<VIS, 0>: equal bit for bit.
template <int VIS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void ka_grad_z(
    DevParams P, Fields F, AdjArrays A, Combine cb, int seg_len, int gx, int gy)
{
    constexpr int R = VisReach<VIS>::R;
    constexpr int TC = 64 + 2 * R, TYR = 4 + 2 * R;
    __shared__ double xt[2][4][4][TC];
    __shared__ double yt[2][4][TYR][64];
    const int t = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int ng = P.ng;
    const int i0 = ng + bx * 64, j0 = ng + by * 4;
    const int i = i0 + tx, j = j0 + ty;
    const int k0 = ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, ng + P.nz);
    const int sx = P.sx, sxy = P.sx * P.sy;
    const int col = j * sx + i;
    // the x-halo cell of this thread (tid < 32 R): array tid / 8R, tile row
    // (tid / 2R) % 4, halo column tid % 2R
    const bool xh_on = tid < 32 * R;
    const int xh_q = xh_on ? tid / (8 * R) : 0;
    const int xh_r = (tid / (2 * R)) % 4;
    const int xh_c = tid % (2 * R);
    const int xh_tc = xh_c < R ? xh_c : 64 + xh_c;
    const int xh_off = (j0 + xh_r) * sx + i0 - R + xh_tc;
    // the y-halo rows of this thread: h = ty, ty + 4, ... < 2R, tile row h
    // (below the tile) or h + 4 (above it)
    constexpr int NYH = (2 * R + 3) / 4;
    int yh_row[NYH], yh_off[NYH];
#pragma unroll
    for (int n = 0; n < NYH; ++n) {
        const int h = ty + 4 * n;
        const int tr = h < R ? h : h + 4;
        yh_row[n] = h < 2 * R ? tr : -1;
        yh_off[n] = (j0 - R + tr) * sx + i;
    }
    const double *wsrc[5] = {F.W[0], F.W[1], F.W[2], F.W[3], F.W[5]};
    double pxo[4], pyo[4], pxh = 0.0, pyh[NYH][4], pz[4], pw[5], py[5], px[5], pc[5];
    const auto prefetch = [&](int kp) {
        const int pl = kp * sxy;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pxo[q] = A.G[3 * q][pl + col];
            pyo[q] = A.G[3 * q + 1][pl + col];
        }
        if (xh_on) pxh = A.G[3 * xh_q][pl + xh_off];
#pragma unroll
        for (int n = 0; n < NYH; ++n)
            if (yh_row[n] >= 0)
#pragma unroll
                for (int q = 0; q < 4; ++q) pyh[n][q] = A.G[3 * q + 1][pl + yh_off[n]];
        const int zn = (kp + R) * sxy + col;
#pragma unroll
        for (int q = 0; q < 4; ++q) pz[q] = A.G[3 * q + 2][zn];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            pw[m] = wsrc[m][pl + col];
            py[m] = A.Y[m][pl + col];
            px[m] = A.X[m][pl + col];
            pc[m] = cb.init ? 0.0 : A.C[m][pl + col];
        }
    };
    double win[4][2 * R + 1];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o) win[q][o] = A.G[3 * q + 2][(k0 - R - 1 + o) * sxy + col];
    prefetch(k0);
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[q][o] = win[q][o + 1];
            win[q][2 * R] = pz[q];
        }
        double (*XT)[4][TC] = xt[k & 1];
        double (*YT)[TYR][64] = yt[k & 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            XT[q][ty][R + tx] = pxo[q];
            YT[q][R + ty][tx] = pyo[q];
        }
        if (xh_on) XT[xh_q][xh_r][xh_tc] = pxh;
#pragma unroll
        for (int n = 0; n < NYH; ++n)
            if (yh_row[n] >= 0)
#pragma unroll
                for (int q = 0; q < 4; ++q) YT[q][yh_row[n]][tx] = pyh[n][q];
        double wv[5], yv[5], xv[5], cv[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) { wv[m] = pw[m]; yv[m] = py[m]; xv[m] = px[m]; cv[m] = pc[m]; }
        __syncthreads();
        if (k + 1 < k1) prefetch(k + 1);
        double lq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double vx[2 * R + 1], vy[2 * R + 1];
#pragma unroll
            for (int o = -R; o <= R; ++o) {
                vx[o + R] = XT[q][ty][R + tx + o];
                vy[o + R] = YT[q][R + ty + o][tx];
            }
            double s = 0.0;
            s += -diff_window<VIS>(vx + R, P.dx);
            s += -diff_window<VIS>(vy + R, P.dy);
            s += -diff_window<VIS>(&win[q][R], P.dz);
            lq[q] = s;
        }
        grad_finish(P, A, cb, id, lq, wv, yv, xv, cv);
    }
}

// one block: out[0] = sum of partial[0..n) in a fixed order
__global__ __launch_bounds__(256) void k_sum_partials(const double *partial, int n, double *out)
{
    __shared__ double red[4];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) s += partial[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (red[0] + red[1]) + (red[2] + red[3]);
}

// volume-mean dissipation: block sums of Phi = tau_ij d_j u_i (BC: the
// closures on non-periodic axes, as the viscous flux takes them)
template <int VIS, int BC>
__global__ __launch_bounds__(256) void k_dissip(DevParams P, Fields F, double *partial)
{
    __shared__ double red[4];
    int bx, by, bz;
    xcd_block3_band(bx, by, bz);
    const int i = P.ng + bx * blockDim.x + threadIdx.x;
    const int j = P.ng + by * blockDim.y + threadIdx.y;
    const int k = P.ng + bz;
    double phi = 0.0;
    if (i < P.ng + P.nx && j < P.ng + P.ny) {
        const int id = (k * P.sy + j) * P.sx + i;
        const int sxy = P.sx * P.sy;
        double g[9];
        const double *src[3] = {F.W[1], F.W[2], F.W[3]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double *c = src[q] + id;
            if constexpr (BC) {
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    g[3 * q + d] = dgen<VIS>(P, src[q], i, j, k, d, d == 0 ? P.dx : d == 1 ? P.dy : P.dz);
            } else {
                g[3 * q + 0] = diff_central<VIS>(c, 1, P.dx);
                g[3 * q + 1] = diff_central<VIS>(c, P.sx, P.dy);
                g[3 * q + 2] = diff_central<VIS>(c, sxy, P.dz);
            }
        }
        const double mu = get_mu(P, F.W[5][id]);
        const double divU = g[0] + g[4] + g[8];
        phi = mu * (2.0 * (g[0] * g[0] + g[4] * g[4] + g[8] * g[8]) - (2.0 / 3.0) * divU * divU +
                    (g[1] + g[3]) * (g[1] + g[3]) + (g[2] + g[6]) * (g[2] + g[6]) + (g[5] + g[7]) * (g[5] + g[7]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) phi += __shfl_xor(phi, o, 64);
    const int lane = (threadIdx.y * blockDim.x + threadIdx.x) & 63;
    const int wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    if (lane == 0) red[wid] = phi;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// -------------------------------------------------------------- launches

static inline dim3 agrid(const DevParams &P) { return dim3((P.nx + 63) / 64, (P.ny + 3) / 4, P.nz); }
static const dim3 kABlock(64, 4, 1);

// column segments of the adjoint sweeps: about 8192 waves in flight over
// the device, segments no shorter than 16 cells
int ad_seg(const DevParams &P, int dir)
{
    const int na = dir == 0 ? P.ny : P.nx;
    const int nb = dir == 2 ? P.ny : P.nz;
    const int nd = dir == 0 ? P.nx : dir == 1 ? P.ny : P.nz;
    const long long cols = (long long)((na + kAdjCols - 1) / kAdjCols) * nb;
    long long nseg = (8192 + cols - 1) / cols;
    const long long maxseg = nd / 16 > 0 ? nd / 16 : 1;
    if (nseg > maxseg) nseg = maxseg;
    if (nseg < 1) nseg = 1;
    return (int)((nd + nseg - 1) / nseg);
}

int ad_blocks(const DevParams &P) { return ((P.nx + 63) / 64) * ((P.ny + 3) / 4) * P.nz; }

bool ad_supported(const DevParams &P)
{
    return P.recon == R_WENO5 || P.recon == R_LINEAR || P.recon == R_MDCD || P.recon == R_MDCD_HYBRID ||
           P.recon == R_UD7 || P.recon == R_UD7L;
}

// some axis is not periodic: the BC forms of the viscous kernels
static bool any_bc(const DevParams &P) { return !(P.periodic[0] && P.periodic[1] && P.periodic[2]); }

// the z-marching forms of the viscous adjoint kernels: a periodic box that
// 64 x 4 blocks tile exactly (every thread reaches the barriers); z
// segments of kAdjSegZ planes
constexpr int kAdjSegZ = 32;
static bool adj_zmarch(const DevParams &P)
{
    return !any_bc(P) && P.adj_zmarch && P.nx % 64 == 0 && P.ny % 4 == 0;
}

template <int RECON, int FVS, int CHAR>
static void tlm_faces_c(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir)
{
    if (dir == 0) hipLaunchKernelGGL((kt_faces<0, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A);
    else if (dir == 1) hipLaunchKernelGGL((kt_faces<1, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A);
    else hipLaunchKernelGGL((kt_faces<2, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A);
}

template <int RECON, int FVS>
static void tlm_faces_r(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int ch)
{
    if (ch) tlm_faces_c<RECON, FVS, 1>(s, P, F, A, dir);
    else tlm_faces_c<RECON, FVS, 0>(s, P, F, A, dir);
}

template <int FVS>
static void tlm_faces_f(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int ch)
{
    switch (P.recon) {
    case R_WENO5: tlm_faces_r<R_WENO5, FVS>(s, P, F, A, dir, ch); break;
    case R_LINEAR: tlm_faces_r<R_LINEAR, FVS>(s, P, F, A, dir, ch); break;
    case R_MDCD: tlm_faces_r<R_MDCD, FVS>(s, P, F, A, dir, ch); break;
    case R_MDCD_HYBRID: tlm_faces_r<R_MDCD_HYBRID, FVS>(s, P, F, A, dir, ch); break;
    case R_UD7: tlm_faces_r<R_UD7, FVS>(s, P, F, A, dir, ch); break;
    default: tlm_faces_r<R_UD7L, FVS>(s, P, F, A, dir, ch); break;
    }
}

// the tangent's ghosts on the non-periodic sides, in apply_boundary's order
static void tlm_fill_bc(hipStream_t s, const DevParams &P, const AdjArrays &A)
{
    for (int face = 0; face < 6; ++face) {
        const int axis = face >> 1;
        if (P.periodic[axis]) continue;
        const int n = P.ng * (axis == 0 ? P.sy : P.sx) * (axis == 2 ? P.sy : P.sz);
        hipLaunchKernelGGL(kt_fill_bc, dim3((n + 255) / 256), dim3(256), 0, s, P, A, face, P.bc[face]);
    }
}

template <int VIS>
static void tlm_visc(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A)
{
    hipLaunchKernelGGL(kt_tc, agrid(P), kABlock, 0, s, P, F, A);
    hipLaunchKernelGGL(kt_vflux<VIS>, agrid(P), kABlock, 0, s, P, F, A);
    hipLaunchKernelGGL(kt_vdiv<VIS>, agrid(P), kABlock, 0, s, P, A);
}

// Z <- R'(U) X for the state in F (TLM); uses Y, T, V as scratch;
// char_recon: the characteristic inviscid faces
hipError_t launch_tlm_rhs(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int char_recon)
{
    if (!ad_supported(P)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kt_c2p, agrid(P), kABlock, 0, s, P, F, A);
    tlm_fill_bc(s, P, A);
    for (int d = 0; d < 3; ++d) {
        if (P.fvs == FVS_SW) tlm_faces_f<FVS_SW>(s, P, F, A, d, char_recon);
        else tlm_faces_f<FVS_LF>(s, P, F, A, d, char_recon);
        if (d == 0) hipLaunchKernelGGL(kt_diff<0>, agrid(P), kABlock, 0, s, P, A, 1);
        else if (d == 1) hipLaunchKernelGGL(kt_diff<1>, agrid(P), kABlock, 0, s, P, A, 0);
        else hipLaunchKernelGGL(kt_diff<2>, agrid(P), kABlock, 0, s, P, A, 0);
    }
    switch (P.vis) {
    case VIS_C8: tlm_visc<VIS_C8>(s, P, F, A); break;
    case VIS_C6: tlm_visc<VIS_C6>(s, P, F, A); break;
    default: tlm_visc<VIS_C4>(s, P, F, A); break;
    }
    return hipGetLastError();
}

// Lines per block of the whole-line adjoint sweep along `dir` (0: the sweep
// falls back to the per-component marching kernel): as many neighbouring
// x-columns as fit 1024 threads and half the device's LDS (two workgroups
// per CU) for the scheme's stencil (two buffers of 1 + S planes per line).
// A whole line must fit the device's LDS.  At 128^3 on the MI355X (160 KiB)
// that is 4 lines per block; 8 lines at one workgroup per CU ran the y sweep
// in 0.24 ms against 0.21 ms.
static long long adj_line_bytes(const DevParams &P, int n)
{
    const int S = P.recon == R_UD7 || P.recon == R_UD7L ? 8 : P.recon == R_LINEAR ? 2 : 6;
    return 2LL * (1 + S) * n * (long long)sizeof(double);
}

int adj_line_tx(const DevParams &P, int dir)
{
    const int n = dir == 0 ? P.nx : dir == 1 ? P.ny : P.nz;
    const long long per_line = adj_line_bytes(P, n);
    if (!P.adj_lines || n > 1024 || per_line > P.lds_max) return 0;
    if (dir == 0) return 1;
    int tx = 16;
    while (tx > 1 && ((long long)tx * n > 1024 || tx * per_line > P.lds_max / 2)) tx >>= 1;
    return tx;
}

// the x adjoint sweep runs first and stores Y when it holds whole rows
bool adj_xrow(const DevParams &P) { return adj_line_tx(P, 0) > 0; }

template <int RECON, int FVS>
static void adj_sweep_r(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int seg)
{
    const int tx = adj_line_tx(P, dir);
    if (tx > 0) {
        const int n = dir == 0 ? P.nx : dir == 1 ? P.ny : P.nz;
        const size_t lds = (size_t)adj_line_bytes(P, n) * tx;
        const dim3 blk(tx, n);
        const int store = dir == 0 ? 1 : 0;
        if (lds > 65536) {
            // beyond the default dynamic LDS limit of a launch
            (void)hipFuncSetAttribute((const void *)ka_sweep_line<0, RECON, FVS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            (void)hipFuncSetAttribute((const void *)ka_sweep_line<1, RECON, FVS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            (void)hipFuncSetAttribute((const void *)ka_sweep_line<2, RECON, FVS>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        }
        if (dir == 0)
            hipLaunchKernelGGL((ka_sweep_line<0, RECON, FVS>), dim3(P.ny, P.nz), blk, lds, s, P, F, A, store);
        else if (dir == 1)
            hipLaunchKernelGGL((ka_sweep_line<1, RECON, FVS>), dim3((P.nx + tx - 1) / tx, P.nz), blk, lds, s, P, F,
                               A, store);
        else
            hipLaunchKernelGGL((ka_sweep_line<2, RECON, FVS>), dim3((P.nx + tx - 1) / tx, P.ny), blk, lds, s, P, F,
                               A, store);
        return;
    }
    const int na = dir == 0 ? P.ny : P.nx;
    const int nb = dir == 2 ? P.ny : P.nz;
    const int nd = dir == 0 ? P.nx : dir == 1 ? P.ny : P.nz;
    dim3 g((na + kAdjCols - 1) / kAdjCols, nb, (nd + seg - 1) / seg);
    if (dir == 0) hipLaunchKernelGGL((ka_sweepc<0, RECON, FVS>), g, dim3(64), 0, s, P, F, A, seg);
    else if (dir == 1) hipLaunchKernelGGL((ka_sweepc<1, RECON, FVS>), g, dim3(64), 0, s, P, F, A, seg);
    else hipLaunchKernelGGL((ka_sweepc<2, RECON, FVS>), g, dim3(64), 0, s, P, F, A, seg);
}

template <int FVS>
static void adj_sweep_f(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int seg)
{
    switch (P.recon) {
    case R_WENO5: adj_sweep_r<R_WENO5, FVS>(s, P, F, A, dir, seg); break;
    case R_LINEAR: adj_sweep_r<R_LINEAR, FVS>(s, P, F, A, dir, seg); break;
    case R_MDCD: adj_sweep_r<R_MDCD, FVS>(s, P, F, A, dir, seg); break;
    case R_MDCD_HYBRID: adj_sweep_r<R_MDCD_HYBRID, FVS>(s, P, F, A, dir, seg); break;
    case R_UD7: adj_sweep_r<R_UD7, FVS>(s, P, F, A, dir, seg); break;
    default: adj_sweep_r<R_UD7L, FVS>(s, P, F, A, dir, seg); break;
    }
}

template <int VIS>
static void adj_vflux(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double w_obj,
                      double *partial)
{
    ArraySet xs{{A.X[1], A.X[2], A.X[3], A.X[4]}, 4};
    fill_set(s, P, xs);
    if (any_bc(P)) hipLaunchKernelGGL((ka_vflux<VIS, 1>), agrid(P), kABlock, 0, s, P, F, A, w_obj, partial);
    else hipLaunchKernelGGL((ka_vflux<VIS, 0>), agrid(P), kABlock, 0, s, P, F, A, w_obj, partial);
}

template <int VIS>
static void adj_grad(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, const Combine &cb)
{
    ArraySet gs{};
    for (int m = 0; m < 12; ++m) gs.p[m] = A.G[m];
    gs.n = 12;
    fill_set(s, P, gs);
    if (any_bc(P)) {
        hipLaunchKernelGGL((ka_grad<VIS, 1>), agrid(P), kABlock, 0, s, P, F, A, cb);
    } else if (adj_zmarch(P)) {
        const int seg = P.nz < kAdjSegZ ? P.nz : kAdjSegZ;
        const int gx = P.nx / 64, gy = P.ny / 4, n = gx * gy * ((P.nz + seg - 1) / seg);
        hipLaunchKernelGGL((ka_grad_z<VIS>), dim3(n), kABlock, 0, s, P, F, A, cb, seg, gx, gy);
    } else {
        hipLaunchKernelGGL((ka_grad<VIS, 0>), agrid(P), kABlock, 0, s, P, F, A, cb);
    }
}

template <int RECON, int FVS, int CHAR>
static void adj_face_c(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, const CharBuf &B, int dir)
{
    const int store = dir == 0 ? 1 : 0;
    if (dir == 0) {
        hipLaunchKernelGGL((ka_face<0, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A, B);
        hipLaunchKernelGGL((ka_gather<0, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A, B, store);
    } else if (dir == 1) {
        hipLaunchKernelGGL((ka_face<1, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A, B);
        hipLaunchKernelGGL((ka_gather<1, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A, B, store);
    } else {
        hipLaunchKernelGGL((ka_face<2, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A, B);
        hipLaunchKernelGGL((ka_gather<2, RECON, FVS, CHAR>), agrid(P), kABlock, 0, s, P, F, A, B, store);
    }
}

template <int FVS, int CHAR>
static void adj_face_f(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, const CharBuf &B, int dir)
{
    switch (P.recon) {
    case R_WENO5: adj_face_c<R_WENO5, FVS, CHAR>(s, P, F, A, B, dir); break;
    case R_LINEAR: adj_face_c<R_LINEAR, FVS, CHAR>(s, P, F, A, B, dir); break;
    case R_MDCD: adj_face_c<R_MDCD, FVS, CHAR>(s, P, F, A, B, dir); break;
    case R_MDCD_HYBRID: adj_face_c<R_MDCD_HYBRID, FVS, CHAR>(s, P, F, A, B, dir); break;
    case R_UD7: adj_face_c<R_UD7, FVS, CHAR>(s, P, F, A, B, dir); break;
    default: adj_face_c<R_UD7L, FVS, CHAR>(s, P, F, A, B, dir); break;
    }
}

// double arrays of the per-face buffer of the face-wise inviscid adjoint:
// the F+ and F- adjoints of every stencil slot and the Roe-state terms of
// two cells
int char_buf_arrays(const DevParams &P)
{
    const int S = P.recon == R_UD7 || P.recon == R_UD7L ? 8 : P.recon == R_LINEAR ? 2 : 6;
    return 10 * S + 10;
}

// the inviscid adjoint along `dir` runs face-wise (per-face buffer): always
// for the characteristic form, and on non-periodic axes
bool adj_by_faces(const DevParams &P, int dir, int char_recon) { return char_recon || !P.periodic[dir]; }

// the x adjoint sweep stores Y (else Y = 0 is needed before it)
bool adj_x_stores(const DevParams &P, int char_recon) { return adj_by_faces(P, 0, char_recon) || adj_xrow(P); }

// Y += W-adjoint of X . R_inv(U) along `dir` (Y = 0 before the first call
// unless adj_x_stores); the face-wise form goes through cb
hipError_t launch_adj_sweep(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int seg,
                            const CharBuf *cb, int char_recon)
{
    if (!ad_supported(P)) return hipErrorInvalidValue;
    if (adj_by_faces(P, dir, char_recon)) {
        if (!cb || !cb->base || cb->n != (long long)P.sx * P.sy * P.sz) return hipErrorInvalidValue;
        if (char_recon) {
            if (P.fvs == FVS_SW) adj_face_f<FVS_SW, 1>(s, P, F, A, *cb, dir);
            else adj_face_f<FVS_LF, 1>(s, P, F, A, *cb, dir);
        } else {
            if (P.fvs == FVS_SW) adj_face_f<FVS_SW, 0>(s, P, F, A, *cb, dir);
            else adj_face_f<FVS_LF, 0>(s, P, F, A, *cb, dir);
        }
        return hipGetLastError();
    }
    if (P.fvs == FVS_SW) adj_sweep_f<FVS_SW>(s, P, F, A, dir, seg);
    else adj_sweep_f<FVS_LF>(s, P, F, A, dir, seg);
    return hipGetLastError();
}

// Viscous flux part of the adjoint RHS: LW += the reverse rule of the
// viscous flux at X, LG <- gradient adjoints; block partials of the 1/Re
// derivative of X . R_vis + w_obj * sum Phi into `partial`.
hipError_t launch_adj_vflux(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double w_obj,
                            double *partial)
{
    switch (P.vis) {
    case VIS_C8: adj_vflux<VIS_C8>(s, P, F, A, w_obj, partial); break;
    case VIS_C6: adj_vflux<VIS_C6>(s, P, F, A, w_obj, partial); break;
    default: adj_vflux<VIS_C4>(s, P, F, A, w_obj, partial); break;
    }
    return hipGetLastError();
}

// Gradient part, which completes LW, then Z = bx X + bl C2P'^T LW (+ the
// accumulator C, updated with a X).
hipError_t launch_adj_grad(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double a,
                           double bx, double bl, int init, int add_acc)
{
    const Combine cb{a, bx, bl, init, add_acc};
    switch (P.vis) {
    case VIS_C8: adj_grad<VIS_C8>(s, P, F, A, cb); break;
    case VIS_C6: adj_grad<VIS_C6>(s, P, F, A, cb); break;
    default: adj_grad<VIS_C4>(s, P, F, A, cb); break;
    }
    return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t s, const double *partial, int n, double *out)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, n, out);
    return hipGetLastError();
}

template <int VIS>
static void dissip_v(hipStream_t s, const DevParams &P, const Fields &F, double *partial)
{
    if (any_bc(P)) hipLaunchKernelGGL((k_dissip<VIS, 1>), agrid(P), kABlock, 0, s, P, F, partial);
    else hipLaunchKernelGGL((k_dissip<VIS, 0>), agrid(P), kABlock, 0, s, P, F, partial);
}

hipError_t launch_dissip(hipStream_t s, const DevParams &P, const Fields &F, double *partial)
{
    switch (P.vis) {
    case VIS_C8: dissip_v<VIS_C8>(s, P, F, partial); break;
    case VIS_C6: dissip_v<VIS_C6>(s, P, F, partial); break;
    default: dissip_v<VIS_C4>(s, P, F, partial); break;
    }
    return hipGetLastError();
}

}  // namespace ns3d
