// ns3d_adjoint.hip -- tangent-linear (TLM) and discrete adjoint of the FP64
// RHS on a periodic box, and the dissipation objective.
//
// Notation: W = (rho, u, v, w, p) with T = p / (rho R); U = conserved.
// R(U) = R_inv + R_vis as in range_kutta.cpp:11-99.  With periodic ghosts
// every face of the reference's face range is a periodic face f (right face
// of cell f, left face of cell f+1), so the adjoint of the flux difference is
//     lambda_F(f) = (L(f+1) - L(f)) / h
// and the central differences satisfy D^T = -D.  All adjoint arrays hold
// interior cells; periodic images are reached by index wrap.
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

__device__ __forceinline__ double dmu_dT(const DevParams &P, double T, double mu)
{
    if (P.mu_law == 1) return 0.0;
    if (P.mu_law == 2) return mu * P.mu_power / T;
    return mu * (1.5 / T - 1.0 / (T + P.sutherland_s));
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

// ------------------------------------------------------------------ TLM

// tangent of the face flux of the right face of every interior cell along
// DIR, from W (ghost-filled) and dW (interior, wrapped): dF[m] arrays
template <int DIR, int RECON, int FVS>
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
    using D1 = Dual<1>;
    D1 fp[S][5], fm[S][5];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int q = awrap(c - MID + s, P.ng, n);
        const int id = DIR == 0 ? (k * P.sy + j) * P.sx + q : DIR == 1 ? (k * P.sy + q) * P.sx + i
                                                                        : (q * P.sy + j) * P.sx + i;
        D1 x[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) { x[m].v = F.W[m][id]; x[m].d[0] = A.Y[m][id]; }
        split_t<FVS, DIR>(P, x[0], x[1], x[2], x[3], x[4], fp[s], fm[s]);
    }
    const int id = (k * P.sy + j) * P.sx + i;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        D1 sp[S], sm[S];
#pragma unroll
        for (int s = 0; s < S; ++s) { sp[s] = fp[s][m]; sm[S - 1 - s] = fm[s][m]; }
        D1 r = recon_t<RECON>(P, sp) + recon_t<RECON>(P, sm);
        A.T[m][id] = r.d[0];
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
    int il = id;
    if (DIR == 0) il = (k * P.sy + j) * P.sx + awrap(i - 1, P.ng, P.nx);
    else if (DIR == 1) il = (k * P.sy + awrap(j - 1, P.ng, P.ny)) * P.sx + i;
    else il = (awrap(k - 1, P.ng, P.nz) * P.sy + j) * P.sx + i;
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
            g[3 * q + d].v = dwrap<VIS>(P, vals[q], i, j, k, d, h[d]);
            g[3 * q + d].d[0] = dwrap<VIS>(P, tans[q], i, j, k, d, h[d]);
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
        for (int d = 0; d < 3; ++d) s += dwrap<VIS>(P, A.V[4 * d + m], i, j, k, d, h[d]);
        A.Z[m + 1][id] += s;
    }
}

// --------------------------------------------------------------- adjoint

// Inviscid adjoint along DIR: one thread per column segment of interior
// cells [c0, c1).  It marches over every periodic face whose stencil holds
// one of its cells, keeping the split fluxes of the S stencil cells and
// their adjoint accumulators in registers; the reconstruction's reverse rule
// spreads lambda_F over the stencil, and when a cell leaves the window its
// accumulated split-flux adjoint goes through the split Jacobian (5-direction
// dual numbers) into LW (rho, u, v, w, p).  Cells are owned by one thread:
// no atomics.
template <int DIR, int RECON, int FVS>
__global__ __launch_bounds__(64) void ka_sweep(DevParams P, Fields F, AdjArrays A, int seg_len)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    const int na = DIR == 0 ? P.ny : P.nx;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const int b = blockIdx.y;
    const int n = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const int c0 = P.ng + blockIdx.z * seg_len;
    const int c1 = min(c0 + seg_len, P.ng + n);
    if (c0 >= c1) return;
    int base, st;
    if (DIR == 0) { base = ((P.ng + b) * P.sy + (P.ng + a)) * P.sx; st = 1; }
    else if (DIR == 1) { base = ((P.ng + b) * P.sy) * P.sx + (P.ng + a); st = P.sx; }
    else { base = (P.ng + b) * P.sx + (P.ng + a); st = P.sx * P.sy; }
    const double inv_d = DIR == 0 ? P.idx : DIR == 1 ? P.idy : P.idz;
    auto cid = [&](int q) { return base + awrap(q, P.ng, n) * st; };

    double fp[S][5], fm[S][5], fpb[S][5], fmb[S][5];
    const int fA = c0 - S + 1 + MID, fB = c1 - 1 + MID;
#pragma unroll
    for (int s = 1; s < S; ++s) {
        const int id = cid(fA - 1 - MID + s);
        split_t<FVS, DIR>(P, F.W[0][id], F.W[1][id], F.W[2][id], F.W[3][id], F.W[4][id], fp[s], fm[s]);
#pragma unroll
        for (int m = 0; m < 5; ++m) { fpb[s][m] = 0.0; fmb[s][m] = 0.0; }
    }
    double ll[5];
    {
        const int id = cid(fA);
#pragma unroll
        for (int m = 0; m < 5; ++m) ll[m] = A.X[m][id];
    }
    for (int f = fA; f <= fB; ++f) {
#pragma unroll
        for (int s = 0; s < S - 1; ++s)
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                fp[s][m] = fp[s + 1][m]; fm[s][m] = fm[s + 1][m];
                fpb[s][m] = fpb[s + 1][m]; fmb[s][m] = fmb[s + 1][m];
            }
        {
            const int id = cid(f - MID + S - 1);
            split_t<FVS, DIR>(P, F.W[0][id], F.W[1][id], F.W[2][id], F.W[3][id], F.W[4][id], fp[S - 1], fm[S - 1]);
#pragma unroll
            for (int m = 0; m < 5; ++m) { fpb[S - 1][m] = 0.0; fmb[S - 1][m] = 0.0; }
        }
        double gF[5];
        {
            const int id = cid(f + 1);
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const double lr = A.X[m][id];
                gF[m] = (lr - ll[m]) * inv_d;
                ll[m] = lr;
            }
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sp[S], sm[S], gb[S];
#pragma unroll
            for (int s = 0; s < S; ++s) { sp[s] = fp[s][m]; sm[S - 1 - s] = fm[s][m]; }
            recon_vjp<RECON>(P, sp, gF[m], gb);
#pragma unroll
            for (int s = 0; s < S; ++s) fpb[s][m] += gb[s];
            recon_vjp<RECON>(P, sm, gF[m], gb);
#pragma unroll
            for (int s = 0; s < S; ++s) fmb[s][m] += gb[S - 1 - s];
        }
        const int q = f - MID;
        if (q >= c0 && q < c1) {
            const int id = base + q * st;
            Dual<5> x[5], Fp[5], Fm[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                x[m].v = F.W[m][id];
#pragma unroll
                for (int t = 0; t < 5; ++t) x[m].d[t] = t == m ? 1.0 : 0.0;
            }
            split_t<FVS, DIR>(P, x[0], x[1], x[2], x[3], x[4], Fp, Fm);
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < 5; ++m) s += fpb[0][m] * Fp[m].d[t] + fmb[0][m] * Fm[m].d[t];
                A.Y[t][id] += s;
            }
        }
    }
}

// Viscous adjoint, pointwise part.  Recomputes the gradients and viscous
// flux at the cell, forms lambda_VF = -D(L) (periodic), applies the reverse
// rule of compute_viscous_flux (ViscousFlux.cpp:282-339) into the gradient
// adjoint LG and LW, adds w_obj * d(Phi)/d(.) of the dissipation objective,
// and writes the block sum of (lambda_VF . VF + w_obj Phi) / mu_ref, the
// derivative with respect to 1/Re of lambda . R_vis + w_obj * sum Phi.
template <int VIS>
__global__ __launch_bounds__(256) void ka_vflux(DevParams P, Fields F, AdjArrays A, double w_obj, double *partial)
{
    __shared__ double red[4];
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    double contrib = 0.0;
    if (i < P.ng + P.nx && j < P.ng + P.ny) {
        const int id = (k * P.sy + j) * P.sx + i;
        const int sxy = P.sx * P.sy;
        double g[12];
        const double *src[4] = {F.W[1], F.W[2], F.W[3], F.W[5]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double *c = src[q] + id;
            g[3 * q + 0] = diff_central<VIS>(c, 1, P.dx);
            g[3 * q + 1] = diff_central<VIS>(c, P.sx, P.dy);
            g[3 * q + 2] = diff_central<VIS>(c, sxy, P.dz);
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
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < 4; ++m) lv[4 * d + m] = -dwrap<VIS>(P, A.X[m + 1], i, j, k, d, h[d]);
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

// gradient adjoint: LW += D^T LG (= -D LG, periodic) for u, v, w, T
template <int VIS>
__global__ __launch_bounds__(256) void ka_grad(DevParams P, Fields F, AdjArrays A)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double h[3] = {P.dx, P.dy, P.dz};
    double lq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) s -= dwrap<VIS>(P, A.G[3 * q + d], i, j, k, d, h[d]);
        lq[q] = s;
    }
    const double rho = F.W[0][id], T = F.W[5][id];
    A.Y[0][id] += -lq[3] * T / rho;
    A.Y[1][id] += lq[0];
    A.Y[2][id] += lq[1];
    A.Y[3][id] += lq[2];
    A.Y[4][id] += lq[3] / (rho * P.Rgas);
}

// lambda_U = C2P'(U)^T LW, then the RK3 reverse combination:
//   acc = (init ? 0 : acc) + a * L_out
//   out = bx * L_out + bl * lambda_U + (add_acc ? acc : 0)
__global__ void ka_combine(DevParams P, Fields F, AdjArrays A, double a, double bx, double bl, int init, int add_acc)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double rho = F.W[0][id], u = F.W[1][id], v = F.W[2][id], w = F.W[3][id];
    const double g1 = P.gamma - 1.0;
    const double br = A.Y[0][id], bu = A.Y[1][id], bv = A.Y[2][id], bw = A.Y[3][id], bp = A.Y[4][id];
    double lu[5];
    lu[0] = br - (u * bu + v * bv + w * bw) / rho + g1 * 0.5 * (u * u + v * v + w * w) * bp;
    lu[1] = bu / rho - g1 * u * bp;
    lu[2] = bv / rho - g1 * v * bp;
    lu[3] = bw / rho - g1 * w * bp;
    lu[4] = g1 * bp;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double lo = A.X[m][id];
        const double acc = (init ? 0.0 : A.C[m][id]) + a * lo;
        A.C[m][id] = acc;
        A.Z[m][id] = bx * lo + bl * lu[m] + (add_acc ? acc : 0.0);
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

// volume-mean dissipation: block sums of Phi = tau_ij d_j u_i
template <int VIS>
__global__ __launch_bounds__(256) void k_dissip(DevParams P, Fields F, double *partial)
{
    __shared__ double red[4];
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    double phi = 0.0;
    if (i < P.ng + P.nx && j < P.ng + P.ny) {
        const int id = (k * P.sy + j) * P.sx + i;
        const int sxy = P.sx * P.sy;
        double g[9];
        const double *src[3] = {F.W[1], F.W[2], F.W[3]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double *c = src[q] + id;
            g[3 * q + 0] = diff_central<VIS>(c, 1, P.dx);
            g[3 * q + 1] = diff_central<VIS>(c, P.sx, P.dy);
            g[3 * q + 2] = diff_central<VIS>(c, sxy, P.dz);
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

int ad_blocks(const DevParams &P) { return ((P.nx + 63) / 64) * ((P.ny + 3) / 4) * P.nz; }

bool ad_supported(const DevParams &P)
{
    return P.recon == R_WENO5 || P.recon == R_LINEAR || P.recon == R_MDCD || P.recon == R_UD7 ||
           P.recon == R_UD7L;
}

template <int RECON, int FVS>
static void tlm_faces_r(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir)
{
    if (dir == 0) hipLaunchKernelGGL((kt_faces<0, RECON, FVS>), agrid(P), kABlock, 0, s, P, F, A);
    else if (dir == 1) hipLaunchKernelGGL((kt_faces<1, RECON, FVS>), agrid(P), kABlock, 0, s, P, F, A);
    else hipLaunchKernelGGL((kt_faces<2, RECON, FVS>), agrid(P), kABlock, 0, s, P, F, A);
}

template <int FVS>
static void tlm_faces_f(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir)
{
    switch (P.recon) {
    case R_WENO5: tlm_faces_r<R_WENO5, FVS>(s, P, F, A, dir); break;
    case R_LINEAR: tlm_faces_r<R_LINEAR, FVS>(s, P, F, A, dir); break;
    case R_MDCD: tlm_faces_r<R_MDCD, FVS>(s, P, F, A, dir); break;
    case R_UD7: tlm_faces_r<R_UD7, FVS>(s, P, F, A, dir); break;
    default: tlm_faces_r<R_UD7L, FVS>(s, P, F, A, dir); break;
    }
}

template <int VIS>
static void tlm_visc(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A)
{
    hipLaunchKernelGGL(kt_tc, agrid(P), kABlock, 0, s, P, F, A);
    hipLaunchKernelGGL(kt_vflux<VIS>, agrid(P), kABlock, 0, s, P, F, A);
    hipLaunchKernelGGL(kt_vdiv<VIS>, agrid(P), kABlock, 0, s, P, A);
}

// Z <- R'(U) X for the state in F (TLM); uses Y, T, V as scratch
hipError_t launch_tlm_rhs(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A)
{
    if (!ad_supported(P)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kt_c2p, agrid(P), kABlock, 0, s, P, F, A);
    for (int d = 0; d < 3; ++d) {
        if (P.fvs == FVS_SW) tlm_faces_f<FVS_SW>(s, P, F, A, d);
        else tlm_faces_f<FVS_LF>(s, P, F, A, d);
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

template <int RECON, int FVS>
static void adj_sweep_r(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int seg)
{
    const int na = dir == 0 ? P.ny : P.nx;
    const int nb = dir == 2 ? P.ny : P.nz;
    const int nd = dir == 0 ? P.nx : dir == 1 ? P.ny : P.nz;
    dim3 g((na + 63) / 64, nb, (nd + seg - 1) / seg);
    if (dir == 0) hipLaunchKernelGGL((ka_sweep<0, RECON, FVS>), g, dim3(64), 0, s, P, F, A, seg);
    else if (dir == 1) hipLaunchKernelGGL((ka_sweep<1, RECON, FVS>), g, dim3(64), 0, s, P, F, A, seg);
    else hipLaunchKernelGGL((ka_sweep<2, RECON, FVS>), g, dim3(64), 0, s, P, F, A, seg);
}

template <int FVS>
static void adj_sweep_f(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir, int seg)
{
    switch (P.recon) {
    case R_WENO5: adj_sweep_r<R_WENO5, FVS>(s, P, F, A, dir, seg); break;
    case R_LINEAR: adj_sweep_r<R_LINEAR, FVS>(s, P, F, A, dir, seg); break;
    case R_MDCD: adj_sweep_r<R_MDCD, FVS>(s, P, F, A, dir, seg); break;
    case R_UD7: adj_sweep_r<R_UD7, FVS>(s, P, F, A, dir, seg); break;
    default: adj_sweep_r<R_UD7L, FVS>(s, P, F, A, dir, seg); break;
    }
}

template <int VIS>
static void adj_visc(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double w_obj,
                     double *partial)
{
    hipLaunchKernelGGL(ka_vflux<VIS>, agrid(P), kABlock, 0, s, P, F, A, w_obj, partial);
    hipLaunchKernelGGL(ka_grad<VIS>, agrid(P), kABlock, 0, s, P, F, A);
}

// Y <- W-adjoint of X^T R(U) (+ w_obj * the dissipation sum), assuming Y = 0
// on entry; block partials of d/d(1/Re) into `partial`
hipError_t launch_adj_rhs(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double w_obj,
                          double *partial, int seg)
{
    if (!ad_supported(P)) return hipErrorInvalidValue;
    for (int d = 0; d < 3; ++d) {
        if (P.fvs == FVS_SW) adj_sweep_f<FVS_SW>(s, P, F, A, d, seg);
        else adj_sweep_f<FVS_LF>(s, P, F, A, d, seg);
    }
    switch (P.vis) {
    case VIS_C8: adj_visc<VIS_C8>(s, P, F, A, w_obj, partial); break;
    case VIS_C6: adj_visc<VIS_C6>(s, P, F, A, w_obj, partial); break;
    default: adj_visc<VIS_C4>(s, P, F, A, w_obj, partial); break;
    }
    return hipGetLastError();
}

hipError_t launch_adj_combine(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double a,
                              double bx, double bl, int init, int add_acc)
{
    hipLaunchKernelGGL(ka_combine, agrid(P), kABlock, 0, s, P, F, A, a, bx, bl, init, add_acc);
    return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t s, const double *partial, int n, double *out)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, n, out);
    return hipGetLastError();
}

hipError_t launch_dissip(hipStream_t s, const DevParams &P, const Fields &F, double *partial)
{
    switch (P.vis) {
    case VIS_C8: hipLaunchKernelGGL(k_dissip<VIS_C8>, agrid(P), kABlock, 0, s, P, F, partial); break;
    case VIS_C6: hipLaunchKernelGGL(k_dissip<VIS_C6>, agrid(P), kABlock, 0, s, P, F, partial); break;
    default: hipLaunchKernelGGL(k_dissip<VIS_C4>, agrid(P), kABlock, 0, s, P, F, partial); break;
    }
    return hipGetLastError();
}

}  // namespace ns3d
