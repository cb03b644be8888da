// ns3d_kernels.hip -- HIP kernels of the FP64 RHS + RK3 path for gfx950 and
// the launch helpers the C ABI (ns3d_api.cpp) calls.
//
// Device state per context (SoA, idx3 layout with ng ghosts, FP64):
//   W[6]  rho u v w p T    primitive state; conserved U = P2C(W) on the fly
//   U0[5]                 conserved state at the start of the RK3 step
//   A[5]                  inviscid accumulator S = (ax + ay) + az, then RHS
//   G[12], VF[12]         velocity/temperature gradients, viscous fluxes
#include <type_traits>

#include "ns3d_device.hpp"
#include "ns3d_internal.hpp"

namespace ns3d {

__device__ __forceinline__ int wrapi(int q, int ng, int n)
{
    return q < ng ? q + n : (q >= ng + n ? q - n : q);
}

// ------------------------------------------------------------- utilities

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// primitiveToConserved's clamp and T over every cell (field_structures.h:436-448)
__global__ void k_prim_setup(DevParams P, Fields F, int ncell)
{
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell) return;
    double rr = F.W[0][id];
    if (rr <= 0.0) { rr = 1e-12; F.W[0][id] = rr; }
    F.W[5][id] = F.W[4][id] / (rr * P.Rgas);
}

// conservedToPrimitive over every cell (field_structures.h:405-427); the
// conserved input sits in src[0..4].
__global__ void k_c2p_all(DevParams P, Fields F, const double *const *src, int ncell)
{
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell) return;
    const double gamma = P.gamma;
    double rr = src[0][id];
    if (rr <= 0.0) rr = 1e-12;
    double u = src[1][id] / rr;
    double v = src[2][id] / rr;
    double w = src[3][id] / rr;
    double kinetic = 0.5 * (u * u + v * v + w * w);
    double e_internal = src[4][id] / rr - kinetic;
    double p = (gamma - 1.0) * rr * e_internal;
    F.W[0][id] = rr; F.W[1][id] = u; F.W[2][id] = v; F.W[3][id] = w; F.W[4][id] = p;
    F.W[5][id] = p / (rr * P.Rgas);
}

// primitiveToConserved (field_structures.h:430-451) into dst[0..4]
__global__ void k_p2c_all(DevParams P, Fields F, double *const *dst, int ncell)
{
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell) return;
    const double gamma = P.gamma;
    double rr = F.W[0][id];
    if (rr <= 0.0) rr = 1e-12;
    double u = F.W[1][id], v = F.W[2][id], w = F.W[3][id], p = F.W[4][id];
    dst[0][id] = rr;
    dst[1][id] = rr * u;
    dst[2][id] = rr * v;
    dst[3][id] = rr * w;
    double kinetic = 0.5 * (u * u + v * v + w * w);
    double e_internal = p / ((gamma - 1.0) * rr);
    dst[4][id] = rr * (e_internal + kinetic);
}

// Periodic ghost fill of the six W arrays: the face regions of
// exchange_halos_physical on one rank (field_structures.h:795-853), cross
// directions restricted to interior cells.  blockIdx.y selects the axis.
__global__ void k_fill_periodic(DevParams P, Fields F)
{
    const int axis = blockIdx.y;
    if (!P.periodic[axis]) return;
    int dst, src;
    if (!periodic_ghost_pair(P, axis, blockIdx.x * blockDim.x + threadIdx.x, dst, src)) return;
#pragma unroll
    for (int m = 0; m < 6; ++m) F.W[m][dst] = F.W[m][src];
}

// apply_wall_bc / apply_symmetry_bc / apply_outflow_bc / apply_inflow_bc
// (boundary.cpp:58-207) on W for one face over its full cross ranges, then T
// of the written cells as primitiveToConserved forms it, with its density
// clamp (field_structures.h:436-448).  Launched face by face in the order of
// apply_boundary (boundary.cpp:24-53), after the periodic fill.
__global__ void k_fill_bc(DevParams P, Fields F, int face, int kind)
{
    const int axis = face >> 1;
    const int na = axis == 0 ? P.sy : P.sx;
    const int nb = axis == 2 ? P.sy : P.sz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.ng * na * nb) return;
    int dst, src, normal;
    bc_cell(P, face, kind, t, dst, src, normal);
    double s[5];
    if (kind == BCK_INFLOW) {
        inflow_state(s);
    } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) s[m] = F.W[m][src];
        if (kind != BCK_OUTFLOW) s[normal] = -s[normal];
    }
    double rr = s[0];
    if (rr <= 0.0) rr = 1e-12;
    F.W[0][dst] = rr;
#pragma unroll
    for (int m = 1; m < 5; ++m) F.W[m][dst] = s[m];
    F.W[5][dst] = s[4] / (rr * P.Rgas);
}

// --------------------------------------------------------- fused sweeps

// One face's flux of one component: the plus and minus reconstructions
// added (InviscousFlux_new.cpp:4-163 through reconstruct_select).  RCP:
// WENO5 by recon_weno5_rcp, bit for bit recon_weno5 on the range of split
// fluxes that weno_rcp_range_ok admits.
template <int RECON, bool RCP>
__device__ __forceinline__ double face_flux(const DevParams &P, const double *sp, const double *sm, double r6)
{
    if constexpr (RCP) {
        double plus_face = recon_weno5_rcp(sp, r6);
        double minus_face = recon_weno5_rcp(sm, r6);
        return plus_face + minus_face;
    } else {
        double plus_face = recon<RECON>(P, sp);
        double minus_face = recon<RECON>(P, sm);
        return plus_face + minus_face;
    }
}
//
// One thread marches one column of cells along the sweep axis DIR over a
// segment [c0, c1), keeping the split fluxes of the S cells of the current
// face stencil in registers: each cell is split once, each face flux is
// reconstructed once (plus one face of warm-up per segment).  Lanes run
// along the fastest cross axis, so for DIR = 1, 2 every load is coalesced.
// The cell result (F(f) - F(f-1)) * inv_d is what range_kutta.cpp:55-97
// forms per direction; ACC = 0 stores it, ACC = 1 adds it to the
// accumulator, so after x, y, z the accumulator holds (ax + ay) + az.
template <int DIR, int RECON, int FVS, int ACC, int MINW>
__global__ __launch_bounds__(64, MINW) void k_sweep(DevParams P, Fields F, int seg_len)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    // grid (nseg, lane tiles, b): after the XCD remap the segments of one
    // column, whose halos overlap, run on the same XCD
    int bs, bt, bb;
    xcd_block3(bs, bt, bb);
    const int na = DIR == 0 ? P.ny : P.nx;
    const int a = bt * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const int b = bb;
    const int nd = DIR == 0 ? P.nx : DIR == 1 ? P.ny : P.nz;
    const int c0 = P.ng + bs * seg_len;
    const int c1 = min(c0 + seg_len, P.ng + nd);
    if (c0 >= c1) return;
    int base, st;
    if (DIR == 0) { base = ((P.ng + b) * P.sy + (P.ng + a)) * P.sx; st = 1; }
    else if (DIR == 1) { base = ((P.ng + b) * P.sy) * P.sx + (P.ng + a); st = P.sx; }
    else { base = (P.ng + b) * P.sx + (P.ng + a); st = P.sx * P.sy; }
    const double inv_d = DIR == 0 ? P.idx : DIR == 1 ? P.idy : P.idz;
    const double *__restrict__ rho = F.W[0];
    const double *__restrict__ uu = F.W[1];
    const double *__restrict__ vv = F.W[2];
    const double *__restrict__ ww = F.W[3];
    const double *__restrict__ pp = F.W[4];

    // WENO5 faces by recon_weno5_rcp while every split flux that has entered
    // the windows of the wave's columns lies in its range (sticky: the
    // literal form from the first flux outside it on)
    constexpr bool WRCP = RECON == R_WENO5;
    const double r6 = WRCP ? rcp_refined_six() : 0.0;
    bool rng_ok = true;
    double fp[S][5], fm[S][5];
#pragma unroll
    for (int s = 1; s < S; ++s) {
        int id = base + (c0 - 2 - MID + s) * st;
        split_state<FVS, DIR>(P, rho[id], uu[id], vv[id], ww[id], pp[id], fp[s], fm[s]);
        if constexpr (WRCP) rng_ok = rng_ok && weno_rcp_range_ok5(fp[s]) && weno_rcp_range_ok5(fm[s]);
    }
    double fprev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // the primitives of the cell entering the window are loaded one face
    // ahead, so their latency hides behind a face's reconstruction
    double nx_[5];
    {
        int id = base + (c0 - 1 - MID + S - 1) * st;
        nx_[0] = rho[id]; nx_[1] = uu[id]; nx_[2] = vv[id]; nx_[3] = ww[id]; nx_[4] = pp[id];
    }
    // y / z: the accumulator of the next cell written, loaded an iteration
    // ahead (a load at the point of the update left its latency exposed)
    double an[5];
    if constexpr (ACC == 1) {
#pragma unroll
        for (int m = 0; m < 5; ++m) an[m] = F.A[m][base + c0 * st];
    }
    double acur[5];
    // window and loads for face f: the cell entering the window is split
    const auto advance = [&](int f) {
#pragma unroll
        for (int s = 0; s < S - 1; ++s)
#pragma unroll
            for (int m = 0; m < 5; ++m) { fp[s][m] = fp[s + 1][m]; fm[s][m] = fm[s + 1][m]; }
        if constexpr (ACC == 1) {
#pragma unroll
            for (int m = 0; m < 5; ++m) acur[m] = an[m];
        }
        double cur[5] = {nx_[0], nx_[1], nx_[2], nx_[3], nx_[4]};
        if (f + 1 < c1) {
            int id = base + (f + 1 - MID + S - 1) * st;
            nx_[0] = rho[id]; nx_[1] = uu[id]; nx_[2] = vv[id]; nx_[3] = ww[id]; nx_[4] = pp[id];
            if constexpr (ACC == 1) {
                if (f >= c0) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) an[m] = F.A[m][base + (f + 1) * st];
                }
            }
        }
        split_state<FVS, DIR>(P, cur[0], cur[1], cur[2], cur[3], cur[4], fp[S - 1], fm[S - 1]);
        if constexpr (WRCP) rng_ok = rng_ok && weno_rcp_range_ok5(fp[S - 1]) && weno_rcp_range_ok5(fm[S - 1]);
    };
    // face f's fluxes, and for f >= c0 the cell's difference stored
    const auto emit = [&](int f, auto rcp) {
        double ff[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sp[S], sm[S];
#pragma unroll
            for (int s = 0; s < S; ++s) { sp[s] = fp[s][m]; sm[s] = fm[S - 1 - s][m]; }
            ff[m] = face_flux<RECON, decltype(rcp)::value>(P, sp, sm, r6);
        }
        if (f >= c0) {
            int idc = base + f * st;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                double ad = (ff[m] - fprev[m]) * inv_d;
                if constexpr (ACC == 0) F.A[m][idc] = ad;
                else F.A[m][idc] = acur[m] + ad;
            }
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) fprev[m] = ff[m];
    };
    int f = c0 - 1;
    if constexpr (WRCP) {
        // the reciprocal form while the whole wave is in range (a uniform
        // exit), then the literal form from the first face that is not: two
        // loops, so the two forms never share a body's registers
        for (; f < c1; ++f) {
            advance(f);
            if (!__all(rng_ok)) break;
            emit(f, std::true_type{});
        }
        if (f < c1) emit(f++, std::false_type{});
    }
    for (; f < c1; ++f) {
        advance(f);
        emit(f, std::false_type{});
    }
}

// x sweep, one thread per cell, one block per full x-row (periodic x).
// Threads split their cell (plus the S-1 halo cells of the row's stencil)
// into LDS (SoA: conflict-free), then each thread reconstructs the right
// face of its cell.  On a periodic axis the row's left boundary face is
// bit-identical to its right boundary face (its stencil cells are the exact
// periodic images), so the first cell takes it from the last thread and no
// face is computed twice.  Loads and stores are coalesced along x.
template <int RECON, int FVS>
__global__ __launch_bounds__(1024) void k_sweep_xrow(DevParams P, Fields F)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    extern __shared__ double lds[];
    const int TX = P.nx;
    const int NSLOT = TX + S - 1;          // split slots: cells ng-MID .. ng+nx-2-MID+S
    double *fpl = lds;                      // [5][NSLOT]
    double *fml = lds + 5 * NSLOT;          // [5][NSLOT]
    double *ffl = lds + 10 * NSLOT;         // [5][TX] face fluxes
    const int t = threadIdx.x;
    const int j = P.ng + blockIdx.x;
    const int k = P.ng + blockIdx.y;
    const int row = (k * P.sy + j) * P.sx;
    const int i0 = P.ng - MID;              // cell of slot 0
    // WENO5 faces by recon_weno5_rcp when every split flux of the row lies
    // in its range, else the literal form for the whole row
    constexpr bool WRCP = RECON == R_WENO5;
    bool rng_ok = true;
    for (int q = t; q < NSLOT; q += blockDim.x) {
        const int id = row + i0 + q;
        double Fp[5], Fm[5];
        split_state<FVS, 0>(P, F.W[0][id], F.W[1][id], F.W[2][id], F.W[3][id], F.W[4][id], Fp, Fm);
        if constexpr (WRCP) rng_ok = rng_ok && weno_rcp_range_ok5(Fp) && weno_rcp_range_ok5(Fm);
#pragma unroll
        for (int m = 0; m < 5; ++m) { fpl[m * NSLOT + q] = Fp[m]; fml[m * NSLOT + q] = Fm[m]; }
    }
    bool row_rcp = false;
    if constexpr (WRCP) row_rcp = __syncthreads_and(rng_ok) != 0;
    else __syncthreads();
    const auto faces = [&](auto rcp) {
        // face ng+t (right face of cell ng+t): stencil cells ng+t-MID .. +S-1
        double r6 = 0.0;
        if constexpr (decltype(rcp)::value) r6 = rcp_refined_six();
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sp[S], sm[S];
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2) {
                sp[s2] = fpl[m * NSLOT + t + s2];
                sm[S - 1 - s2] = fml[m * NSLOT + t + s2];
            }
            ffl[m * TX + t] = face_flux<RECON, decltype(rcp)::value>(P, sp, sm, r6);
        }
    };
    if (t < TX) {
        if constexpr (WRCP) {
            if (row_rcp) faces(std::true_type{});
            else faces(std::false_type{});
        } else {
            faces(std::false_type{});
        }
    }
    __syncthreads();
    if (t < TX) {
        const int tl = t == 0 ? TX - 1 : t - 1;
        const int id = row + P.ng + t;
#pragma unroll
        for (int m = 0; m < 5; ++m) F.A[m][id] = (ffl[m * TX + t] - ffl[m * TX + tl]) * P.idx;
    }
}

// x and y derivatives of u, v, w, T at cell id (ViscousFlux.cpp:240-279)
template <int VIS>
__device__ __forceinline__ void gradxy(const DevParams &P, const double *const src[4], int id, int dxl, int dxh,
                                       int dyl, int dyh, double g[8])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double *c = src[q] + id;
        g[2 * q + 0] = diff_central_bc<VIS>(c, 1, P.dx, dxl, dxh);
        g[2 * q + 1] = diff_central_bc<VIS>(c, P.sx, P.dy, dyl, dyh);
    }
}

// The viscous fluxes of cell id from its gradients g (du/dx, du/dy, du/dz,
// dv/dx, ..., dT/dz), velocity and temperature (ViscousFlux.cpp:282-339)
__device__ __forceinline__ void vflux_store(const DevParams &P, const Fields &F, int id, const double g[12],
                                            double u, double v, double w, double T)
{
    const double Pr = P.Pr, Cp = P.Cp;
    double mu = get_mu(P, T);
    double du_dx = g[0], du_dy = g[1], du_dz = g[2];
    double dv_dx = g[3], dv_dy = g[4], dv_dz = g[5];
    double dw_dx = g[6], dw_dy = g[7], dw_dz = g[8];
    double dT_dx = g[9], dT_dy = g[10], dT_dz = g[11];
    double divU = du_dx + dv_dy + dw_dz;
    double tau_xx = 2.0 * mu * du_dx - (2.0 / 3.0) * mu * divU;
    double tau_yy = 2.0 * mu * dv_dy - (2.0 / 3.0) * mu * divU;
    double tau_zz = 2.0 * mu * dw_dz - (2.0 / 3.0) * mu * divU;
    double tau_xy = mu * (du_dy + dv_dx);
    double tau_xz = mu * (du_dz + dw_dx);
    double tau_yz = mu * (dv_dz + dw_dy);
    double qx = mu * Cp / Pr * dT_dx;
    double qy = mu * Cp / Pr * dT_dy;
    double qz = mu * Cp / Pr * dT_dz;
    F.VF[0][id] = tau_xx;
    F.VF[1][id] = tau_xy;
    F.VF[2][id] = tau_xz;
    F.VF[3][id] = u * tau_xx + v * tau_xy + w * tau_xz + qx;
    F.VF[5][id] = tau_yy;
    F.VF[6][id] = tau_yz;
    F.VF[7][id] = u * tau_xy + v * tau_yy + w * tau_yz + qy;
    F.VF[10][id] = tau_zz;
    F.VF[11][id] = u * tau_xz + v * tau_yz + w * tau_zz + qz;
}

// Gradients (ViscousFlux.cpp:240-279) fused with the viscous flux
// (ViscousFlux.cpp:282-339): one thread per (i, j) column marching along z
// with the R planes above and below held in registers, so the z stencil
// reads each plane once; x and y neighbours come through the caches.  A
// plane's x / y derivatives are taken in the iteration that brings it into
// the z window (R planes ahead of the centre) and held in a register delay
// line: taken R planes later, the plane's lines had left L2 and were read
// from HBM twice.  Writes only the viscous-flux arrays (9 distinct: the
// stress is symmetric and the context aliases VF[4], VF[8], VF[9] to VF[1],
// VF[2], VF[6]).  Same arithmetic and order as k_grad + k_vflux.
template <int VIS, int MINB, int BC>
__global__ __launch_bounds__(256, MINB) void k_gradflux(DevParams P, Fields F, int seg_len)
{
    constexpr int R = VisReach<VIS>::R;
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    const int i = P.ng + bx * blockDim.x + threadIdx.x;
    const int j = P.ng + by * blockDim.y + threadIdx.y;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int k0 = P.ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, P.ng + P.nz);
    const int sxy = P.sx * P.sy;
    const int col = j * P.sx + i;
    const double *const src[4] = {F.W[1], F.W[2], F.W[3], F.W[5]};
    int dxl = R, dxh = R, dyl = R, dyh = R, dzl = R, dzh = R;
    if constexpr (BC) {
        axis_dist(P.periodic[0], i, P.ng, P.nx, R, dxl, dxh);
        axis_dist(P.periodic[1], j, P.ng, P.ny, R, dyl, dyh);
    }
    double win[4][2 * R + 1];
    double gxy[R + 1][8];     // slot s: x / y derivatives of plane k + s
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o) win[q][o] = src[q][(k0 - R - 1 + o) * sxy + col];
#pragma unroll
    for (int s = 1; s <= R; ++s)
        if (k0 + s - 1 < k1) gradxy<VIS>(P, src, (k0 + s - 1) * sxy + col, dxl, dxh, dyl, dyh, gxy[s]);
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
        if constexpr (BC) axis_dist(P.periodic[2], k, P.ng, P.nz, R, dzl, dzh);
#pragma unroll
        for (int s = 0; s < R; ++s)
#pragma unroll
            for (int q = 0; q < 8; ++q) gxy[s][q] = gxy[s + 1][q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[q][o] = win[q][o + 1];
            win[q][2 * R] = src[q][id + R * sxy];
        }
        // past the segment the slot is never used: recompute the last plane
        // (in cache) rather than branch; the branch made the compiler
        // serialise these loads (0.93 ms against 1.23 ms at 256^3)
        gradxy<VIS>(P, src, min(k + R, k1 - 1) * sxy + col, dxl, dxh, dyl, dyh, gxy[R]);
        double g[12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            g[3 * q + 0] = gxy[0][2 * q];
            g[3 * q + 1] = gxy[0][2 * q + 1];
            g[3 * q + 2] = diff_window_bc<VIS>(&win[q][R], P.dz, dzl, dzh);
        }
        vflux_store(P, F, id, g, win[0][R], win[1][R], win[2][R], win[3][R]);
    }
}

// The gradient/flux pass on a box periodic in x and y with each entering
// plane's x / y neighbourhood staged in LDS: a 64 x 4 workgroup writes its
// plane tile with an R-cell halo in x (own rows) and y (own columns; the
// halo cells are the primitive arrays' ghost layers) into one of two
// buffers, one barrier per plane, and takes the x / y derivatives from
// there by diff_central on the tile.  The own values are the z window's
// entering plane; the halo costs up to 2 + 1 loads per thread and quantity
// instead of the 4R stencil loads of k_gradflux's cached form.  The next
// plane's values are loaded before this plane's derivative and flux
// arithmetic.  The same arithmetic in the same order as k_gradflux<VIS, *,
// 0>: equal bit for bit.  Needs nx % 64 == 0 and ny % 4 == 0 (no thread
// leaves before a barrier).
template <int VIS, int TY>
__global__ __launch_bounds__(64 * TY) void k_gradflux_lds(DevParams P, Fields F, int seg_len, int gx, int gy)
{
    constexpr int R = VisReach<VIS>::R;
    constexpr int TR = TY + 2 * R, TC = 64 + 2 * R;   // tile rows, columns
    __shared__ double tile[2][4][TR][TC];
    const int t = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int i0 = P.ng + bx * 64, j0 = P.ng + by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int k0 = P.ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, P.ng + P.nz);
    const int sx = P.sx, sxy = P.sx * P.sy;
    const int col = j * sx + i;
    const double *const src[4] = {F.W[1], F.W[2], F.W[3], F.W[5]};
    // halo cells this thread stages: y-halo rows h = ty, ty + TY, ... < 2R
    // (tile row h, or h + TY above the tile), and one x-halo cell (tid < 8 R TY:
    // quantity tid / (2R TY), own row (tid / 2R) % TY, halo column tid % 2R)
    constexpr int NYH = (2 * R + TY - 1) / TY;
    int yh_row[NYH], yh_off[NYH];
#pragma unroll
    for (int n = 0; n < NYH; ++n) {
        const int h = ty + TY * n;
        const int tr = h < R ? h : h + TY;
        yh_row[n] = h < 2 * R ? tr : -1;
        yh_off[n] = (j0 - R + tr) * sx + i;
    }
    const bool xh_on = tid < 8 * R * TY;
    const int xh_q = xh_on ? tid / (2 * R * TY) : 0;
    const int xh_r = (tid / (2 * R)) % TY;
    const int xh_c = tid % (2 * R);
    const int xh_tc = xh_c < R ? xh_c : 64 + xh_c;
    const int xh_off = (j0 + xh_r) * sx + i0 - R + xh_tc;
    double win[4][2 * R + 1];
    double gxy[R + 1][8];     // slot s: x / y derivatives of plane k + s
    double nyh[NYH][4], nxh = 0.0, nown[4];
    const auto prefetch = [&](int p, bool halo) {
#pragma unroll
        for (int q = 0; q < 4; ++q) nown[q] = src[q][p * sxy + col];
        if (halo) {
#pragma unroll
            for (int n = 0; n < NYH; ++n)
                if (yh_row[n] >= 0)
#pragma unroll
                    for (int q = 0; q < 4; ++q) nyh[n][q] = src[q][p * sxy + yh_off[n]];
            if (xh_on) nxh = src[xh_q][p * sxy + xh_off];
        }
    };
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o) win[q][o] = src[q][(k0 - R - 1 + o) * sxy + col];
#pragma unroll
    for (int s = 1; s <= R; ++s)
        if (k0 + s - 1 < k1) gradxy<VIS>(P, src, (k0 + s - 1) * sxy + col, R, R, R, R, gxy[s]);
    prefetch(k0 + R, k0 + R < k1);
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
        const int p = k + R;   // the entering plane
#pragma unroll
        for (int s = 0; s < R; ++s)
#pragma unroll
            for (int q = 0; q < 8; ++q) gxy[s][q] = gxy[s + 1][q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[q][o] = win[q][o + 1];
            win[q][2 * R] = nown[q];
        }
        const bool stage = p < k1;   // uniform over the workgroup
        if (stage) {
            double (*T)[TR][TC] = tile[p & 1];
#pragma unroll
            for (int q = 0; q < 4; ++q) T[q][R + ty][R + tx] = nown[q];
#pragma unroll
            for (int n = 0; n < NYH; ++n)
                if (yh_row[n] >= 0)
#pragma unroll
                    for (int q = 0; q < 4; ++q) T[q][yh_row[n]][R + tx] = nyh[n][q];
            if (xh_on) T[xh_q][R + xh_r][xh_tc] = nxh;
            __syncthreads();
        }
        if (p + 1 < k1 + R) prefetch(p + 1, p + 1 < k1);
        if (stage) {
            double (*T)[TR][TC] = tile[p & 1];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double *c = &T[q][R + ty][R + tx];
                gxy[R][2 * q + 0] = diff_central<VIS>(c, 1, P.dx);
                gxy[R][2 * q + 1] = diff_central<VIS>(c, TC, P.dy);
            }
        }
        double g[12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            g[3 * q + 0] = gxy[0][2 * q];
            g[3 * q + 1] = gxy[0][2 * q + 1];
            g[3 * q + 2] = diff_window<VIS>(&win[q][R], P.dz);
        }
        vflux_store(P, F, id, g, win[0][R], win[1][R], win[2][R], win[3][R]);
    }
}

// ------------------------------------------------ face-flux path (unfused)

// build_eigen_matrices_new (InviscousFlux_new.cpp:532-625) and
// computeRoeAveragedState_new (:628-659)
__device__ void char_matrices(double gamma, const double ul[5], const double ur[5],
                              double nx, double ny, double nz, double Lm[5][5], double Rm[5][5])
{
    double rho_L = ul[0], u_L = ul[1], v_L = ul[2], w_L = ul[3], p_L = ul[4];
    double E_L = rho_L * (0.5 * (u_L * u_L + v_L * v_L + w_L * w_L) + p_L / ((gamma - 1.0) * rho_L));
    double rho_R = ur[0], u_R = ur[1], v_R = ur[2], w_R = ur[3], p_R = ur[4];
    double E_R = rho_R * (0.5 * (u_R * u_R + v_R * v_R + w_R * w_R) + p_R / ((gamma - 1.0) * rho_R));
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    double denom = 1.0 / (sqrt_rho_L + sqrt_rho_R);
    double u0 = (sqrt_rho_L * u_L + sqrt_rho_R * u_R) * denom;
    double v0 = (sqrt_rho_L * v_L + sqrt_rho_R * v_R) * denom;
    double w0 = (sqrt_rho_L * w_L + sqrt_rho_R * w_R) * denom;
    double H_L = (E_L + p_L) / rho_L;
    double H_R = (E_R + p_R) / rho_R;
    double h0 = (sqrt_rho_L * H_L + sqrt_rho_R * H_R) * denom;
    double V = 0.5 * (u0 * u0 + v0 * v0 + w0 * w0);
    double un = u0 * nx + v0 * ny + w0 * nz;
    double c = sqrt((gamma - 1.0) * (h0 - V));
    Rm[0][0] = 1.0; Rm[1][0] = u0 - c * nx; Rm[2][0] = v0 - c * ny; Rm[3][0] = w0 - c * nz; Rm[4][0] = h0 - un * c;
    Rm[0][1] = 1.0; Rm[1][1] = u0; Rm[2][1] = v0; Rm[3][1] = w0; Rm[4][1] = V;
    Rm[0][2] = 0.0; Rm[1][2] = nz; Rm[2][2] = nx; Rm[3][2] = ny; Rm[4][2] = u0 * nz + v0 * nx + w0 * ny;
    Rm[0][3] = 0.0; Rm[1][3] = ny; Rm[2][3] = nz; Rm[3][3] = nx; Rm[4][3] = u0 * ny + v0 * nz + w0 * nx;
    Rm[0][4] = 1.0; Rm[1][4] = u0 + c * nx; Rm[2][4] = v0 + c * ny; Rm[3][4] = w0 + c * nz; Rm[4][4] = h0 + un * c;
    double chi = 0.5 / c;
    double K = (gamma - 1.0) / (c * c);
    double K1 = K / 2;
    Lm[0][0] = K1 * V + chi * un; Lm[0][1] = -chi * nx - K1 * u0; Lm[0][2] = -chi * ny - K1 * v0;
    Lm[0][3] = -chi * nz - K1 * w0; Lm[0][4] = K1;
    Lm[1][0] = 1.0 - K * V; Lm[1][1] = K * u0; Lm[1][2] = K * v0; Lm[1][3] = K * w0; Lm[1][4] = -K;
    Lm[2][0] = -(u0 * nz + v0 * nx + w0 * ny); Lm[2][1] = nz; Lm[2][2] = nx; Lm[2][3] = ny; Lm[2][4] = 0.0;
    Lm[3][0] = -(u0 * ny + v0 * nz + w0 * nx); Lm[3][1] = ny; Lm[3][2] = nz; Lm[3][3] = nx; Lm[3][4] = 0.0;
    Lm[4][0] = K1 * V - chi * un; Lm[4][1] = chi * nx - K1 * u0; Lm[4][2] = chi * ny - K1 * v0;
    Lm[4][3] = chi * nz - K1 * w0; Lm[4][4] = K1;
}

// One thread per face: du_inviscous (InviscousFlux_new.cpp:4-163) or, with
// CHAR, du_inviscous_character (:165-437).  Output in the reference's face
// layouts idx_fx / idx_fy / idx_fz (field_structures.h:72-81).
template <int DIR, int RECON, int FVS, int CHAR>
__global__ __launch_bounds__(256) void k_faces(DevParams P, Fields F, FaceArrays out)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    const int ng = P.ng;
    // fastest coordinate x-like, then y, then z; the face axis has n+1 faces
    const int n0 = DIR == 0 ? P.nx + 1 : P.nx;
    const int n1 = DIR == 1 ? P.ny + 1 : P.ny;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int t1 = blockIdx.y * blockDim.y + threadIdx.y;
    const int t2 = blockIdx.z;
    if (t0 >= n0 || t1 >= n1) return;
    int i = (DIR == 0 ? ng - 1 : ng) + t0;
    int j = (DIR == 1 ? ng - 1 : ng) + t1;
    int k = (DIR == 2 ? ng - 1 : ng) + t2;
    const int st = DIR == 0 ? 1 : DIR == 1 ? P.sx : P.sx * P.sy;
    const int idf = (k * P.sy + j) * P.sx + i;   // cell left of the face
    double fp[S][5], fm[S][5];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        int id = idf + (s - MID) * st;
        split_state<FVS, DIR>(P, F.W[0][id], F.W[1][id], F.W[2][id], F.W[3][id], F.W[4][id], fp[s], fm[s]);
    }
    double wface[5];
    if constexpr (CHAR) {
        constexpr double nxv = DIR == 0 ? 1.0 : 0.0, nyv = DIR == 1 ? 1.0 : 0.0, nzv = DIR == 2 ? 1.0 : 0.0;
        double ul[5], ur[5], Lm[5][5], Rm[5][5];
        const int ir = idf + st;
        ul[0] = F.W[0][idf]; ul[1] = F.W[1][idf]; ul[2] = F.W[2][idf]; ul[3] = F.W[3][idf]; ul[4] = F.W[4][idf];
        ur[0] = F.W[0][ir]; ur[1] = F.W[1][ir]; ur[2] = F.W[2][ir]; ur[3] = F.W[3][ir]; ur[4] = F.W[4][ir];
        char_matrices(P.gamma, ul, ur, nxv, nyv, nzv, Lm, Rm);
        double wchar[5];
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            double sp[S], sm[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                double ap = 0.0, am = 0.0;
#pragma unroll
                for (int r = 0; r < 5; ++r) { ap += Lm[n][r] * fp[s][r]; am += Lm[n][r] * fm[s][r]; }
                sp[s] = ap;
                sm[S - 1 - s] = am;
            }
            double plus_face = recon<RECON>(P, sp);
            double minus_face = recon<RECON>(P, sm);
            wchar[n] = plus_face + minus_face;
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < 5; ++r) sum += Rm[n][r] * wchar[r];
            wface[n] = sum;
        }
    } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sp[S], sm[S];
#pragma unroll
            for (int s = 0; s < S; ++s) { sp[s] = fp[s][m]; sm[s] = fm[S - 1 - s][m]; }
            double plus_face = recon<RECON>(P, sp);
            double minus_face = recon<RECON>(P, sm);
            wface[m] = plus_face + minus_face;
        }
    }
    int fid;
    if (DIR == 0) fid = (k * P.sy + j) * (P.sx - 1) + i;
    else if (DIR == 1) fid = (k * (P.sy - 1) + j) * P.sx + i;
    else fid = (k * P.sy + j) * P.sx + i;
#pragma unroll
    for (int m = 0; m < 5; ++m) out.f[DIR * 5 + m][fid] = wface[m];
}

// Flux difference of range_kutta.cpp:51-98 from the face arrays into the
// accumulator: A = ((fx_r-fx_l)*idx + (fy_t-fy_b)*idy) + (fz_f-fz_b)*idz.
__global__ void k_face_diff(DevParams P, Fields F, FaceArrays ff)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const int fxr = (k * P.sy + j) * (P.sx - 1) + i;
    const int fyt = (k * (P.sy - 1) + j) * P.sx + i;
    const int fzf = (k * P.sy + j) * P.sx + i;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        double fx_r = ff.f[m][fxr], fx_l = ff.f[m][fxr - 1];
        double fy_t = ff.f[5 + m][fyt], fy_b = ff.f[5 + m][fyt - P.sx];
        double fz_f = ff.f[10 + m][fzf], fz_b = ff.f[10 + m][fzf - P.sx * P.sy];
        F.A[m][id] = (fx_r - fx_l) * P.idx + (fy_t - fy_b) * P.idy + (fz_f - fz_b) * P.idz;
    }
}

// ------------------------------------------------------------- viscous

// compute_gradients (ViscousFlux.cpp:240-279): full order on periodic axes,
// the closures of choose_scheme near the ends of the others.
template <int VIS>
__global__ __launch_bounds__(256) void k_grad(DevParams P, Fields F)
{
    constexpr int R = VisReach<VIS>::R;
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const int sxy = P.sx * P.sy;
    int dxl, dxh, dyl, dyh, dzl, dzh;
    axis_dist(P.periodic[0], i, P.ng, P.nx, R, dxl, dxh);
    axis_dist(P.periodic[1], j, P.ng, P.ny, R, dyl, dyh);
    axis_dist(P.periodic[2], k, P.ng, P.nz, R, dzl, dzh);
    const double *src[4] = {F.W[1], F.W[2], F.W[3], F.W[5]};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double *c = src[q] + id;
        F.G[3 * q + 0][id] = diff_central_bc<VIS>(c, 1, P.dx, dxl, dxh);
        F.G[3 * q + 1][id] = diff_central_bc<VIS>(c, P.sx, P.dy, dyl, dyh);
        F.G[3 * q + 2][id] = diff_central_bc<VIS>(c, sxy, P.dz, dzl, dzh);
    }
}

// compute_viscous_flux (ViscousFlux.cpp:282-339)
__global__ __launch_bounds__(256) void k_vflux(DevParams P, Fields F)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int id = (k * P.sy + j) * P.sx + i;
    const double Pr = P.Pr, Cp = P.Cp;
    double mu = get_mu(P, F.W[5][id]);
    double du_dx = F.G[0][id], du_dy = F.G[1][id], du_dz = F.G[2][id];
    double dv_dx = F.G[3][id], dv_dy = F.G[4][id], dv_dz = F.G[5][id];
    double dw_dx = F.G[6][id], dw_dy = F.G[7][id], dw_dz = F.G[8][id];
    double dT_dx = F.G[9][id], dT_dy = F.G[10][id], dT_dz = F.G[11][id];
    double divU = du_dx + dv_dy + dw_dz;
    double tau_xx = 2.0 * mu * du_dx - (2.0 / 3.0) * mu * divU;
    double tau_yy = 2.0 * mu * dv_dy - (2.0 / 3.0) * mu * divU;
    double tau_zz = 2.0 * mu * dw_dz - (2.0 / 3.0) * mu * divU;
    double tau_xy = mu * (du_dy + dv_dx);
    double tau_xz = mu * (du_dz + dw_dx);
    double tau_yz = mu * (dv_dz + dw_dy);
    double qx = mu * Cp / Pr * dT_dx;
    double qy = mu * Cp / Pr * dT_dy;
    double qz = mu * Cp / Pr * dT_dz;
    double u = F.W[1][id], v = F.W[2][id], w = F.W[3][id];
    F.VF[0][id] = tau_xx;
    F.VF[1][id] = tau_xy;
    F.VF[2][id] = tau_xz;
    F.VF[3][id] = u * tau_xx + v * tau_xy + w * tau_xz + qx;
    F.VF[4][id] = tau_xy;
    F.VF[5][id] = tau_yy;
    F.VF[6][id] = tau_yz;
    F.VF[7][id] = u * tau_xy + v * tau_yy + w * tau_yz + qy;
    F.VF[8][id] = tau_xz;
    F.VF[9][id] = tau_yz;
    F.VF[10][id] = tau_zz;
    F.VF[11][id] = u * tau_xz + v * tau_yz + w * tau_zz + qz;
}

// The end of a divergence/stage pass at cell id: r[1..4] holds the viscous
// divergence (r[0] = 0) and the RHS is r - a, a the inviscid accumulator
// (range_kutta.cpp:11-99).  MODE 0 stores the RHS in A; MODE 1..3 is the
// RK3 stage (range_kutta.cpp:101-161) from the cell's primitives w (rho, u,
// v, w, p; P2C as field_structures.h:430-451) and, for stages 2 and 3, the
// step's start state u0: the new primitives (C2P) and T are stored, and the
// non-physical scan of the stages in P.check_mask reports the cell.
template <int MODE>
__device__ __forceinline__ void vdiv_finish64(const DevParams &P, const Fields &F, int id, double r[5],
                                              const double a[5], const double w5[5], const double u0[5], double dt,
                                              unsigned long long *flag)
{
#pragma unroll
    for (int m = 0; m < 5; ++m) r[m] += -(a[m]);
    if constexpr (MODE == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) F.A[m][id] = r[m];
    } else {
        const double gamma = P.gamma;
        double rr = w5[0], u = w5[1], v = w5[2], w = w5[3], p = w5[4];
        double U[5];
        U[0] = rr;
        U[1] = rr * u;
        U[2] = rr * v;
        U[3] = rr * w;
        {
            double kinetic = 0.5 * (u * u + v * v + w * w);
            double e_internal = p / ((gamma - 1.0) * rr);
            U[4] = rr * (e_internal + kinetic);
        }
        double Un[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            if constexpr (MODE == 1) {
                F.U0[m][id] = U[m];
                Un[m] = U[m] + dt * r[m];
            } else if constexpr (MODE == 2) {
                Un[m] = 0.75 * u0[m] + 0.25 * (U[m] + dt * r[m]);
            } else {
                Un[m] = (1.0 / 3.0) * u0[m] + (2.0 / 3.0) * (U[m] + dt * r[m]);
            }
        }
        double rn = Un[0];
        if (rn <= 0.0) rn = 1e-12;
        double un = Un[1] / rn;
        double vn = Un[2] / rn;
        double wn = Un[3] / rn;
        double kinetic = 0.5 * (un * un + vn * vn + wn * wn);
        double e_internal = Un[4] / rn - kinetic;
        double pn = (gamma - 1.0) * rn * e_internal;
        double Tn = pn / (rn * P.Rgas);
        F.W[0][id] = rn; F.W[1][id] = un; F.W[2][id] = vn; F.W[3][id] = wn; F.W[4][id] = pn;
        F.W[5][id] = Tn;
        // the scan is compiled into every stage and enabled by a runtime
        // mask (the third stage's, range_kutta.cpp:150-161): written for
        // the third stage only, the compiler scheduled the second
        // stage's loads in small groups (1.68 ms against 1.18 ms at 256^3)
        if ((P.check_mask >> MODE) & 1) {
            if (!(rn > 0.0) || !(pn > 0.0) || !(Tn > 0.0))
                atomicMin(flag, nonphysical_key(flag, id));
        }
    }
}

// compute_vis_flux (ViscousFlux.cpp:341-386) + the RHS combine of
// range_kutta.cpp:61-97, and for MODE 1..3 the RK3 stage update of
// range_kutta.cpp:109-147 with the C2P of field_structures.h:405-427.
// One thread per (i, j) column marching along z, the z derivative of Gv
// taken from a register window of the R planes above and below.
// Viscous-flux ghosts on periodic axes are read as periodic images (index
// wrap), which is what exchange_halos_viscous_flux stores there on one
// periodic rank; on the other axes the closures read interior values only.
// MODE 0: A <- rhs.  MODE s = 1,2,3: W <- C2P(stage s update), U0 written at
// stage 1; stage 3 flags non-physical cells (range_kutta.cpp:150-161).
template <int VIS, int MODE, int BC>
__global__ __launch_bounds__(256) void k_vdiv_z(DevParams P, Fields F, double dt, int seg_len,
                                                unsigned long long *flag, int dev_dt)
{
    constexpr int R = VisReach<VIS>::R;
    const int ng = P.ng;
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    const int i = ng + bx * blockDim.x + threadIdx.x;
    const int j = ng + by * blockDim.y + threadIdx.y;
    if (i >= ng + P.nx || j >= ng + P.ny) return;
    const int k0 = ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, ng + P.nz);
    if (MODE != 0 && dev_dt) dt = cfl_dt(P, flag);
    const int sxy = P.sx * P.sy;
    const int col = j * P.sx + i;
    int oxc[2 * R + 1], oyc[2 * R + 1];
#pragma unroll
    for (int o = -R; o <= R; ++o) {
        oxc[o + R] = j * P.sx + wrapi(i + o, ng, P.nx);
        oyc[o + R] = wrapi(j + o, ng, P.ny) * P.sx + i;
    }
    double win[4][2 * R + 1];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o)
            win[m][o] = F.VF[8 + m][wrapi(k0 - R - 1 + o, ng, P.nz) * sxy + col];
    int dxl = R, dxh = R, dyl = R, dyh = R, dzl = R, dzh = R;
    if constexpr (BC) {
        axis_dist(P.periodic[0], i, ng, P.nx, R, dxl, dxh);
        axis_dist(P.periodic[1], j, ng, P.ny, R, dyl, dyh);
    }
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
        const int kn = wrapi(k + R, ng, P.nz) * sxy + col;
        if constexpr (BC) axis_dist(P.periodic[2], k, ng, P.nz, R, dzl, dzh);
        double r[5];
        r[0] = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[m][o] = win[m][o + 1];
            win[m][2 * R] = F.VF[8 + m][kn];
            double wx[2 * R + 1], wy[2 * R + 1];
#pragma unroll
            for (int o = 0; o < 2 * R + 1; ++o) {
                wx[o] = F.VF[m][k * sxy + oxc[o]];
                wy[o] = F.VF[4 + m][k * sxy + oyc[o]];
            }
            double acc = 0.0;
            acc += diff_window_bc<VIS>(wx + R, P.dx, dxl, dxh);
            acc += diff_window_bc<VIS>(wy + R, P.dy, dyl, dyh);
            acc += diff_window_bc<VIS>(&win[m][R], P.dz, dzl, dzh);
            r[m + 1] = acc;
        }
        double a[5], w5[5], u0[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            a[m] = F.A[m][id];
            if constexpr (MODE != 0) w5[m] = F.W[m][id];
            if constexpr (MODE >= 2) u0[m] = F.U0[m][id];
        }
        vdiv_finish64<MODE>(P, F, id, r, a, w5, u0, dt, flag);
    }
}

// The divergence/stage pass on a periodic box with each plane's x / y
// viscous fluxes staged in LDS: a 64 x 4 workgroup writes its plane's x
// fluxes with an R-cell x halo and its y fluxes with an R-row y halo
// (periodic images: the flux arrays hold no ghosts) into one of two
// buffers, one barrier per plane, and takes both derivatives from there.
// Everything the next plane needs (own fluxes, halos, the entering z-flux
// plane, A, the primitives, U0) is loaded before this plane's arithmetic, so
// a plane's loads overlap the previous plane's arithmetic.  The same
// arithmetic in the same order as k_vdiv_z<VIS, MODE, 0>: equal bit for
// bit.  Needs nx % 64 == 0 and ny % 4 == 0 (no thread leaves before a
// barrier).
template <int VIS, int MODE, int TY>
__global__ __launch_bounds__(64 * TY) void k_vdiv_lds(DevParams P, Fields F, double dt, int seg_len, int gx, int gy,
                                                  unsigned long long *flag, int dev_dt)
{
    constexpr int R = VisReach<VIS>::R;
    constexpr int TC = 64 + 2 * R, TYR = TY + 2 * R;
    __shared__ double fxt[2][4][TY][TC];    // x fluxes: tile rows, R-cell x halo
    __shared__ double fyt[2][4][TYR][64];   // y fluxes: tile columns, R-row y halo
    const int ng = P.ng;
    const int t = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int i0 = ng + bx * 64, j0 = ng + by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int k0 = ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, ng + P.nz);
    if (MODE != 0 && dev_dt) dt = cfl_dt(P, flag);
    const int sx = P.sx, sxy = P.sx * P.sy;
    const int col = j * sx + i;
    // the x-halo cell of this thread (tid < 8 R TY): flux tid / (2R TY), tile
    // row (tid / 2R) % TY, halo column tid % 2R
    const bool xh_on = tid < 8 * R * TY;
    const int xh_q = xh_on ? tid / (2 * R * TY) : 0;
    const int xh_r = (tid / (2 * R)) % TY;
    const int xh_c = tid % (2 * R);
    const int xh_tc = xh_c < R ? xh_c : 64 + xh_c;
    const int xh_off = (j0 + xh_r) * sx + wrapi(i0 - R + xh_tc, ng, P.nx);
    // the y-halo rows of this thread: h = ty, ty + TY, ... < 2R, tile row h
    // (below the tile) or h + TY (above it)
    constexpr int NYH = (2 * R + TY - 1) / TY;
    int yh_row[NYH], yh_off[NYH];
#pragma unroll
    for (int n = 0; n < NYH; ++n) {
        const int h = ty + TY * n;
        const int tr = h < R ? h : h + TY;
        yh_row[n] = h < 2 * R ? tr : -1;
        yh_off[n] = wrapi(j0 - R + tr, ng, P.ny) * sx + i;
    }
    double pfx[4], pfy[4], pxh = 0.0, pyh[NYH][4], pz[4], pa[5], pw[5], pu0[5];
    const auto prefetch = [&](int kp) {
        const int pl = kp * sxy;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pfx[q] = F.VF[q][pl + col];
            pfy[q] = F.VF[4 + q][pl + col];
        }
        if (xh_on) pxh = F.VF[xh_q][pl + xh_off];
#pragma unroll
        for (int n = 0; n < NYH; ++n)
            if (yh_row[n] >= 0)
#pragma unroll
                for (int q = 0; q < 4; ++q) pyh[n][q] = F.VF[4 + q][pl + yh_off[n]];
        const int zn = wrapi(kp + R, ng, P.nz) * sxy + col;
#pragma unroll
        for (int m = 0; m < 4; ++m) pz[m] = F.VF[8 + m][zn];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            pa[m] = F.A[m][pl + col];
            if constexpr (MODE != 0) pw[m] = F.W[m][pl + col];
            if constexpr (MODE >= 2) pu0[m] = F.U0[m][pl + col];
        }
    };
    double win[4][2 * R + 1];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o) win[m][o] = F.VF[8 + m][wrapi(k0 - R - 1 + o, ng, P.nz) * sxy + col];
    prefetch(k0);
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[m][o] = win[m][o + 1];
            win[m][2 * R] = pz[m];
        }
        double (*FX)[TY][TC] = fxt[k & 1];
        double (*FY)[TYR][64] = fyt[k & 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            FX[q][ty][R + tx] = pfx[q];
            FY[q][R + ty][tx] = pfy[q];
        }
        if (xh_on) FX[xh_q][xh_r][xh_tc] = pxh;
#pragma unroll
        for (int n = 0; n < NYH; ++n)
            if (yh_row[n] >= 0)
#pragma unroll
                for (int q = 0; q < 4; ++q) FY[q][yh_row[n]][tx] = pyh[n][q];
        double a[5], w5[5], u0[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            a[m] = pa[m];
            w5[m] = MODE != 0 ? pw[m] : 0.0;
            u0[m] = MODE >= 2 ? pu0[m] : 0.0;
        }
        __syncthreads();
        if (k + 1 < k1) prefetch(k + 1);
        double r[5];
        r[0] = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            double wx[2 * R + 1], wy[2 * R + 1];
#pragma unroll
            for (int o = -R; o <= R; ++o) {
                wx[o + R] = FX[m][ty][R + tx + o];
                wy[o + R] = FY[m][R + ty + o][tx];
            }
            double acc = 0.0;
            acc += diff_window<VIS>(wx + R, P.dx);
            acc += diff_window<VIS>(wy + R, P.dy);
            acc += diff_window<VIS>(&win[m][R], P.dz);
            r[m + 1] = acc;
        }
        vdiv_finish64<MODE>(P, F, id, r, a, w5, u0, dt, flag);
    }
}

// ------------------------------------------------------ point functions

// split_state of n states (one per thread) for the batched point API
template <int FVS, int DIR>
__global__ __launch_bounds__(256) void k_split_batch(DevParams P, long long n, SplitBatch b)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double Fp[5], Fm[5];
    split_state<FVS, DIR>(P, b.w[0][t], b.w[1][t], b.w[2][t], b.w[3][t], b.w[4][t], Fp, Fm);
#pragma unroll
    for (int m = 0; m < 5; ++m) { b.fp[m][t] = Fp[m]; b.fm[m][t] = Fm[m]; }
}

// reconstruct_select on n stencils of S values (reversed for flag < 0)
template <int RECON>
__global__ __launch_bounds__(256) void k_recon_batch(DevParams P, long long n, const double *st, double flag,
                                                     double *out)
{
    constexpr int S = Stencil<RECON>::S;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s[S];
#pragma unroll
    for (int q = 0; q < S; ++q) s[q] = st[t * S + (flag < 0.0 ? S - 1 - q : q)];
    out[t] = recon<RECON>(P, s);
}

// the averaging fallback of reconstruction.cpp:77-84 for other schemes
__global__ __launch_bounds__(256) void k_recon_fallback(long long n, int width, const double *st, double flag,
                                                        double *out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double *v = st + t * width;
    if (width >= 2) {
        const int mid = width / 2;
        out[t] = flag >= 0.0 ? 0.5 * (v[mid] + v[mid - 1]) : 0.5 * (v[mid] + v[mid + 1]);
    } else {
        out[t] = width == 0 ? 0.0 : v[0];
    }
}

// ----------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// compute_timestep (timestep.cpp:5-33): min over interior cells.  A bounded
// grid strides over the interior; each block folds its minimum (a positive
// double) into *out with one unsigned 64-bit atomicMin on its bits.
__global__ __launch_bounds__(256) void k_cfl(DevParams P, Fields F, unsigned long long *out)
{
    __shared__ double red[4];
    const int nxy = P.nx * P.ny;
    const int total = nxy * P.nz;
    double m = 1e9;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int k = t / nxy;
        const int r = t - k * nxy;
        const int j = r / P.nx;
        const int i = r - j * P.nx;
        const int id = ((P.ng + k) * P.sy + (P.ng + j)) * P.sx + (P.ng + i);
        m = fmin(m, cfl_cell(P, F.W[0][id], F.W[1][id], F.W[2][id], F.W[3][id], F.W[4][id]));
    }
    m = wave_min(m);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
        atomicMin(out, (unsigned long long)__double_as_longlong(b));
    }
}

// Plain streaming copy dst <- src of n doubles (two per thread per
// iteration, 16-byte accesses): the reference point for achievable HBM
// bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long n2)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// compute_total_energy (postprocess.cpp:103-128): per-block partial sums of
// 0.5*rho*|u|^2*dV, added on the host in block order.
__global__ __launch_bounds__(256) void k_ke(DevParams P, Fields F, double *partial)
{
    __shared__ double red[4];
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = P.ng + blockIdx.z;
    double s = 0.0;
    if (i < P.ng + P.nx && j < P.ng + P.ny) {
        const int id = (k * P.sy + j) * P.sx + i;
        double rho = F.W[0][id], u = F.W[1][id], v = F.W[2][id], w = F.W[3][id];
        double kinetic = 0.5 * rho * (u * u + v * v + w * w);
        s = kinetic * (P.dx * P.dy * P.dz);
    }
    s = wave_sum(s);
    const int lane = (threadIdx.y * blockDim.x + threadIdx.x) & 63;
    const int wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
    if (lane == 0) red[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double b = 0.0;
        for (int q = 0; q < (int)(blockDim.x * blockDim.y) / 64; ++q) b += red[q];
        partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = b;
    }
}

// ------------------------------------------------------ launch helpers

static inline dim3 cell_grid(const DevParams &P)
{
    return dim3((P.nx + 63) / 64, (P.ny + 3) / 4, P.nz);
}
static const dim3 kCellBlock(64, 4, 1);

// end of an RK3 step on the device: advance the step counter of the
// non-physical key and, for the device time step, the simulated time
// t += cfl * min (maincycle.cpp's t += dt, in the same order); dev_dt 2:
// the third stage reduced the next step's CFL minimum, which becomes the
// current one, and its slot starts over
__global__ void k_step_end(unsigned long long *flag, double cfl, int dev_dt)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    flag[FLAG_STEP] = flag[FLAG_STEP] + 1ull;
    if (dev_dt) {
        const double t = __longlong_as_double((long long)flag[FLAG_TIME]) +
                         cfl * __longlong_as_double((long long)flag[FLAG_CFL]);
        flag[FLAG_TIME] = (unsigned long long)__double_as_longlong(t);
    }
    if (dev_dt == 2) {
        flag[FLAG_CFL] = flag[FLAG_CFL_NEXT];
        flag[FLAG_CFL_NEXT] = (unsigned long long)__double_as_longlong(1e9);
    }
}

hipError_t launch_step_end(hipStream_t s, unsigned long long *flag, double cfl, int dev_dt)
{
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, s, flag, cfl, dev_dt);
    return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n)
{
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_prim_setup(hipStream_t s, const DevParams &P, const Fields &F)
{
    int n = P.sx * P.sy * P.sz;
    hipLaunchKernelGGL(k_prim_setup, dim3((n + 255) / 256), dim3(256), 0, s, P, F, n);
    return hipGetLastError();
}

hipError_t launch_c2p_all(hipStream_t s, const DevParams &P, const Fields &F, const double *const *src)
{
    int n = P.sx * P.sy * P.sz;
    hipLaunchKernelGGL(k_c2p_all, dim3((n + 255) / 256), dim3(256), 0, s, P, F, src, n);
    return hipGetLastError();
}

hipError_t launch_p2c_all(hipStream_t s, const DevParams &P, const Fields &F, double *const *dst)
{
    int n = P.sx * P.sy * P.sz;
    hipLaunchKernelGGL(k_p2c_all, dim3((n + 255) / 256), dim3(256), 0, s, P, F, dst, n);
    return hipGetLastError();
}

hipError_t launch_fill_ghosts(hipStream_t s, const DevParams &P, const Fields &F)
{
    int m = P.ny * P.nz;
    if (P.nx * P.nz > m) m = P.nx * P.nz;
    if (P.nx * P.ny > m) m = P.nx * P.ny;
    int total = 2 * P.ng * m;
    hipLaunchKernelGGL(k_fill_periodic, dim3((total + 255) / 256, 3), dim3(256), 0, s, P, F);
    for (int face = 0; face < 6; ++face) {
        const int axis = face >> 1;
        if (P.periodic[axis] || P.bc[face] == BCK_PERIODIC) continue;
        const int n = P.ng * (axis == 0 ? P.sy : P.sx) * (axis == 2 ? P.sy : P.sz);
        hipLaunchKernelGGL(k_fill_bc, dim3((n + 255) / 256), dim3(256), 0, s, P, F, face, P.bc[face]);
    }
    return hipGetLastError();
}

using SweepFn = void (*)(DevParams, Fields, int);

// x stores (ACC = 0), y and z accumulate (ACC = 1); MINW is the waves-per-SIMD
// floor the register allocation must allow (2: no scheme spills; 3: they
// spill).
template <int RECON, int FVS, int MINW>
static SweepFn pick_sweep(int dir)
{
    if (dir == 0) return (SweepFn)k_sweep<0, RECON, FVS, 0, MINW>;
    if (dir == 1) return (SweepFn)k_sweep<1, RECON, FVS, 1, MINW>;
    return (SweepFn)k_sweep<2, RECON, FVS, 1, MINW>;
}

template <int FVS, int MINW>
static SweepFn pick_sweep_r(int recon, int dir)
{
    switch (recon) {
    case R_WENO5: return pick_sweep<R_WENO5, FVS, MINW>(dir);
    case R_LINEAR: return pick_sweep<R_LINEAR, FVS, MINW>(dir);
    case R_MDCD: return pick_sweep<R_MDCD, FVS, MINW>(dir);
    case R_MDCD_HYBRID: return pick_sweep<R_MDCD_HYBRID, FVS, MINW>(dir);
    case R_UD7: return pick_sweep<R_UD7, FVS, MINW>(dir);
    case R_UD7L: return pick_sweep<R_UD7L, FVS, MINW>(dir);
    default: return nullptr;
    }
}

hipError_t launch_sweep(hipStream_t s, const DevParams &P, const Fields &F, int dir, int seg_len, int minw)
{
    SweepFn fn;
    const int fk = fvs_kind(P);
    // 0: the default, 2.  At 3 every scheme spills since the y / z sweeps
    // hold the next cell's accumulator (WENO5 at 3 without it: 1.22 ms per
    // y / z sweep at 256^3; at 2 with it: 1.10 ms)
    if (minw == 0) minw = 2;
    if (minw == 3)
        fn = fk == FVS_SW ? pick_sweep_r<FVS_SW, 3>(P.recon, dir)
             : fk == FVS_SWE ? pick_sweep_r<FVS_SWE, 3>(P.recon, dir) : pick_sweep_r<FVS_LF, 3>(P.recon, dir);
    else
        fn = fk == FVS_SW ? pick_sweep_r<FVS_SW, 2>(P.recon, dir)
             : fk == FVS_SWE ? pick_sweep_r<FVS_SWE, 2>(P.recon, dir) : pick_sweep_r<FVS_LF, 2>(P.recon, dir);
    if (!fn) return hipErrorInvalidValue;
    const int na = dir == 0 ? P.ny : P.nx;
    const int nb = dir == 2 ? P.ny : P.nz;
    const int nd = dir == 0 ? P.nx : dir == 1 ? P.ny : P.nz;
    const int nseg = (nd + seg_len - 1) / seg_len;
    hipLaunchKernelGGL(fn, dim3(nseg, (na + 63) / 64, nb), dim3(64), 0, s, P, F, seg_len);
    return hipGetLastError();
}

using XrowFn = void (*)(DevParams, Fields);

template <int FVS>
static XrowFn pick_xrow(int recon)
{
    switch (recon) {
    case R_WENO5: return k_sweep_xrow<R_WENO5, FVS>;
    case R_LINEAR: return k_sweep_xrow<R_LINEAR, FVS>;
    case R_MDCD: return k_sweep_xrow<R_MDCD, FVS>;
    case R_MDCD_HYBRID: return k_sweep_xrow<R_MDCD_HYBRID, FVS>;
    case R_UD7: return k_sweep_xrow<R_UD7, FVS>;
    case R_UD7L: return k_sweep_xrow<R_UD7L, FVS>;
    default: return nullptr;
    }
}

static size_t xrow_lds_bytes(const DevParams &P)
{
    const int S = P.recon == R_WENO5 || P.recon == R_MDCD || P.recon == R_MDCD_HYBRID ? 6
                  : P.recon == R_LINEAR ? 2 : 8;
    return (size_t)(10 * (P.nx + S - 1) + 5 * P.nx) * sizeof(double);
}

// full-row blocks need a periodic x axis and a row that fits in the LDS of
// one workgroup
bool xrow_supported(const DevParams &P)
{
    return P.periodic[0] && P.nx <= 1024 && xrow_lds_bytes(P) <= (size_t)P.lds_max;
}

hipError_t launch_sweep_xrow(hipStream_t s, const DevParams &P, const Fields &F)
{
    const int fk = fvs_kind(P);
    XrowFn fn = fk == FVS_SW ? pick_xrow<FVS_SW>(P.recon)
                : fk == FVS_SWE ? pick_xrow<FVS_SWE>(P.recon) : pick_xrow<FVS_LF>(P.recon);
    if (!fn || !xrow_supported(P)) return hipErrorInvalidValue;
    const size_t lds = xrow_lds_bytes(P);
    const int threads = ((P.nx + 63) / 64) * 64;
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(fn, dim3(P.ny, P.nz), dim3(threads), lds, s, P, F);
    return hipGetLastError();
}

template <int MINB, int BC>
static hipError_t gradflux_vis(hipStream_t s, const DevParams &P, const Fields &F, int seg_len)
{
    dim3 g((P.nx + 63) / 64, (P.ny + 3) / 4, (P.nz + seg_len - 1) / seg_len);
    switch (P.vis) {
    case VIS_C8: hipLaunchKernelGGL((k_gradflux<VIS_C8, MINB, BC>), g, kCellBlock, 0, s, P, F, seg_len); break;
    case VIS_C6: hipLaunchKernelGGL((k_gradflux<VIS_C6, MINB, BC>), g, kCellBlock, 0, s, P, F, seg_len); break;
    case VIS_C4: hipLaunchKernelGGL((k_gradflux<VIS_C4, MINB, BC>), g, kCellBlock, 0, s, P, F, seg_len); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// the closure-aware variant only where an axis is not periodic
template <int TY>
static hipError_t gradflux_lds_vis(hipStream_t s, const DevParams &P, const Fields &F, int seg_len)
{
    const int gx = P.nx / 64, gy = P.ny / TY, gz = (P.nz + seg_len - 1) / seg_len;
    const dim3 g(gx * gy * gz), b(64, TY);
    switch (P.vis) {
    case VIS_C8: hipLaunchKernelGGL((k_gradflux_lds<VIS_C8, TY>), g, b, 0, s, P, F, seg_len, gx, gy); break;
    case VIS_C6: hipLaunchKernelGGL((k_gradflux_lds<VIS_C6, TY>), g, b, 0, s, P, F, seg_len, gx, gy); break;
    case VIS_C4: hipLaunchKernelGGL((k_gradflux_lds<VIS_C4, TY>), g, b, 0, s, P, F, seg_len, gx, gy); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// the LDS-staged form where it applies (a periodic box with nx % 64 == 0):
// tiles of lds = 8 rows where 8 divides ny, else of 4 where 4 does; lds = 0
// or any other box: k_gradflux with its waves-per-EU floor minb
hipError_t launch_gradflux(hipStream_t s, const DevParams &P, const Fields &F, int seg_len, int minb, int lds)
{
    const bool bc = !(P.periodic[0] && P.periodic[1] && P.periodic[2]);
    if (lds && !bc && P.nx % 64 == 0) {
        if (lds == 8 && P.ny % 8 == 0) return gradflux_lds_vis<8>(s, P, F, seg_len);
        if (P.ny % 4 == 0) return gradflux_lds_vis<4>(s, P, F, seg_len);
    }
    if (minb == 2) return bc ? gradflux_vis<2, 1>(s, P, F, seg_len) : gradflux_vis<2, 0>(s, P, F, seg_len);
    return bc ? gradflux_vis<1, 1>(s, P, F, seg_len) : gradflux_vis<1, 0>(s, P, F, seg_len);
}

template <int VIS, int BC>
static void vdivz_mode(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt, int seg,
                       unsigned long long *flag, int dev_dt)
{
    dim3 g((P.nx + 63) / 64, (P.ny + 3) / 4, (P.nz + seg - 1) / seg);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_vdiv_z<VIS, 0, BC>), g, kCellBlock, 0, s, P, F, dt, seg, flag, dev_dt); break;
    case 1: hipLaunchKernelGGL((k_vdiv_z<VIS, 1, BC>), g, kCellBlock, 0, s, P, F, dt, seg, flag, dev_dt); break;
    case 2: hipLaunchKernelGGL((k_vdiv_z<VIS, 2, BC>), g, kCellBlock, 0, s, P, F, dt, seg, flag, dev_dt); break;
    default: hipLaunchKernelGGL((k_vdiv_z<VIS, 3, BC>), g, kCellBlock, 0, s, P, F, dt, seg, flag, dev_dt); break;
    }
}

template <int BC>
static hipError_t vdivz_vis(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt, int seg,
                            unsigned long long *flag, int dev_dt)
{
    switch (P.vis) {
    case VIS_C8: vdivz_mode<VIS_C8, BC>(s, P, F, mode, dt, seg, flag, dev_dt); break;
    case VIS_C6: vdivz_mode<VIS_C6, BC>(s, P, F, mode, dt, seg, flag, dev_dt); break;
    case VIS_C4: vdivz_mode<VIS_C4, BC>(s, P, F, mode, dt, seg, flag, dev_dt); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// the closure-aware variant only where an axis is not periodic
template <int VIS, int TY>
static void vdiv_lds_mode(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt, int seg,
                          unsigned long long *flag, int dev_dt)
{
    const int gx = P.nx / 64, gy = P.ny / TY, gz = (P.nz + seg - 1) / seg;
    const dim3 g(gx * gy * gz), b(64, TY);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_vdiv_lds<VIS, 0, TY>), g, b, 0, s, P, F, dt, seg, gx, gy, flag, dev_dt); break;
    case 1: hipLaunchKernelGGL((k_vdiv_lds<VIS, 1, TY>), g, b, 0, s, P, F, dt, seg, gx, gy, flag, dev_dt); break;
    case 2: hipLaunchKernelGGL((k_vdiv_lds<VIS, 2, TY>), g, b, 0, s, P, F, dt, seg, gx, gy, flag, dev_dt); break;
    default: hipLaunchKernelGGL((k_vdiv_lds<VIS, 3, TY>), g, b, 0, s, P, F, dt, seg, gx, gy, flag, dev_dt); break;
    }
}

template <int TY>
static hipError_t vdiv_lds_vis(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt, int seg,
                               unsigned long long *flag, int dev_dt)
{
    switch (P.vis) {
    case VIS_C8: vdiv_lds_mode<VIS_C8, TY>(s, P, F, mode, dt, seg, flag, dev_dt); break;
    case VIS_C6: vdiv_lds_mode<VIS_C6, TY>(s, P, F, mode, dt, seg, flag, dev_dt); break;
    case VIS_C4: vdiv_lds_mode<VIS_C4, TY>(s, P, F, mode, dt, seg, flag, dev_dt); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// the LDS-staged form where it applies (a periodic box with nx % 64 == 0):
// tiles of lds = 8 rows where 8 divides ny, else of 4 where 4 does; lds = 0
// or any other box: k_vdiv_z
hipError_t launch_vdiv_z(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt, int seg,
                         unsigned long long *flag, int dev_dt, int lds)
{
    const bool per = P.periodic[0] && P.periodic[1] && P.periodic[2];
    if (lds && per && P.nx % 64 == 0) {
        if (lds == 8 && P.ny % 8 == 0) return vdiv_lds_vis<8>(s, P, F, mode, dt, seg, flag, dev_dt);
        if (P.ny % 4 == 0) return vdiv_lds_vis<4>(s, P, F, mode, dt, seg, flag, dev_dt);
    }
    if (per) return vdivz_vis<0>(s, P, F, mode, dt, seg, flag, dev_dt);
    return vdivz_vis<1>(s, P, F, mode, dt, seg, flag, dev_dt);
}

using FaceFn = void (*)(DevParams, Fields, FaceArrays);

template <int RECON, int FVS, int CHAR>
static FaceFn pick_faces(int dir)
{
    if (dir == 0) return k_faces<0, RECON, FVS, CHAR>;
    if (dir == 1) return k_faces<1, RECON, FVS, CHAR>;
    return k_faces<2, RECON, FVS, CHAR>;
}

template <int FVS, int CHAR>
static FaceFn pick_faces_r(int recon, int dir)
{
    switch (recon) {
    case R_WENO5: return pick_faces<R_WENO5, FVS, CHAR>(dir);
    case R_LINEAR: return pick_faces<R_LINEAR, FVS, CHAR>(dir);
    case R_MDCD: return pick_faces<R_MDCD, FVS, CHAR>(dir);
    case R_MDCD_HYBRID: return pick_faces<R_MDCD_HYBRID, FVS, CHAR>(dir);
    case R_UD7: return pick_faces<R_UD7, FVS, CHAR>(dir);
    case R_UD7L: return pick_faces<R_UD7L, FVS, CHAR>(dir);
    default: return nullptr;
    }
}

hipError_t launch_faces(hipStream_t s, const DevParams &P, const Fields &F, const FaceArrays &out,
                        int dir, int char_recon)
{
    FaceFn fn;
    const int fk = fvs_kind(P);
    if (fk == FVS_SW) fn = char_recon ? pick_faces_r<FVS_SW, 1>(P.recon, dir) : pick_faces_r<FVS_SW, 0>(P.recon, dir);
    else if (fk == FVS_SWE)
        fn = char_recon ? pick_faces_r<FVS_SWE, 1>(P.recon, dir) : pick_faces_r<FVS_SWE, 0>(P.recon, dir);
    else fn = char_recon ? pick_faces_r<FVS_LF, 1>(P.recon, dir) : pick_faces_r<FVS_LF, 0>(P.recon, dir);
    if (!fn) return hipErrorInvalidValue;
    const int n0 = dir == 0 ? P.nx + 1 : P.nx;
    const int n1 = dir == 1 ? P.ny + 1 : P.ny;
    const int n2 = dir == 2 ? P.nz + 1 : P.nz;
    hipLaunchKernelGGL(fn, dim3((n0 + 63) / 64, (n1 + 3) / 4, n2), kCellBlock, 0, s, P, F, out);
    return hipGetLastError();
}

hipError_t launch_face_diff(hipStream_t s, const DevParams &P, const Fields &F, const FaceArrays &ff)
{
    hipLaunchKernelGGL(k_face_diff, cell_grid(P), kCellBlock, 0, s, P, F, ff);
    return hipGetLastError();
}

hipError_t launch_grad(hipStream_t s, const DevParams &P, const Fields &F)
{
    switch (P.vis) {
    case VIS_C8: hipLaunchKernelGGL(k_grad<VIS_C8>, cell_grid(P), kCellBlock, 0, s, P, F); break;
    case VIS_C6: hipLaunchKernelGGL(k_grad<VIS_C6>, cell_grid(P), kCellBlock, 0, s, P, F); break;
    case VIS_C4: hipLaunchKernelGGL(k_grad<VIS_C4>, cell_grid(P), kCellBlock, 0, s, P, F); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_vflux(hipStream_t s, const DevParams &P, const Fields &F)
{
    hipLaunchKernelGGL(k_vflux, cell_grid(P), kCellBlock, 0, s, P, F);
    return hipGetLastError();
}

template <int FVS>
static void split_batch_dir(hipStream_t s, const DevParams &P, long long n, const SplitBatch &b, int dir)
{
    const dim3 g((unsigned)((n + 255) / 256));
    if (dir == 0) hipLaunchKernelGGL((k_split_batch<FVS, 0>), g, dim3(256), 0, s, P, n, b);
    else if (dir == 1) hipLaunchKernelGGL((k_split_batch<FVS, 1>), g, dim3(256), 0, s, P, n, b);
    else hipLaunchKernelGGL((k_split_batch<FVS, 2>), g, dim3(256), 0, s, P, n, b);
}

hipError_t launch_split_batch(hipStream_t s, const DevParams &P, long long n, const SplitBatch &b, int dir)
{
    const int fk = fvs_kind(P);
    if (fk == FVS_SW) split_batch_dir<FVS_SW>(s, P, n, b, dir);
    else if (fk == FVS_SWE) split_batch_dir<FVS_SWE>(s, P, n, b, dir);
    else split_batch_dir<FVS_LF>(s, P, n, b, dir);
    return hipGetLastError();
}

hipError_t launch_recon_batch(hipStream_t s, const DevParams &P, long long n, int width, const double *st,
                              double flag, double *out)
{
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    switch (P.recon) {
    case R_WENO5: hipLaunchKernelGGL(k_recon_batch<R_WENO5>, g, b, 0, s, P, n, st, flag, out); break;
    case R_LINEAR: hipLaunchKernelGGL(k_recon_batch<R_LINEAR>, g, b, 0, s, P, n, st, flag, out); break;
    case R_MDCD: hipLaunchKernelGGL(k_recon_batch<R_MDCD>, g, b, 0, s, P, n, st, flag, out); break;
    case R_MDCD_HYBRID: hipLaunchKernelGGL(k_recon_batch<R_MDCD_HYBRID>, g, b, 0, s, P, n, st, flag, out); break;
    case R_UD7: hipLaunchKernelGGL(k_recon_batch<R_UD7>, g, b, 0, s, P, n, st, flag, out); break;
    case R_UD7L: hipLaunchKernelGGL(k_recon_batch<R_UD7L>, g, b, 0, s, P, n, st, flag, out); break;
    default: hipLaunchKernelGGL(k_recon_fallback, g, b, 0, s, n, width, st, flag, out); break;
    }
    return hipGetLastError();
}

hipError_t launch_cfl(hipStream_t s, const DevParams &P, const Fields &F, unsigned long long *out)
{
    const long total = (long)P.nx * P.ny * P.nz;
    long blocks = (total + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_cfl, dim3((unsigned)blocks), dim3(256), 0, s, P, F, out);
    return hipGetLastError();
}

// The same copy with one 16-byte element per thread (a flat grid) and
// nontemporal loads and stores: the best streaming rate measured on the
// device, reported beside the 8 TB/s peak.
__global__ __launch_bounds__(256) void k_copy_nt(const double2 *__restrict__ src, double2 *__restrict__ dst, long n2)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) {
        double2 v;
        v.x = __builtin_nontemporal_load(&src[t].x);
        v.y = __builtin_nontemporal_load(&src[t].y);
        __builtin_nontemporal_store(v.x, &dst[t].x);
        __builtin_nontemporal_store(v.y, &dst[t].y);
    }
}

hipError_t launch_copy_nt(hipStream_t s, const double *src, double *dst, long n)
{
    const long n2 = n / 2;
    const long blocks = (n2 + 255) / 256;
    if (blocks > 2147483647L) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy_nt, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(src), reinterpret_cast<double2 *>(dst), n2);
    return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, const double *src, double *dst, long n)
{
    const long n2 = n / 2;
    long blocks = (n2 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(src), reinterpret_cast<double2 *>(dst), n2);
    return hipGetLastError();
}

int ke_blocks(const DevParams &P)
{
    dim3 g = cell_grid(P);
    return (int)(g.x * g.y * g.z);
}

hipError_t launch_ke(hipStream_t s, const DevParams &P, const Fields &F, double *partial)
{
    hipLaunchKernelGGL(k_ke, cell_grid(P), kCellBlock, 0, s, P, F, partial);
    return hipGetLastError();
}

}  // namespace ns3d
