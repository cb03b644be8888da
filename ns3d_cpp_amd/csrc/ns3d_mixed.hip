// ns3d_mixed.hip -- mixed-precision RHS + RK3 path (NS3D_PREC_MIXED).
//
// State and arithmetic split:
//   FP64  conserved state U, the FVS split, the face fluxes and their
//         differences, the RK3 stage update, and p/T recovery from E minus
//         the kinetic energy (C2P).  The RK3 stages hold their increment
//         over U in FP32.
//   FP32  a shadow of the primitives stored as deviations from the
//         nondimensional reference state (rho-1, u, v, w, p-p_inf, T-1;
//         p_inf = 1/(gamma Ma^2)), the reconstructions (weights and smoothness
//         indicators included), the gradients, viscous fluxes and their
//         divergence, and the inviscid accumulator.
// The FP32 reconstructions act on split-flux residuals F(s) - F(ref), with
// the reference taken at the face's upwind cell and added back in FP64:
// every reconstruction here is translation invariant (weights depend on
// differences only, linear stencils sum to 1), so the low-Mach pressure
// part of the fluxes never passes through FP32.
#include "ns3d_device.hpp"
#include "ns3d_internal.hpp"

namespace ns3d {

// ------------------------------------------------ FP32 reconstructions

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ float recon32_weno5(const float *s)
{
    float f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
    const float eps = 1e-6f;
    float d0 = f0 - 2.0f * f1 + f2, e0 = f0 - 4.0f * f1 + 3.0f * f2;
    float d1 = f1 - 2.0f * f2 + f3, e1 = f1 - f3;
    float d2 = f2 - 2.0f * f3 + f4, e2 = 3.0f * f2 - 4.0f * f3 + f4;
    float b0 = (13.0f / 12.0f) * d0 * d0 + 0.25f * e0 * e0;
    float b1 = (13.0f / 12.0f) * d1 * d1 + 0.25f * e1 * e1;
    float b2 = (13.0f / 12.0f) * d2 * d2 + 0.25f * e2 * e2;
    float t0 = eps + b0, t1 = eps + b1, t2 = eps + b2;
    float a0 = 0.1f * rcpf(t0 * t0);
    float a1 = 0.6f * rcpf(t1 * t1);
    float a2 = 0.3f * rcpf(t2 * t2);
    float inv = rcpf(a0 + a1 + a2) * (1.0f / 6.0f);
    float p0 = 2.0f * f0 - 7.0f * f1 + 11.0f * f2;
    float p1 = -f1 + 5.0f * f2 + 2.0f * f3;
    float p2 = 2.0f * f2 + 5.0f * f3 - f4;
    return (a0 * p0 + a1 * p1 + a2 * p2) * inv;
}

__device__ __forceinline__ float recon32_mdcd(const float *s, float disp, float diss)
{
    return (0.5f * disp + 0.5f * diss) * s[0] +
           (-1.5f * disp - 2.5f * diss - 1.0f / 12.0f) * s[1] +
           (7.0f / 12.0f + disp + 5.0f * diss) * s[2] +
           (7.0f / 12.0f + disp - 5.0f * diss) * s[3] +
           (-1.5f * disp + 2.5f * diss - 1.0f / 12.0f) * s[4] +
           (0.5f * disp - 0.5f * diss) * s[5];
}

__device__ __forceinline__ float recon32_ud7(const float *s)
{
    return (-1.0f / 140.0f) * s[0] + (5.0f / 84.0f) * s[1] - (101.0f / 420.0f) * s[2]
           + (319.0f / 420.0f) * s[3] + (107.0f / 210.0f) * s[4]
           - (19.0f / 210.0f) * s[5] + (1.0f / 105.0f) * s[6];
}

__device__ __forceinline__ float recon32_ud7l(const float *s, float alpha)
{
    float b = 1.0f - alpha;
    float a0 = (-3.0f / 420.0f) * alpha - (3.0f / 840.0f) * b;
    float a1 = (25.0f / 420.0f) * alpha + (29.0f / 840.0f) * b;
    float a2 = (-101.0f / 420.0f) * alpha - (139.0f / 840.0f) * b;
    float a3 = (319.0f / 420.0f) * alpha + (533.0f / 840.0f) * b;
    float a4 = (214.0f / 420.0f) * alpha + (533.0f / 840.0f) * b;
    float a5 = (-38.0f / 420.0f) * alpha - (139.0f / 840.0f) * b;
    float a6 = (4.0f / 420.0f) * alpha + (29.0f / 840.0f) * b;
    float a7 = (-3.0f / 840.0f) * b;
    return a0 * s[0] + a1 * s[1] + a2 * s[2] + a3 * s[3] + a4 * s[4] + a5 * s[5] + a6 * s[6] + a7 * s[7];
}

template <int RECON>
__device__ __forceinline__ float recon32(const DevParams &P, const float *s)
{
    if constexpr (RECON == R_WENO5) return recon32_weno5(s);
    else if constexpr (RECON == R_LINEAR) return s[0];
    else if constexpr (RECON == R_MDCD) return recon32_mdcd(s, (float)P.mdcd_disp, (float)P.mdcd_diss);
    else if constexpr (RECON == R_MDCD_HYBRID) return 0.0f;   // never used: see face_mixed
    else if constexpr (RECON == R_UD7) return recon32_ud7(s);
    else return recon32_ud7l(s, (float)P.ud7l_alpha);
}

// Face flux of component m from FP64 split values of the S stencil cells:
// residuals against the upwind cell of each side go through FP32.
template <int RECON, int S>
__device__ __forceinline__ double face_mixed(const DevParams &P, const double *fp, const double *fm)
{
    constexpr int MID = (S - 1) / 2;
    if constexpr (RECON == R_MDCD_HYBRID) {
        // its 6-point indicator beta3 cancels terms ~2e7 times larger than
        // its value: input rounding to FP32 alone would leave O(1) relative
        // error in it, so this option is reconstructed in FP64 throughout
        double dm[S];
#pragma unroll
        for (int s = 0; s < S; ++s) dm[S - 1 - s] = fm[s];
        return recon_mdcd_hybrid(fp, P.mdcd_disp, P.mdcd_diss) + recon_mdcd_hybrid(dm, P.mdcd_disp, P.mdcd_diss);
    }
    const double rp = fp[MID], rm = fm[MID + 1 < S ? MID + 1 : S - 1];
    float sp[S], sm[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        sp[s] = (float)(fp[s] - rp);
        sm[S - 1 - s] = (float)(fm[s] - rm);
    }
    return ((double)recon32<RECON>(P, sp) + rp) + ((double)recon32<RECON>(P, sm) + rm);
}

__device__ __forceinline__ void shadow_prims(const DevParams &P, const MFields &M, int id, double &rho, double &u,
                                             double &v, double &w, double &p)
{
    rho = P.rho_ref + (double)M.S[0][id];
    u = (double)M.S[1][id];
    v = (double)M.S[2][id];
    w = (double)M.S[3][id];
    p = P.p_ref + (double)M.S[4][id];
}

// ---------------------------------------------------------------- sweeps

// x sweep with FP64 split values in LDS, one block per full x-row (as
// k_sweep_xrow); used for MDCD_HYBRID, which reconstructs in FP64.  PER = 0
// (x not periodic) computes the left face of the first cell as well.
template <int RECON, int FVS, int PER>
__global__ __launch_bounds__(1024) void km_sweep_xrow64(DevParams P, MFields M)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    constexpr int SH = PER ? 0 : 1;
    extern __shared__ double lds[];
    const int TX = P.nx;
    const int NSLOT = TX + S - 1 + SH;
    const int NF = TX + SH;                 // faces, the first right of cell ng - SH
    double *fpl = lds;
    double *fml = lds + 5 * NSLOT;
    double *ffl = lds + 10 * NSLOT;
    const int t = threadIdx.x;
    const int j = P.ng + blockIdx.x;
    const int k = P.ng + blockIdx.y;
    const int row = (k * P.sy + j) * P.sx;
    const int i0 = P.ng - MID - SH;
    for (int q = t; q < NSLOT; q += blockDim.x) {
        double rho, u, v, w, p, Fp[5], Fm[5];
        shadow_prims(P, M, row + i0 + q, rho, u, v, w, p);
        split_state<FVS, 0>(P, rho, u, v, w, p, Fp, Fm);
#pragma unroll
        for (int m = 0; m < 5; ++m) { fpl[m * NSLOT + q] = Fp[m]; fml[m * NSLOT + q] = Fm[m]; }
    }
    __syncthreads();
    for (int f = t; f < NF; f += blockDim.x) {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sp[S], sm[S];
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2) { sp[s2] = fpl[m * NSLOT + f + s2]; sm[s2] = fml[m * NSLOT + f + s2]; }
            ffl[m * NF + f] = face_mixed<RECON, S>(P, sp, sm);
        }
    }
    __syncthreads();
    if (t < TX) {
        const int fr = t + SH;
        const int fl = PER ? (t == 0 ? TX - 1 : t - 1) : t;
        const int id = row + P.ng + t;
#pragma unroll
        for (int m = 0; m < 5; ++m) M.A[m][id] = (float)((ffl[m * NF + fr] - ffl[m * NF + fl]) * P.idx);
    }
}

// y / z sweeps with an FP64 split window (as k_sweep); used for MDCD_HYBRID.
template <int DIR, int RECON, int FVS>
__global__ __launch_bounds__(64, 2) void km_sweep64(DevParams P, MFields M, int seg_len)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    int bs, bt, bb;
    xcd_block3(bs, bt, bb);   // grid (nseg, lane tiles, b), as km_sweep's launch
    const int a = bt * blockDim.x + threadIdx.x;
    if (a >= P.nx) return;
    const int b = bb;
    const int nd = DIR == 1 ? P.ny : P.nz;
    const int c0 = P.ng + bs * seg_len;
    const int c1 = min(c0 + seg_len, P.ng + nd);
    if (c0 >= c1) return;
    int base, st;
    if (DIR == 1) { base = ((P.ng + b) * P.sy) * P.sx + (P.ng + a); st = P.sx; }
    else { base = (P.ng + b) * P.sx + (P.ng + a); st = P.sx * P.sy; }
    const double inv_d = DIR == 1 ? P.idy : P.idz;
    double fp[S][5], fm[S][5];
#pragma unroll
    for (int s = 1; s < S; ++s) {
        double rho, u, v, w, p;
        shadow_prims(P, M, base + (c0 - 2 - MID + s) * st, rho, u, v, w, p);
        split_state<FVS, DIR>(P, rho, u, v, w, p, fp[s], fm[s]);
    }
    float nx_[5];
    {
        const int id = base + (c0 - 1 - MID + S - 1) * st;
#pragma unroll
        for (int q = 0; q < 5; ++q) nx_[q] = M.S[q][id];
    }
    double fprev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int f = c0 - 1; f < c1; ++f) {
#pragma unroll
        for (int s = 0; s < S - 1; ++s)
#pragma unroll
            for (int m = 0; m < 5; ++m) { fp[s][m] = fp[s + 1][m]; fm[s][m] = fm[s + 1][m]; }
        {
            float cur[5] = {nx_[0], nx_[1], nx_[2], nx_[3], nx_[4]};
            if (f + 1 < c1) {
                const int id = base + (f + 1 - MID + S - 1) * st;
#pragma unroll
                for (int q = 0; q < 5; ++q) nx_[q] = M.S[q][id];
            }
            split_state<FVS, DIR>(P, P.rho_ref + (double)cur[0], (double)cur[1], (double)cur[2], (double)cur[3],
                                  P.p_ref + (double)cur[4], fp[S - 1], fm[S - 1]);
        }
        double ff[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            double sp[S], sm[S];
#pragma unroll
            for (int s = 0; s < S; ++s) { sp[s] = fp[s][m]; sm[s] = fm[s][m]; }
            ff[m] = face_mixed<RECON, S>(P, sp, sm);
        }
        if (f >= c0) {
            const int idc = base + f * st;
#pragma unroll
            for (int m = 0; m < 5; ++m)
                M.A[m][idc] = (float)((double)M.A[m][idc] + (ff[m] - fprev[m]) * inv_d);
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) fprev[m] = ff[m];
    }
}

// ------------------------------------- reconstruction in difference form
//
// The split fluxes are formed in FP64 (their low-Mach pressure and enthalpy
// parts are large) and each cell's difference to its neighbour along the
// sweep, d(c) = F(c) - F(c-1), is rounded to FP32 once.  A face value is the
// upwind cell's value plus a reconstruction of the differences; every scheme
// here is translation invariant (WENO5 weights see differences only, linear
// stencils sum to 1), so that is the same face value.  The flux difference of
// cell c then needs no FP64 face value at all:
//   F(c) - F(c-1) = d+(c) + d-(c+1) + [rel(c) - rel(c-1)],
// rel(f) = rel+(f) + rel-(f), the two sides' reconstructed differences.
// D[1..S-1] are the stencil's differences in reconstruct_select's
// orientation: D[t] = f[t] - f[t-1].

// WENO5-JS face value minus f[2] from D[1..4]: one reciprocal for the
// weights, after scaling the (eps + beta)^2 near their maximum (an exact
// power of two from the exponent bits: the y / z sweeps ran 2.5 % faster
// than with a second, quarter-rate, reciprocal).  The weights see only
// ratios of the T_k and of the alpha_k, so t_k = 4 (eps + beta_k) (the 1/4
// of beta's second square folded away) and alpha_k scaled by 10 (linear
// weights 1, 6, 3) leave them unchanged and save a multiply each.  The 1/6
// of the candidate polynomials goes into the reciprocal's argument: a float
// 1/6 applied after it scaled every reconstructed difference of a run by
// the same factor 1 + 3.0e-8, and that bias accumulated over the steps.
// The first term of a WENO5 smoothness indicator, scaled by 4 (the weights
// see ratios only): 4 eps + (13/3) s^2, s the stencil's second difference.
// A three-cell stencil's s is the same at the three faces that use it, so the
// marching sweeps carry it from face to face.
__device__ __forceinline__ float weno5_u(float s)
{
    return fmaf((13.0f / 3.0f) * s, s, 4e-6f);
}

// rel32_weno5 with the smoothness indicators' first terms u0, u1, u2 given
__device__ __forceinline__ float rel32_weno5_u(const float *D, float u0, float u1, float u2)
{
    const float D1 = D[1], D2 = D[2], D3 = D[3], D4 = D[4];
    const float r0 = 3.0f * D2 - D1, r1 = D2 + D3, r2 = D4 - 3.0f * D3;
    const float t0 = fmaf(r0, r0, u0);
    const float t1 = fmaf(r1, r1, u1);
    const float t2 = fmaf(r2, r2, u2);
    float T0 = t0 * t0, T1 = t1 * t1, T2 = t2 * t2;
    // scale by the power of two 2^(127 - e), e the biased exponent of the
    // largest T (T >= 1.6e-11 > 0, so e is in 91..254): the largest scaled T
    // is in [1, 2), exactly, with integer ops instead of a reciprocal
    const unsigned e = (__float_as_uint(fmaxf(T0, fmaxf(T1, T2))) >> 23) & 0xffu;
    const float im = __uint_as_float((254u - e) << 23);
    T0 *= im; T1 *= im; T2 *= im;
    // 10 alpha_k = 10 d_k / T_k, scaled by T0 T1 T2 (the largest T is < 2)
    const float a0 = T1 * T2, a1 = 6.0f * T0 * T2, a2 = 3.0f * T0 * T1;
    const float q0 = 2.0f * r0 - D2;          // 5 D2 - 2 D1
    const float q1 = D2 + 2.0f * D3;
    const float q2 = 4.0f * D3 - D4;
    return (a0 * q0 + a1 * q1 + a2 * q2) * rcpf(6.0f * (a0 + a1 + a2));
}

__device__ __forceinline__ float rel32_weno5(const float *D)
{
    return rel32_weno5_u(D, weno5_u(D[2] - D[1]), weno5_u(D[3] - D[2]), weno5_u(D[4] - D[3]));
}

// coefficients e[t] of a linear scheme in difference form:
// sum_s c_s f_s - f_MID = sum_t e_t D_t, e_t = sum_{s>=t} c_s (t > MID),
// -sum_{s<t} c_s (t <= MID)
template <int RECON>
__device__ __forceinline__ void diff_coeffs(const DevParams &P, float *e)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    double c[S];
    if constexpr (RECON == R_LINEAR) {
        c[0] = 1.0; c[1] = 0.0;
    } else if constexpr (RECON == R_MDCD) {
        const double disp = P.mdcd_disp, diss = P.mdcd_diss;
        c[0] = 0.5 * disp + 0.5 * diss;
        c[1] = -1.5 * disp - 2.5 * diss - 1.0 / 12.0;
        c[2] = 7.0 / 12.0 + disp + 5.0 * diss;
        c[3] = 7.0 / 12.0 + disp - 5.0 * diss;
        c[4] = -1.5 * disp + 2.5 * diss - 1.0 / 12.0;
        c[5] = 0.5 * disp - 0.5 * diss;
    } else if constexpr (RECON == R_UD7) {
        // 420 c: e comes out as integers, exact in FP32 (rel_scale applies
        // the 1/420)
        c[0] = -3.0; c[1] = 25.0; c[2] = -101.0; c[3] = 319.0;
        c[4] = 214.0; c[5] = -38.0; c[6] = 4.0; c[7] = 0.0;
    } else {
        const double a = P.ud7l_alpha, b = 1.0 - a;
        c[0] = -3.0 / 420.0 * a - 3.0 / 840.0 * b;
        c[1] = 25.0 / 420.0 * a + 29.0 / 840.0 * b;
        c[2] = -101.0 / 420.0 * a - 139.0 / 840.0 * b;
        c[3] = 319.0 / 420.0 * a + 533.0 / 840.0 * b;
        c[4] = 214.0 / 420.0 * a + 533.0 / 840.0 * b;
        c[5] = -38.0 / 420.0 * a - 139.0 / 840.0 * b;
        c[6] = 4.0 / 420.0 * a + 29.0 / 840.0 * b;
        c[7] = -3.0 / 840.0 * b;
    }
#pragma unroll
    for (int t = 1; t < S; ++t) {
        double acc = 0.0;
        if (t > MID) {
#pragma unroll
            for (int q = 0; q < S; ++q) acc += q >= t ? c[q] : 0.0;
        } else {
#pragma unroll
            for (int q = 0; q < S; ++q) acc -= q < t ? c[q] : 0.0;
        }
        e[t] = (float)acc;
    }
}

// The reconstructed differences of both sides of a face, scaled: UD7's
// weights are applied as integers (diff_coeffs) and their 1/420 here as a
// float pair.  Weights rounded to FP32 would change the scheme by the same
// relative amount at every face of a run (its dissipation, in effect),
// which accumulates over the steps instead of averaging out.  MDCD's and
// UD7L's weights depend on run parameters and keep their FP32 rounding.
template <int RECON>
__device__ __forceinline__ float rel_scale(float r)
{
    if constexpr (RECON == R_UD7) {
        constexpr F2 K = {(float)(1.0 / 420.0), (float)(1.0 / 420.0 - (double)(float)(1.0 / 420.0))};
        return mul2(r, K);
    } else {
        return r;
    }
}

template <int RECON>
__device__ __forceinline__ float rel32(const float *e, const float *D)
{
    constexpr int S = Stencil<RECON>::S;
    if constexpr (RECON == R_WENO5) {
        return rel32_weno5(D);
    } else {
        float r = e[1] * D[1];
#pragma unroll
        for (int t = 2; t < S; ++t) r += e[t] * D[t];
        return r;
    }
}

// split_state with reciprocals instead of divisions and explicit FMAs (this
// path does not reproduce the reference's rounding).  Contraction is off
// here and every FMA is written out: the sweeps call it from several places
// and difference the results of neighbouring cells, so a uniform state must
// give bitwise equal values whichever call site evaluated them (free-stream
// preservation).  Without eps the eigenvalue split is max(l, 0), min(l, 0),
// which 0.5 (l +- |l|) equals up to the sign of a zero.
template <int FVS, int DIR>
__device__ __forceinline__ void split_fast(const DevParams &P, double rho, double u, double v, double w, double p,
                                           double Fp[5], double Fm[5])
{
#pragma clang fp contract(off)
    const double gamma = P.gamma;
    const double irho = rcp64(rho);
    const double V = 0.5 * fma(w, w, fma(v, v, u * u));
    const double ig1 = 1.0 / (gamma - 1.0);
    const double E = fma(rho, V, p * ig1);
    const double c2 = gamma * p * irho;
    const double c = sqrt(c2);
    const double un = DIR == 0 ? u : (DIR == 1 ? v : w);
    if constexpr (FVS == FVS_SW || FVS == FVS_SWE) {
        double l1p, l1m, l2p, l2m, l3p, l3m;
        if constexpr (FVS == FVS_SWE) {
            lam_split_fvs<FVS>(un - c, P.sw_eps, l1p, l1m);
            lam_split_fvs<FVS>(un, P.sw_eps, l2p, l2m);
            lam_split_fvs<FVS>(un + c, P.sw_eps, l3p, l3m);
        } else {
            l1p = fmax(un - c, 0.0); l1m = fmin(un - c, 0.0);
            l2p = fmax(un, 0.0); l2m = fmin(un, 0.0);
            l3p = fmax(un + c, 0.0); l3m = fmin(un + c, 0.0);
        }
        const double H = (E + p) * irho;
        const double ga1 = rho * (0.5 / gamma), ga2 = 2.0 * (gamma - 1.0);
        const double S1p = l1p + l3p, S1m = l1m + l3m;
        const double S2p = (l3p - l1p) * c, S2m = (l3m - l1m) * c;
        const double Gp = fma(ga2, l2p, S1p), Gm = fma(ga2, l2m, S1m);   // S1 + ga2 l2
        const double g2p = ga2 * l2p, g2m = ga2 * l2m;
        Fp[0] = ga1 * Gp;
        Fp[1] = ga1 * (DIR == 0 ? fma(Gp, u, S2p) : Gp * u);
        Fp[2] = ga1 * (DIR == 1 ? fma(Gp, v, S2p) : Gp * v);
        Fp[3] = ga1 * (DIR == 2 ? fma(Gp, w, S2p) : Gp * w);
        Fp[4] = ga1 * fma(S1p, H, fma(g2p, V, S2p * un));
        Fm[0] = ga1 * Gm;
        Fm[1] = ga1 * (DIR == 0 ? fma(Gm, u, S2m) : Gm * u);
        Fm[2] = ga1 * (DIR == 1 ? fma(Gm, v, S2m) : Gm * v);
        Fm[3] = ga1 * (DIR == 2 ? fma(Gm, w, S2m) : Gm * w);
        Fm[4] = ga1 * fma(S1m, H, fma(g2m, V, S2m * un));
    } else {
        const double a = fabs(un) + c;
        const double F1 = rho * un;
        const double F5 = (E + p) * un;
        Fp[0] = 0.5 * (F1 + a * rho);
        Fp[1] = 0.5 * (F1 * u + (DIR == 0 ? p : 0.0) + a * rho * u);
        Fp[2] = 0.5 * (F1 * v + (DIR == 1 ? p : 0.0) + a * rho * v);
        Fp[3] = 0.5 * (F1 * w + (DIR == 2 ? p : 0.0) + a * rho * w);
        Fp[4] = 0.5 * (F5 + a * E);
        Fm[0] = 0.5 * (F1 - a * rho);
        Fm[1] = 0.5 * (F1 * u + (DIR == 0 ? p : 0.0) - a * rho * u);
        Fm[2] = 0.5 * (F1 * v + (DIR == 1 ? p : 0.0) - a * rho * v);
        Fm[3] = 0.5 * (F1 * w + (DIR == 2 ? p : 0.0) - a * rho * w);
        Fm[4] = 0.5 * (F5 - a * E);
    }
}

// x sweep, one block per full periodic x-row.  Each thread splits one cell in
// FP64 and differences it with the left neighbour's split values, taken from
// the next lower lane: wave w covers cells 63w - 1 .. 63w + 62 of the row's
// slots, its first lane only supplying the left neighbour of the second.  LDS
// holds the FP32 differences of the row and its ghost cells, and rel per
// face.  Launched with 64 * ceil((nx + S) / 63) threads.
template <int RECON, int FVS, int PER>
__global__ __launch_bounds__(1024) void km_sweep_xrow(DevParams P, MFields M)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    extern __shared__ float ldsf[];
    const int TX = P.nx;
    const int NSLOT = TX + S;                 // cells ng - MID - 1 .. ng + TX + S - MID - 2
    float *dl = ldsf;                         // [10][NSLOT] differences (slot 0 unused)
    float *rl = dl + 10 * NSLOT;              // [5][NF] rel per face
    constexpr int SH = PER ? 0 : 1;
    const int NF = TX + SH;                   // faces, the first right of cell ng - SH
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int j = P.ng + blockIdx.x;
    const int k = P.ng + blockIdx.y;
    const int row = (k * P.sy + j) * P.sx;
    const int i0 = P.ng - MID - 1;
    {
        const int q = 63 * (t >> 6) + lane - 1;
        const int qc = q < 0 ? 0 : (q < NSLOT ? q : NSLOT - 1);
        double rho, u, v, w, p, F[10];
        shadow_prims(P, M, row + i0 + qc, rho, u, v, w, p);
        split_fast<FVS, 0>(P, rho, u, v, w, p, F, F + 5);
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            const double L = __shfl_up(F[m], 1, 64);
            if (lane > 0 && q > 0 && q < NSLOT) dl[m * NSLOT + q] = (float)(F[m] - L);
        }
    }
    __syncthreads();
    float e[S];
    diff_coeffs<RECON == R_WENO5 ? R_MDCD : RECON>(P, e);   // unused for WENO5
    if (t < NF) {
        // face right of cell ng + c, c = t - SH: stencil cells ng + c - MID + s
        // = slots c + 1 + s (c = -1 reads slots 1 .. S - 1)
        const int c = t - SH;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            float D[S], Dm[S];
#pragma unroll
            for (int tt = 1; tt < S; ++tt) {
                D[tt] = dl[m * NSLOT + c + 1 + tt];
                Dm[tt] = -dl[(5 + m) * NSLOT + c + 1 + S - tt];
            }
            rl[m * NF + t] = rel_scale<RECON>(rel32<RECON>(e, D) + rel32<RECON>(e, Dm));
        }
    }
    __syncthreads();
    if (t < TX) {
        const int fr = t + SH;
        const int fl = PER ? (t == 0 ? TX - 1 : t - 1) : t;
        const int id = row + P.ng + t;
        // 1/dx as a float pair: 1/dx rounded once to float would scale
        // every inviscid term of a run by the same factor (1 - 4.0e-8 for
        // dx = 2 pi / 2^k), an error that accumulates over the steps
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const float dF = dl[m * NSLOT + t + MID + 1] + dl[(5 + m) * NSLOT + t + MID + 2] +
                             (rl[m * NF + fr] - rl[m * NF + fl]);
            M.A[m][id] = mul2(dF, P.id2[0]);
        }
    }
}

// y / z sweeps: one thread marches a column segment keeping, per component
// and side, the S - T0 FP32 differences of the stencil window and the FP64
// split value of the last cell (to difference the next one); accumulates
// into A.  The plus side's window is in registers; the minus side's is a
// per-lane ring in LDS (window position t of face n in slot (t - T0 + n) mod
// L), which rotates by moving its base rather than by register moves and
// leaves room for 4 waves per SIMD (the 8-point schemes spilled at 4 with
// both sides in registers and ran at 3: their sweeps are 32 % faster this
// way, WENO5's 3.5 %).  Each lane touches only its own column of the ring:
// no barrier.  Measured slower: both sides in the ring (12.8 KB per wave
// caps the CU at 3 waves per SIMD), forcing 5 waves (spills), and unrolling
// the march to rotate registers by renaming (165 VGPRs, 3 waves): the sweeps
// are VALU-bound and need the fourth wave.  The code's shape (the push / get
// helpers) is the one measured: the compiler's register use follows it
// (127 VGPRs for WENO5; a rewrite without them took 172, or spilled when
// held to 128).
template <int DIR, int RECON, int FVS>
__global__ __launch_bounds__(64) void km_sweep(DevParams P, MFields M, int seg_len)
{
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    constexpr int T0 = S == 2 ? 0 : 1;       // lowest window position that is read
    constexpr int L = S - T0;
    constexpr int R0 = 5;                     // components 5..9 (minus side) are in the ring
    __shared__ float ring[L][5][64];
    // WENO5: the minus side's smoothness first terms of the last two faces,
    // slot f & 1 for face f (weno5_u).  Not with the eps split, whose extra
    // registers would leave the carry at 146 VGPRs (3 waves per SIMD)
    constexpr bool CARRY = RECON == R_WENO5 && FVS != FVS_SWE;
    __shared__ float uring[CARRY ? 2 : 1][5][64];
    int bs, bt, bb;
    xcd_block3(bs, bt, bb);   // grid (nseg, lane tiles, b), as the launch
    const int lane = threadIdx.x;
    const int a = bt * blockDim.x + lane;
    if (a >= P.nx) return;
    const int b = bb;
    const int nd = DIR == 1 ? P.ny : P.nz;
    const int c0 = P.ng + bs * seg_len;
    const int c1 = min(c0 + seg_len, P.ng + nd);
    if (c0 >= c1) return;
    int base, st;
    if (DIR == 1) { base = ((P.ng + b) * P.sy) * P.sx + (P.ng + a); st = P.sx; }
    else { base = (P.ng + b) * P.sx + (P.ng + a); st = P.sx * P.sy; }
    float e[S];
    diff_coeffs<RECON == R_WENO5 ? R_MDCD : RECON>(P, e);
    float wp[5][S];                            // plus side: wp[m][t], positions T0..S-1
    int n0 = L - 1;                            // face index mod L; face -1 here
    const auto slot = [&n0](int t) { const int q = t - T0 + n0; return q >= L ? q - L : q; };
    // window value of component m (0..9) at position t
    const auto get = [&](int m, int t) -> float {
        if (m >= R0) return ring[slot(t)][m - R0][lane];
        return wp[m][t];
    };
    double prev[10];
    {
        double rho, u, v, w, p, Fp[5], Fm[5];
        shadow_prims(P, M, base + (c0 - 2 - MID + T0) * st, rho, u, v, w, p);
        split_fast<FVS, DIR>(P, rho, u, v, w, p, Fp, Fm);
#pragma unroll
        for (int m = 0; m < 5; ++m) { prev[m] = Fp[m]; prev[5 + m] = Fm[m]; }
    }
    // the newest position's values: the plus side after a shift, the minus
    // side into the slot of position S - 1
    const auto push = [&](const double *Fp, const double *Fm) {
#pragma unroll
        for (int t = T0; t < S - 1; ++t)
#pragma unroll
            for (int m = 0; m < 5; ++m) wp[m][t] = wp[m][t + 1];
        const int q = slot(S - 1);
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            const double Fv = m < 5 ? Fp[m] : Fm[m - 5];
            const float d = (float)(Fv - prev[m]);
            if (m >= R0) ring[q][m - R0][lane] = d;
            else wp[m][S - 1] = d;
            prev[m] = Fv;
        }
    };
    // positions T0 + 1 .. S - 1 as of face -1
#pragma unroll
    for (int t = T0 + 1; t < S; ++t) {
        double rho, u, v, w, p, Fp[5], Fm[5];
        shadow_prims(P, M, base + (c0 - 2 - MID + t) * st, rho, u, v, w, p);
        split_fast<FVS, DIR>(P, rho, u, v, w, p, Fp, Fm);
        {
            const int q = t - T0 - 1;   // slot(t) at face -1
#pragma unroll
            for (int m = 0; m < 5; ++m) wp[m][t] = (float)(Fp[m] - prev[m]);
#pragma unroll
            for (int m = 0; m < 10; ++m) {
                const double Fv = m < 5 ? Fp[m] : Fm[m - 5];
                if (m >= R0) ring[q][m - R0][lane] = (float)(Fv - prev[m]);
                prev[m] = Fv;
            }
        }
    }
    float nx_[5];
    {
        const int id = base + (c0 - 1 - MID + S - 1) * st;
#pragma unroll
        for (int q = 0; q < 5; ++q) nx_[q] = M.S[q][id];
    }
    float relprev[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    // WENO5: each face's three-cell stencils share their second difference
    // with the neighbouring faces.  After the push of face f the plus side's
    // stencils are positions (1,2,3), (2,3,4), (3,4,5) and only the last is
    // new; the minus side's are (5,4,3), (4,3,2), (3,2,1) read backwards and
    // only the first is new.  The plus side carries the older two terms in
    // up0 / up1, the minus side in uring.  Before the first face's push the
    // window holds positions 2..5 of face -1's numbering shifted down by one.
    float up0[5], up1[5];
    if constexpr (CARRY) {
        const int f0 = c0 - 1;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            up0[m] = weno5_u(wp[m][3] - wp[m][2]);
            up1[m] = weno5_u(wp[m][4] - wp[m][3]);
            uring[(f0 + 1) & 1][m][lane] = weno5_u(get(5 + m, 5) - get(5 + m, 4));   // s1 of face f0
            uring[f0 & 1][m][lane] = weno5_u(get(5 + m, 4) - get(5 + m, 3));         // s2 of face f0
        }
    }
    // the accumulator of the next cell written, loaded an iteration ahead
    // (a load at the point of the update left its latency exposed)
    float an[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) an[m] = M.A[m][base + c0 * st];
    for (int f = c0 - 1; f < c1; ++f) {
        n0 = n0 + 1 == L ? 0 : n0 + 1;
        float acur[5] = {an[0], an[1], an[2], an[3], an[4]};
        {
            float cur[5] = {nx_[0], nx_[1], nx_[2], nx_[3], nx_[4]};
            if (f + 1 < c1) {
                const int id = base + (f + 1 - MID + S - 1) * st;
#pragma unroll
                for (int q = 0; q < 5; ++q) nx_[q] = M.S[q][id];
                if (f >= c0) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) an[m] = M.A[m][base + (f + 1) * st];
                }
            }
            double Fp[5], Fm[5];
            split_fast<FVS, DIR>(P, P.rho_ref + (double)cur[0], (double)cur[1], (double)cur[2], (double)cur[3],
                                 P.p_ref + (double)cur[4], Fp, Fm);
            push(Fp, Fm);
        }
        float rel[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            float D[S], Dm[S];
#pragma unroll
            for (int t = 1; t < S; ++t) {
                D[t] = get(m, t);
                Dm[t] = -get(5 + m, S - t);
            }
            if constexpr (CARRY) {
                const float up2 = weno5_u(D[4] - D[3]);
                const float um0 = weno5_u(Dm[2] - Dm[1]);
                const float um1 = uring[(f + 1) & 1][m][lane];
                const float um2 = uring[f & 1][m][lane];
                uring[f & 1][m][lane] = um0;
                rel[m] = rel32_weno5_u(D, up0[m], up1[m], up2) + rel32_weno5_u(Dm, um0, um1, um2);
                up0[m] = up1[m];
                up1[m] = up2;
            } else {
                rel[m] = rel_scale<RECON>(rel32<RECON>(e, D) + rel32<RECON>(e, Dm));
            }
        }
        if (f >= c0) {
            const int idc = base + f * st;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const float dF = get(m, MID) + get(5 + m, MID + 1) + (rel[m] - relprev[m]);
                M.A[m][idc] = acur[m] + mul2(dF, P.id2[DIR]);   // 1/d as a pair, as in the x sweep
            }
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) relprev[m] = rel[m];
    }
}

// --------------------------------------------------------------- viscous

// mu(T) / mu0 in FP32 (get_mu of the FP64 path with its constant factor
// mu0 = mu, or mu (1 + S) for Sutherland, left out: the callers apply mu0
// as a float pair)
__device__ __forceinline__ float mu_shape32(const DevParams &P, float T)
{
    if (P.mu_law == 1) return 1.0f;
    if (P.mu_law == 2) return powf(T, (float)P.mu_power);
    return (T * sqrtf(T)) / ((T + P.suth.h) + P.suth.l);
}

// integer-weight central sums: the derivative is the sum times 1 / (K d),
// K = 840 (C8), 60 (C6), 12 (C4), with the scale as a float pair (rounded
// fractional weights such as 0.8f, or 1/d rounded once, would bias every
// derivative of a run by the same factor)
template <int VIS>
__device__ __forceinline__ float dsum32(const float *v)
{
    if constexpr (VIS == VIS_C8) {
        return 672.0f * (v[1] - v[-1]) - 168.0f * (v[2] - v[-2]) + 32.0f * (v[3] - v[-3]) - 3.0f * (v[4] - v[-4]);
    } else if constexpr (VIS == VIS_C6) {
        return 45.0f * (v[1] - v[-1]) - 9.0f * (v[2] - v[-2]) + (v[3] - v[-3]);
    } else {
        return 8.0f * (v[1] - v[-1]) - (v[2] - v[-2]);
    }
}

// sc: the axis' row of DevParams::dsc
template <int VIS>
__device__ __forceinline__ float diff32(const float *v, const F2 *sc)
{
    return mul2(dsum32<VIS>(v), sc[VIS == VIS_C8 ? 0 : (VIS == VIS_C6 ? 1 : 2)]);
}

// the closures of choose_scheme on non-periodic axes (as diff_closure) in
// FP32; periodic axes pass dlo = dhi = R
__device__ __forceinline__ float diff32_closure(const float *v, const F2 *sc, int dlo, int dhi)
{
    if (dlo == 0) return mul2(-3.0f * v[0] + 4.0f * v[1] - v[2], sc[3]);
    if (dhi == 0) return mul2(-3.0f * v[0] + 4.0f * v[-1] - v[-2], sc[3]);
    const int d = dlo < dhi ? dlo : dhi;
    if (d == 1) return mul2(v[1] - v[-1], sc[3]);
    if (d == 2) return diff32<VIS_C4>(v, sc);
    return diff32<VIS_C6>(v, sc);
}

template <int VIS>
__device__ __forceinline__ float diff32_bc(const float *v, const F2 *sc, int dlo, int dhi)
{
    constexpr int R = VisReach<VIS>::R;
    if (dlo >= R && dhi >= R) return diff32<VIS>(v, sc);
    return diff32_closure(v, sc, dlo, dhi);
}

// x and y derivatives of the four shadow quantities at cell id
template <int VIS>
__device__ __forceinline__ void gradxy32(const float *const src[4], int id, int sx, const F2 *scx, const F2 *scy,
                                         int dxl, int dxh, int dyl, int dyh, float g[8])
{
    constexpr int R = VisReach<VIS>::R;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float wx[2 * R + 1], wy[2 * R + 1];
#pragma unroll
        for (int o = -R; o <= R; ++o) {
            wx[o + R] = src[q][id + o];
            wy[o + R] = src[q][id + o * sx];
        }
        g[2 * q + 0] = diff32_bc<VIS>(wx + R, scx, dxl, dxh);
        g[2 * q + 1] = diff32_bc<VIS>(wy + R, scy, dyl, dyh);
    }
}

// the viscous flux of one cell from its 12 gradients g (u, v, w, T-1 along
// x, y, z) and its shadow u, v, w, T-1, into the 9 distinct flux arrays:
// tau = mu0 * f(T) * (2 S - 2/3 div u I), q = mk0 f(T) grad T, with the
// constants mu0, mk0 and 2/3 as float pairs
__device__ __forceinline__ void vflux32_store(const DevParams &P, const MFields &M, int id, const float g[12], float u,
                                              float v, float w, float t1)
{
    constexpr F2 TWO3 = {(float)(2.0 / 3.0), (float)(2.0 / 3.0 - (double)(float)(2.0 / 3.0))};
    const float f = mu_shape32(P, 1.0f + t1);
    const float d3 = mul2(g[0] + g[4] + g[8], TWO3);
    const float txx = mul2(f * (2.0f * g[0] - d3), P.mu0);
    const float tyy = mul2(f * (2.0f * g[4] - d3), P.mu0);
    const float tzz = mul2(f * (2.0f * g[8] - d3), P.mu0);
    const float txy = mul2(f * (g[1] + g[3]), P.mu0);
    const float txz = mul2(f * (g[2] + g[6]), P.mu0);
    const float tyz = mul2(f * (g[5] + g[7]), P.mu0);
    M.VF[0][id] = txx;
    M.VF[1][id] = txy;
    M.VF[2][id] = txz;
    M.VF[3][id] = u * txx + v * txy + w * txz + mul2(f * g[9], P.mk0);
    M.VF[5][id] = tyy;
    M.VF[6][id] = tyz;
    M.VF[7][id] = u * txy + v * tyy + w * tyz + mul2(f * g[10], P.mk0);
    M.VF[10][id] = tzz;
    M.VF[11][id] = u * txz + v * tyz + w * tzz + mul2(f * g[11], P.mk0);
}

// gradients + viscous flux in FP32 from the shadow (u, v, w, T-1),
// z-marching register window as k_gradflux.  The stress tensor is symmetric:
// VF[4], VF[8], VF[9] (txy, txz, tyz of the y and z fluxes) alias VF[1],
// VF[2], VF[6], so 9 arrays are written.
// A plane enters the z window R iterations before it is the centre plane;
// its x / y derivatives are taken in that same iteration and held in a
// register delay line, so each plane's neighbours are read while its lines
// are still in L2 (taking them R planes later re-fetched the plane from HBM:
// the XCD streams about 1 MB per plane step through its 4 MB L2).
template <int VIS, int BC>
__device__ __forceinline__ void gradflux_tile(const DevParams &P, const MFields &M, int seg_len, int bx, int by, int bz)
{
    constexpr int R = VisReach<VIS>::R;
    const int i = P.ng + bx * blockDim.x + threadIdx.x;
    const int j = P.ng + by * blockDim.y + threadIdx.y;
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int k0 = P.ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, P.ng + P.nz);
    const int sxy = P.sx * P.sy;
    const int col = j * P.sx + i;
    const float *const src[4] = {M.S[1], M.S[2], M.S[3], M.S[5]};
    int dxl = R, dxh = R, dyl = R, dyh = R, dzl = R, dzh = R;
    if constexpr (BC) {
        axis_dist(P.periodic[0], i, P.ng, P.nx, R, dxl, dxh);
        axis_dist(P.periodic[1], j, P.ng, P.ny, R, dyl, dyh);
    }
    float win[4][2 * R + 1];
    float gxy[R + 1][8];      // slot s: x / y derivatives of plane k + s
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o) win[q][o] = src[q][(k0 - R - 1 + o) * sxy + col];
#pragma unroll
    for (int s = 1; s <= R; ++s)
        if (k0 + s - 1 < k1) gradxy32<VIS>(src, (k0 + s - 1) * sxy + col, P.sx, P.dsc[0], P.dsc[1], dxl, dxh, dyl, dyh, gxy[s]);
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
        // the skip past the segment end makes the compiler serialise these
        // loads, but keeps the kernel at 126 VGPRs (4 waves per SIMD); the
        // unconditional form keeps them in flight at 160 (3 waves) and ran
        // 10 % slower
        if (k + R < k1) gradxy32<VIS>(src, id + R * sxy, P.sx, P.dsc[0], P.dsc[1], dxl, dxh, dyl, dyh, gxy[R]);
        float g[12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            g[3 * q + 0] = gxy[0][2 * q];
            g[3 * q + 1] = gxy[0][2 * q + 1];
            g[3 * q + 2] = diff32_bc<VIS>(&win[q][R], P.dsc[2], dzl, dzh);
        }
        vflux32_store(P, M, id, g, win[0][R], win[1][R], win[2][R], win[3][R]);
    }
}

// rows j and j + 1 of a periodic box per thread: each y-neighbour load
// serves both rows (2 R + 2 loads per quantity instead of 2 (2 R + 1)),
// and a 64 x 4 workgroup covers 8 rows, so its y halo is 2x its rows
// instead of 3x.  Interior cells only (periodic axes: no closures).
template <int VIS>
__device__ __forceinline__ void gradxy32_rows2(const float *const src[4], int id, int sx, const F2 *scx,
                                               const F2 *scy, float g0[8], float g1[8])
{
    constexpr int R = VisReach<VIS>::R;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float wx0[2 * R + 1], wx1[2 * R + 1], wy[2 * R + 2];
#pragma unroll
        for (int o = -R; o <= R; ++o) {
            wx0[o + R] = src[q][id + o];
            wx1[o + R] = src[q][id + sx + o];
        }
#pragma unroll
        for (int o = -R; o <= R + 1; ++o) wy[o + R] = src[q][id + o * sx];
        g0[2 * q + 0] = diff32<VIS>(wx0 + R, scx);
        g1[2 * q + 0] = diff32<VIS>(wx1 + R, scx);
        g0[2 * q + 1] = diff32<VIS>(wy + R, scy);
        g1[2 * q + 1] = diff32<VIS>(wy + R + 1, scy);
    }
}

template <int VIS>
__device__ __forceinline__ void gradflux_tile_rows2(const DevParams &P, const MFields &M, int seg_len, int bx, int by,
                                                    int bz)
{
    constexpr int R = VisReach<VIS>::R;
    const int i = P.ng + bx * blockDim.x + threadIdx.x;
    const int j = P.ng + 2 * (by * blockDim.y + threadIdx.y);   // ny is even: j + 1 is interior too
    if (i >= P.ng + P.nx || j >= P.ng + P.ny) return;
    const int k0 = P.ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, P.ng + P.nz);
    const int sx = P.sx, sxy = P.sx * P.sy;
    const int col = j * sx + i;
    const float *const src[4] = {M.S[1], M.S[2], M.S[3], M.S[5]};
    float win[2][4][2 * R + 1];
    float gxy[R + 1][2][8];   // slot s: x / y derivatives of plane k + s, rows j and j + 1
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int o = 1; o < 2 * R + 1; ++o) win[r][q][o] = src[q][(k0 - R - 1 + o) * sxy + col + r * sx];
#pragma unroll
    for (int s = 1; s <= R; ++s)
        if (k0 + s - 1 < k1) gradxy32_rows2<VIS>(src, (k0 + s - 1) * sxy + col, sx, P.dsc[0], P.dsc[1], gxy[s][0], gxy[s][1]);
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
#pragma unroll
        for (int s = 0; s < R; ++s)
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int q = 0; q < 8; ++q) gxy[s][r][q] = gxy[s + 1][r][q];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int o = 0; o < 2 * R; ++o) win[r][q][o] = win[r][q][o + 1];
                win[r][q][2 * R] = src[q][id + R * sxy + r * sx];
            }
        if (k + R < k1) gradxy32_rows2<VIS>(src, id + R * sxy, sx, P.dsc[0], P.dsc[1], gxy[R][0], gxy[R][1]);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            float g[12];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                g[3 * q + 0] = gxy[0][r][2 * q];
                g[3 * q + 1] = gxy[0][r][2 * q + 1];
                g[3 * q + 2] = diff32<VIS>(&win[r][q][R], P.dsc[2]);
            }
            vflux32_store(P, M, id + r * sx, g, win[r][0][R], win[r][1][R], win[r][2][R], win[r][3][R]);
        }
    }
}

// The two-row pass with each entering plane's x / y neighbourhood staged in
// LDS: a 64 x TY workgroup (64 x 2TY cells) writes its plane tile with an R-cell
// halo in x (own rows) and y (64 columns) into one of two LDS buffers, one
// barrier per plane, and takes the x / y derivatives from there.  The own
// values are the z window's entering plane; the halo costs 2R + 1 loads per
// thread and plane instead of the 2 (2R + 1) + 2R + 2 stencil loads per
// quantity of the cached form.  The next plane's values are loaded before the
// derivative and flux arithmetic of this one, which hides their latency.
// TY = 4 or 8 thread rows (taller tiles fetch fewer halo rows per row).
// Needs nx % 64 == 0 and ny % 2TY == 0 (no thread leaves before a barrier).
template <int VIS, int TY>
__global__ __launch_bounds__(64 * TY) void km_gradflux_lds(DevParams P, MFields M, int seg_len, int gx, int gy)
{
    constexpr int R = VisReach<VIS>::R;
    constexpr int TR = 2 * TY + 2 * R, TC = 64 + 2 * R;   // tile rows, columns
    __shared__ float tile[2][4][TR][TC];
    const int t = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int i0 = P.ng + bx * 64, j0 = P.ng + by * 2 * TY;
    const int i = i0 + tx, j = j0 + 2 * ty;
    const int k0 = P.ng + bz * seg_len;
    const int k1 = min(k0 + seg_len, P.ng + P.nz);
    const int sx = P.sx, sxy = P.sx * P.sy;
    const int col = j * sx + i;
    const float *const src[4] = {M.S[1], M.S[2], M.S[3], M.S[5]};
    // halo cells this thread stages: y-halo rows h = ty, ty + TY, ... < 2R
    // (tile row h or h + 2 TY), and one x-halo cell (tid < 16 R TY: quantity
    // tid / (4 R TY), own row (tid / 2R) % 2 TY, halo column tid % 2R)
    constexpr int NYH = (2 * R + TY - 1) / TY;
    int yh_row[NYH], yh_off[NYH];
#pragma unroll
    for (int n = 0; n < NYH; ++n) {
        const int h = ty + TY * n;
        const int tr = h < R ? h : h + 2 * TY;
        yh_row[n] = h < 2 * R ? tr : -1;
        yh_off[n] = (j0 - R + tr) * sx + i;
    }
    const bool xh_on = tid < 16 * R * TY;
    const int xh_q = xh_on ? tid / (4 * R * TY) : 0;
    const int xh_r = (tid / (2 * R)) % (2 * TY);
    const int xh_c = tid % (2 * R);
    const int xh_tc = xh_c < R ? xh_c : 64 + xh_c;
    const int xh_off = (j0 + xh_r) * sx + i0 - R + xh_tc;
    float win[2][4][2 * R + 1];
    float gxy[R + 1][2][8];
    float nyh[NYH][4], nxh = 0.0f, nown[2][4];
    const auto prefetch = [&](int p, bool halo) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) nown[r][q] = src[q][p * sxy + col + r * sx];
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
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int o = 1; o < 2 * R + 1; ++o) win[r][q][o] = src[q][(k0 - R - 1 + o) * sxy + col + r * sx];
#pragma unroll
    for (int s = 1; s <= R; ++s)
        if (k0 + s - 1 < k1) gradxy32_rows2<VIS>(src, (k0 + s - 1) * sxy + col, sx, P.dsc[0], P.dsc[1], gxy[s][0], gxy[s][1]);
    prefetch(k0 + R, k0 + R < k1);
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
        const int p = k + R;   // the entering plane
#pragma unroll
        for (int s = 0; s < R; ++s)
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int q = 0; q < 8; ++q) gxy[s][r][q] = gxy[s + 1][r][q];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int o = 0; o < 2 * R; ++o) win[r][q][o] = win[r][q][o + 1];
                win[r][q][2 * R] = nown[r][q];
            }
        const bool stage = p < k1;   // uniform over the workgroup
        if (stage) {
            float (*T)[TR][TC] = tile[p & 1];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) T[q][R + 2 * ty + r][R + tx] = nown[r][q];
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
            float (*T)[TR][TC] = tile[p & 1];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float wx0[2 * R + 1], wx1[2 * R + 1], wy[2 * R + 2];
#pragma unroll
                for (int o = -R; o <= R; ++o) {
                    wx0[o + R] = T[q][R + 2 * ty][R + tx + o];
                    wx1[o + R] = T[q][R + 2 * ty + 1][R + tx + o];
                }
#pragma unroll
                for (int o = -R; o <= R + 1; ++o) wy[o + R] = T[q][R + 2 * ty + o][R + tx];
                gxy[R][0][2 * q + 0] = diff32<VIS>(wx0 + R, P.dsc[0]);
                gxy[R][1][2 * q + 0] = diff32<VIS>(wx1 + R, P.dsc[0]);
                gxy[R][0][2 * q + 1] = diff32<VIS>(wy + R, P.dsc[1]);
                gxy[R][1][2 * q + 1] = diff32<VIS>(wy + R + 1, P.dsc[1]);
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            float g[12];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                g[3 * q + 0] = gxy[0][r][2 * q];
                g[3 * q + 1] = gxy[0][r][2 * q + 1];
                g[3 * q + 2] = diff32<VIS>(&win[r][q][R], P.dsc[2]);
            }
            vflux32_store(P, M, id + r * sx, g, win[r][0][R], win[r][1][R], win[r][2][R], win[r][3][R]);
        }
    }
}

// one workgroup per tile of blockDim.x x blockDim.y column segments (1D
// grid, XCD remap as xcd_block3)
template <int VIS, int BC>
__global__ __launch_bounds__(256) void km_gradflux(DevParams P, MFields M, int seg_len, int gx, int gy)
{
    const int t = xcd_logical(blockIdx.x, gridDim.x);
    gradflux_tile<VIS, BC>(P, M, seg_len, t % gx, (t / gx) % gy, t / (gx * gy));
}

template <int VIS>
__global__ __launch_bounds__(256) void km_gradflux_rows2(DevParams P, MFields M, int seg_len, int gx, int gy)
{
    const int t = xcd_logical(blockIdx.x, gridDim.x);
    gradflux_tile_rows2<VIS>(P, M, seg_len, t % gx, (t / gx) % gy, t / (gx * gy));
}

// the end of a cell's divergence pass: the RHS into `out` (MODE 0), or the
// RK3 stage update + C2P into U / dU / the shadow (MODE 1..3); with cfl_next
// the third stage also folds the new state's cell into the CFL maxima, from
// the shadow values it stores (the values the CFL pass would read)
// PF: the caller passes the cell's U (u0) and dU (du0) in every mode (else
// only the third stage does, and the first two load them here)
// sh: where given (MODE 1..3), also receives the six shadow values stored
template <int MODE, bool PF = false>
__device__ __forceinline__ void vdiv_finish(const DevParams &P, const MFields &M, int id, const double r[5],
                                            const double u0[5], const float du0[5], double dt, unsigned long long *flag,
                                            bool cfl_next, double &cx, double &cy, double &cz, float *sh = nullptr)
{
    if constexpr (MODE == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) M.out[m][id] = r[m];
    } else {
        // SSP-RK3 (range_kutta.cpp:102-147) in increment form on the
        // step's initial state U:
        //   U1 = U + d1,  d1 = dt L(U)
        //   U2 = 3/4 U + 1/4 (U1 + dt L(U1)) = U + d2,  d2 = (d1 + dt L(U1)) / 4
        //   U3 = 1/3 U + 2/3 (U2 + dt L(U2)) = U + 2/3 (d2 + dt L(U2))
        // U stays in FP64 and is read each stage but written only by the
        // third; the increments are held in FP32 (their rounding is ~6e-8
        // of one step's change).  The stage state is formed from the
        // rounded increment, so the shadow matches what the next stage
        // sees, and the third stage's increment is kept for the residual
        // monitor (exactly U3 - U up to U3's own rounding).
        const double gamma = P.gamma;
        double Un[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            float d;
            if constexpr (MODE == 1) d = (float)(dt * r[m]);
            else if constexpr (MODE == 2) d = (float)(0.25 * ((double)(PF ? du0[m] : M.dU[m][id]) + dt * r[m]));
            else d = (float)((2.0 / 3.0) * ((double)du0[m] + dt * r[m]));
            M.dU[m][id] = d;
            Un[m] = (MODE == 3 || PF ? u0[m] : M.U[m][id]) + (double)d;
        }
        double rn = Un[0];
        if (rn <= 0.0) rn = 1e-12;
        const double un = Un[1] / rn, vn = Un[2] / rn, wn = Un[3] / rn;
        const double kinetic = 0.5 * (un * un + vn * vn + wn * wn);
        const double pn = (gamma - 1.0) * rn * (Un[4] / rn - kinetic);
        const double Tn = pn / (rn * P.Rgas);
        if constexpr (MODE == 3) {
            M.U[0][id] = rn;
#pragma unroll
            for (int m = 1; m < 5; ++m) M.U[m][id] = Un[m];
        }
        const float s0 = (float)(rn - P.rho_ref), s1 = (float)un, s2 = (float)vn, s3 = (float)wn;
        const float s4 = (float)(pn - P.p_ref);
        M.S[0][id] = s0;
        M.S[1][id] = s1;
        M.S[2][id] = s2;
        M.S[3][id] = s3;
        M.S[4][id] = s4;
        M.S[5][id] = (float)(Tn - 1.0);
        if (sh) {
            sh[0] = s0; sh[1] = s1; sh[2] = s2; sh[3] = s3; sh[4] = s4;
            sh[5] = (float)(Tn - 1.0);
        }
        if constexpr (MODE == 3) {
            if (!(rn > 0.0) || !(pn > 0.0) || !(Tn > 0.0))
                atomicMin(flag + FLAG_BAD, nonphysical_key(flag, id));
            if (cfl_next)
                cfl_accum(P, P.rho_ref + (double)s0, (double)s1, (double)s2, (double)s3, P.p_ref + (double)s4,
                          cx, cy, cz);
        }
    }
}

// divergence of the FP32 viscous fluxes (periodic wrap on periodic axes,
// the closures near the ends of the others), the RHS in FP64, and the stage
// update + C2P in FP64 (MODE 1..3) or the RHS into `rhs_out` (MODE 0).
// Occupancy 2: with the register budget of 2 waves per SIMD the compiler
// keeps the ~80 loads of an iteration in flight; at its default target of 4
// it waits for them in small groups (0.74 against 0.81 ms per launch at
// 256^3).
// dt: the argument (dev_dt 0) or cfl * the CFL minimum in the flag words;
// dev_dt 2 (the grid tiles the interior exactly, whole waves): the third
// stage also reduces the new state's CFL minimum into the next step's slot,
// which replaces the CFL pass at the start of the next step.
template <int VIS, int MODE, int BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void km_vdiv_z(DevParams P, MFields M, double dt, int seg_len,
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
        oxc[o + R] = j * P.sx + wrapi_m(i + o, ng, P.nx);
        oyc[o + R] = wrapi_m(j + o, ng, P.ny) * P.sx + i;
    }
    float win[4][2 * R + 1];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o)
            win[m][o] = M.VF[8 + m][wrapi_m(k0 - R - 1 + o, ng, P.nz) * sxy + col];
    int dxl = R, dxh = R, dyl = R, dyh = R, dzl = R, dzh = R;
    if constexpr (BC) {
        axis_dist(P.periodic[0], i, ng, P.nx, R, dxl, dxh);
        axis_dist(P.periodic[1], j, ng, P.ny, R, dyl, dyh);
    }
    const bool cfl_next = MODE == 3 && dev_dt == 2;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
        const int kn = wrapi_m(k + R, ng, P.nz) * sxy + col;
        if constexpr (BC) axis_dist(P.periodic[2], k, ng, P.nz, R, dzl, dzh);
        // every load of the cell's divergence first, then the arithmetic:
        // the scheduling barrier keeps the loads one burst
        float wx[4][2 * R + 1], wy[4][2 * R + 1], an[5];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[m][o] = win[m][o + 1];
            win[m][2 * R] = M.VF[8 + m][kn];
#pragma unroll
            for (int o = 0; o < 2 * R + 1; ++o) {
                wx[m][o] = M.VF[m][k * sxy + oxc[o]];
                wy[m][o] = M.VF[4 + m][k * sxy + oyc[o]];
            }
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) an[m] = M.A[m][id];
        // the third stage takes U and dU into the burst as well (u0, du0);
        // the first two read them in vdiv_finish (measured both ways: 0.90 against 1.10 ms
        // for the third stage at 256^3, 0.60 / 0.66 against 0.70 / 0.73 ms
        // for the first two)
        double u0[5];
        float du0[5];
        if constexpr (MODE == 3) {
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                u0[m] = M.U[m][id];
                du0[m] = M.dU[m][id];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        double r[5];
        r[0] = -(double)an[0];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const float vd = diff32_bc<VIS>(wx[m] + R, P.dsc[0], dxl, dxh) + diff32_bc<VIS>(wy[m] + R, P.dsc[1], dyl, dyh) +
                             diff32_bc<VIS>(&win[m][R], P.dsc[2], dzl, dzh);
            r[m + 1] = (double)vd - (double)an[m + 1];
        }
        vdiv_finish<MODE>(P, M, id, r, u0, du0, dt, flag, cfl_next, cx, cy, cz);
    }
    if (cfl_next) {
        const double cm = wave_min(cfl_bound(P, cx, cy, cz));
        if (__lane_id() == 0) atomicMin(flag + FLAG_CFL_NEXT, (unsigned long long)__double_as_longlong(cm));
    }
}

// The LDS-staged stage pass on a periodic box, one z segment of a 64 x TY
// tile per workgroup (TY = 4 or 8) (km_vdiv_lds; with ZS, km_vdiv_zs).
//
// km_vdiv_z's arithmetic with each plane's x / y stencil fluxes staged in
// LDS: the workgroup writes its plane's x fluxes with an R-cell x halo and
// its y fluxes with an R-row y halo (periodic images of interior cells: the
// flux arrays hold no ghosts) into one of two buffers, one barrier per
// plane, and takes both derivatives from there.  Everything the next plane
// needs (own fluxes, halos, the entering z-flux plane, A, U, dU) is loaded
// before this plane's arithmetic, so the loads of a plane overlap the
// arithmetic of the previous one: about 35 loads per cell instead of 91.
// The same arithmetic in the same order as km_vdiv_z<VIS, MODE, 0>: equal
// bit for bit.  Needs nx % 64 == 0 and ny % TY == 0 (no thread leaves before
// a barrier).
// IMG (MODE 1..3): the stage also writes the new shadow's periodic ghost
// images (the face layers of the fill pass), so that pass can be skipped.
// ZS (MODE 1..3, with IMG): the same stage's z sweep folded in.  Each thread
// also marches its column's z sweep (the state and arithmetic of
// km_sweep<2>, its minus-side ring 64 TY lanes wide), which at plane k yields
// that cell's z flux difference, added to the accumulator of the x and y
// sweeps as km_sweep<2> adds it.  The accumulator's z update is then
// neither stored nor read back (40 B per cell), and the sweep's arithmetic
// runs beside the stage's loads.  The z sweep of a segment reads the input
// shadow M.S across the segment's ends, which other workgroups' stages
// overwrite: the new shadow goes to So (a second buffer, the next stage's
// input).  Equal bit for bit to km_sweep<2> followed by the stage.
template <int VIS, int MODE, bool IMG, bool ZS, int RECON, int FVS, int TY>
__device__ __forceinline__ void vdiv_lds_pass(const DevParams &P, const MFields &M, const Shadow &So, double dt,
                                              int seg_len, int gx, int gy, unsigned long long *flag, int dev_dt)
{
    constexpr int R = VisReach<VIS>::R;
    constexpr int TC = 64 + 2 * R, TYR = TY + 2 * R;
    __shared__ float fxt[2][4][TY][TC];    // x fluxes: tile rows, R-cell x halo
    __shared__ float fyt[2][4][TYR][64];   // y fluxes: tile columns, R-row y halo
    // the z sweep (km_sweep<2>): per-lane ring of the minus side's window
    constexpr int S = Stencil<RECON>::S;
    constexpr int MID = (S - 1) / 2;
    constexpr int T0 = S == 2 ? 0 : 1;
    constexpr int L = S - T0;
    constexpr int R0 = 5;
    constexpr bool CARRY = ZS && RECON == R_WENO5 && FVS != FVS_SWE;
    constexpr int NL = ZS ? 64 * TY : 1;
    __shared__ float ring[ZS ? L : 1][5][NL];
    __shared__ float uring[CARRY ? 2 : 1][5][NL];
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
    MFields Mo = M;   // the stage's outputs; with ZS the shadow into So
    if constexpr (ZS) {
#pragma unroll
        for (int q = 0; q < 6; ++q) Mo.S[q] = So.S[q];
    }

    // ---- z sweep state (km_sweep<2> with base = col, st = sxy)
    const int lane = ZS ? tid : 0;
    const int st = sxy;
    float e[S];
    float wp[5][S];
    int n0 = L - 1;
    double prev[10];
    float nx_[5];
    float relprev[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float up0[5], up1[5];
    const auto slot = [&n0](int tt) { const int q = tt - T0 + n0; return q >= L ? q - L : q; };
    const auto get = [&](int m, int tt) -> float {
        if (m >= R0) return ring[slot(tt)][m - R0][lane];
        return wp[m][tt];
    };
    const auto push = [&](const double *Fp, const double *Fm) {
#pragma unroll
        for (int tt = T0; tt < S - 1; ++tt)
#pragma unroll
            for (int m = 0; m < 5; ++m) wp[m][tt] = wp[m][tt + 1];
        const int q = slot(S - 1);
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            const double Fv = m < 5 ? Fp[m] : Fm[m - 5];
            const float d = (float)(Fv - prev[m]);
            if (m >= R0) ring[q][m - R0][lane] = d;
            else wp[m][S - 1] = d;
            prev[m] = Fv;
        }
    };
    // one step of the z sweep: face f; for f >= k0 the cell's flux difference
    // (times 1/dz as the pair of km_sweep) added to a[]
    const auto zstep = [&](int f, float *a) {
        n0 = n0 + 1 == L ? 0 : n0 + 1;
        {
            float cur[5] = {nx_[0], nx_[1], nx_[2], nx_[3], nx_[4]};
            if (f + 1 < k1) {
                const int id = col + (f + 1 - MID + S - 1) * st;
#pragma unroll
                for (int q = 0; q < 5; ++q) nx_[q] = M.S[q][id];
            }
            double Fp[5], Fm[5];
            split_fast<FVS, 2>(P, P.rho_ref + (double)cur[0], (double)cur[1], (double)cur[2], (double)cur[3],
                               P.p_ref + (double)cur[4], Fp, Fm);
            push(Fp, Fm);
        }
        float rel[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            float D[S], Dm[S];
#pragma unroll
            for (int tt = 1; tt < S; ++tt) {
                D[tt] = get(m, tt);
                Dm[tt] = -get(5 + m, S - tt);
            }
            if constexpr (CARRY) {
                const float up2 = weno5_u(D[4] - D[3]);
                const float um0 = weno5_u(Dm[2] - Dm[1]);
                const float um1 = uring[(f + 1) & 1][m][lane];
                const float um2 = uring[f & 1][m][lane];
                uring[f & 1][m][lane] = um0;
                rel[m] = rel32_weno5_u(D, up0[m], up1[m], up2) + rel32_weno5_u(Dm, um0, um1, um2);
                up0[m] = up1[m];
                up1[m] = up2;
            } else {
                rel[m] = rel_scale<RECON>(rel32<RECON>(e, D) + rel32<RECON>(e, Dm));
            }
        }
        if (f >= k0) {
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const float dF = get(m, MID) + get(5 + m, MID + 1) + (rel[m] - relprev[m]);
                a[m] = a[m] + mul2(dF, P.id2[2]);
            }
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) relprev[m] = rel[m];
    };
    if constexpr (ZS) {
        diff_coeffs<RECON == R_WENO5 ? R_MDCD : RECON>(P, e);
        {
            double rho, u, v, w, p, Fp[5], Fm[5];
            shadow_prims(P, M, col + (k0 - 2 - MID + T0) * st, rho, u, v, w, p);
            split_fast<FVS, 2>(P, rho, u, v, w, p, Fp, Fm);
#pragma unroll
            for (int m = 0; m < 5; ++m) { prev[m] = Fp[m]; prev[5 + m] = Fm[m]; }
        }
#pragma unroll
        for (int tt = T0 + 1; tt < S; ++tt) {
            double rho, u, v, w, p, Fp[5], Fm[5];
            shadow_prims(P, M, col + (k0 - 2 - MID + tt) * st, rho, u, v, w, p);
            split_fast<FVS, 2>(P, rho, u, v, w, p, Fp, Fm);
            const int q = tt - T0 - 1;
#pragma unroll
            for (int m = 0; m < 5; ++m) wp[m][tt] = (float)(Fp[m] - prev[m]);
#pragma unroll
            for (int m = 0; m < 10; ++m) {
                const double Fv = m < 5 ? Fp[m] : Fm[m - 5];
                if (m >= R0) ring[q][m - R0][lane] = (float)(Fv - prev[m]);
                prev[m] = Fv;
            }
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) nx_[q] = M.S[q][col + (k0 - 1 - MID + S - 1) * st];
        if constexpr (CARRY) {
            const int f0 = k0 - 1;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                up0[m] = weno5_u(wp[m][3] - wp[m][2]);
                up1[m] = weno5_u(wp[m][4] - wp[m][3]);
                uring[(f0 + 1) & 1][m][lane] = weno5_u(get(5 + m, 5) - get(5 + m, 4));   // s1 of face f0
                uring[f0 & 1][m][lane] = weno5_u(get(5 + m, 4) - get(5 + m, 3));         // s2 of face f0
            }
        }
        float dummy[5];
        zstep(k0 - 1, dummy);
    }

    // ---- the stage pass
    // the x-halo cell of this thread (tid < 8 R TY): flux tid / (2R TY), tile
    // row (tid / 2R) % TY, halo column tid % 2R
    const bool xh_on = tid < 8 * R * TY;
    const int xh_q = xh_on ? tid / (2 * R * TY) : 0;
    const int xh_r = (tid / (2 * R)) % TY;
    const int xh_c = tid % (2 * R);
    const int xh_tc = xh_c < R ? xh_c : 64 + xh_c;
    const int xh_off = (j0 + xh_r) * sx + wrapi_m(i0 - R + xh_tc, ng, P.nx);
    // the y-halo rows of this thread: h = ty, ty + TY, ... < 2R, tile row h
    // (below the tile) or h + TY (above it)
    constexpr int NYH = (2 * R + TY - 1) / TY;
    int yh_row[NYH], yh_off[NYH];
#pragma unroll
    for (int n = 0; n < NYH; ++n) {
        const int h = ty + TY * n;
        const int tr = h < R ? h : h + TY;
        yh_row[n] = h < 2 * R ? tr : -1;
        yh_off[n] = wrapi_m(j0 - R + tr, ng, P.ny) * sx + i;
    }
    // the y fluxes are VF[4..7] with VF[4] = VF[1]: comps 1..3 from VF[5..7]
    float pfx[4], pfy[3], pxh = 0.0f, pyh[NYH][4], pz[4], pa[5], pdu[5];
    double pu[5];
    const auto prefetch = [&](int kp) {
        const int pl = kp * sxy;
#pragma unroll
        for (int q = 0; q < 4; ++q) pfx[q] = M.VF[q][pl + col];
#pragma unroll
        for (int q = 0; q < 3; ++q) pfy[q] = M.VF[5 + q][pl + col];
        if (xh_on) pxh = M.VF[xh_q][pl + xh_off];
#pragma unroll
        for (int n = 0; n < NYH; ++n)
            if (yh_row[n] >= 0) {
                pyh[n][0] = M.VF[1][pl + yh_off[n]];
#pragma unroll
                for (int q = 1; q < 4; ++q) pyh[n][q] = M.VF[4 + q][pl + yh_off[n]];
            }
        const int zn = wrapi_m(kp + R, ng, P.nz) * sxy + col;
#pragma unroll
        for (int m = 0; m < 4; ++m) pz[m] = M.VF[8 + m][zn];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            pa[m] = M.A[m][pl + col];
            if constexpr (MODE != 0) pu[m] = M.U[m][pl + col];
            if constexpr (MODE >= 2) pdu[m] = M.dU[m][pl + col];
        }
    };
    float win[4][2 * R + 1];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int o = 1; o < 2 * R + 1; ++o)
            win[m][o] = M.VF[8 + m][wrapi_m(k0 - R - 1 + o, ng, P.nz) * sxy + col];
    prefetch(k0);
    const bool cfl_next = MODE == 3 && dev_dt == 2;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int id = k * sxy + col;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int o = 0; o < 2 * R; ++o) win[m][o] = win[m][o + 1];
            win[m][2 * R] = pz[m];
        }
        float (*FX)[TY][TC] = fxt[k & 1];
        float (*FY)[TYR][64] = fyt[k & 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) FX[q][ty][R + tx] = pfx[q];
        if (xh_on) FX[xh_q][xh_r][xh_tc] = pxh;
        FY[0][R + ty][tx] = pfx[1];
#pragma unroll
        for (int q = 1; q < 4; ++q) FY[q][R + ty][tx] = pfy[q - 1];
#pragma unroll
        for (int n = 0; n < NYH; ++n)
            if (yh_row[n] >= 0)
#pragma unroll
                for (int q = 0; q < 4; ++q) FY[q][yh_row[n]][tx] = pyh[n][q];
        float an[5], du0[5];
        double u0[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            an[m] = pa[m];
            u0[m] = MODE != 0 ? pu[m] : 0.0;
            du0[m] = MODE >= 2 ? pdu[m] : 0.0f;
        }
        __syncthreads();
        if (k + 1 < k1) prefetch(k + 1);
        if constexpr (ZS) zstep(k, an);
        double r[5];
        r[0] = -(double)an[0];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            float wx[2 * R + 1], wy[2 * R + 1];
#pragma unroll
            for (int o = -R; o <= R; ++o) {
                wx[o + R] = FX[m][ty][R + tx + o];
                wy[o + R] = FY[m][R + ty + o][tx];
            }
            const float vd = diff32<VIS>(wx + R, P.dsc[0]) + diff32<VIS>(wy + R, P.dsc[1]) +
                             diff32<VIS>(&win[m][R], P.dsc[2]);
            r[m + 1] = (double)vd - (double)an[m + 1];
        }
        if constexpr (IMG && MODE != 0) {
            float sh[6];
            vdiv_finish<MODE, true>(P, Mo, id, r, u0, du0, dt, flag, cfl_next, cx, cy, cz, sh);
            // the cell's images in the ghost layers of each axis: i < 2 ng is
            // the right ghost i + n, i >= n the left ghost i - n
            const int im[3] = {i, j, k};
            const int nn[3] = {P.nx, P.ny, P.nz};
            const int str[3] = {1, sx, sxy};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (im[a] < 2 * ng) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) Mo.S[q][id + nn[a] * str[a]] = sh[q];
                }
                if (im[a] >= nn[a]) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) Mo.S[q][id - nn[a] * str[a]] = sh[q];
                }
            }
        } else {
            vdiv_finish<MODE, true>(P, Mo, id, r, u0, du0, dt, flag, cfl_next, cx, cy, cz);
        }
    }
    if (cfl_next) {
        const double cm = wave_min(cfl_bound(P, cx, cy, cz));
        if (__lane_id() == 0) atomicMin(flag + FLAG_CFL_NEXT, (unsigned long long)__double_as_longlong(cm));
    }
}

// Occupancy 2 for both: at the default target of 4 waves the compiler waits
// for the ~35 loads of a plane in small groups
template <int VIS, int MODE, bool IMG, int TY>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(2, 2))) void km_vdiv_lds(
    DevParams P, MFields M, double dt, int seg_len, int gx, int gy, unsigned long long *flag, int dev_dt)
{
    vdiv_lds_pass<VIS, MODE, IMG, false, R_WENO5, FVS_SW, TY>(P, M, Shadow{}, dt, seg_len, gx, gy, flag, dev_dt);
}

template <int VIS, int MODE, int RECON, int FVS, int TY>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(2, 2))) void km_vdiv_zs(
    DevParams P, MFields M, Shadow So, double dt, int seg_len, int gx, int gy, unsigned long long *flag, int dev_dt)
{
    static_assert(MODE != 0, "the folded z sweep belongs to a stage");
    vdiv_lds_pass<VIS, MODE, true, true, RECON, FVS, TY>(P, M, So, dt, seg_len, gx, gy, flag, dev_dt);
}

// ------------------------------------------------------ state conversion

// primitives (src[0..4], FP64, every cell) -> conserved U (FP64, P2C of
// field_structures.h:430-451) and the FP32 deviation shadow
__global__ void km_from_prims(DevParams P, MFields M, const double *const *src, int ncell, int write_u)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell) return;
    const double gamma = P.gamma;
    double rr = src[0][id];
    if (rr <= 0.0) rr = 1e-12;
    const double u = src[1][id], v = src[2][id], w = src[3][id], p = src[4][id];
    if (write_u) {
        M.U[0][id] = rr;
        M.U[1][id] = rr * u;
        M.U[2][id] = rr * v;
        M.U[3][id] = rr * w;
        const double kinetic = 0.5 * (u * u + v * v + w * w);
        M.U[4][id] = rr * (p / ((gamma - 1.0) * rr) + kinetic);
    }
    const double T = p / (rr * P.Rgas);
    M.S[0][id] = (float)(rr - P.rho_ref);
    M.S[1][id] = (float)u;
    M.S[2][id] = (float)v;
    M.S[3][id] = (float)w;
    M.S[4][id] = (float)(p - P.p_ref);
    M.S[5][id] = (float)(T - 1.0);
}

// conserved U (FP64) -> primitives rho,u,v,w,p,T into dst[0..5] (C2P)
__global__ void km_to_prims(DevParams P, MFields M, double *const *dst, int ncell)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell) return;
    double rr = M.U[0][id];
    if (rr <= 0.0) rr = 1e-12;
    const double u = M.U[1][id] / rr, v = M.U[2][id] / rr, w = M.U[3][id] / rr;
    const double kinetic = 0.5 * (u * u + v * v + w * w);
    const double p = (P.gamma - 1.0) * rr * (M.U[4][id] / rr - kinetic);
    dst[0][id] = rr; dst[1][id] = u; dst[2][id] = v; dst[3][id] = w; dst[4][id] = p;
    dst[5][id] = p / (rr * P.Rgas);
}

// FP64 primitives (rho, u, v, w, p, T) of the FP64 state U + dU (add_du) or
// U at the interior cells: C2P of the state whose FP32 rounding is the
// shadow (the stage kernel's C2P).  The FP64 derivative kernels take them.
__global__ void km_state_prims64(DevParams P, MFields M, Prims64 out, int add_du)
{
    const int i = P.ng + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = P.ng + blockIdx.y;
    const int k = P.ng + blockIdx.z;
    if (i >= P.ng + P.nx) return;
    const int id = (k * P.sy + j) * P.sx + i;
    double Un[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) Un[m] = M.U[m][id] + (add_du ? (double)M.dU[m][id] : 0.0);
    double rn = Un[0];
    if (rn <= 0.0) rn = 1e-12;
    const double un = Un[1] / rn, vn = Un[2] / rn, wn = Un[3] / rn;
    const double kinetic = 0.5 * (un * un + vn * vn + wn * wn);
    const double pn = (P.gamma - 1.0) * rn * (Un[4] / rn - kinetic);
    out.p[0][id] = rn;
    out.p[1][id] = un;
    out.p[2][id] = vn;
    out.p[3][id] = wn;
    out.p[4][id] = pn;
    out.p[5][id] = pn / (rn * P.Rgas);
}

// periodic ghost fill of the six FP32 shadow arrays (faces only)
__global__ void km_fill_periodic(DevParams P, MFields M)
{
    const int axis = blockIdx.y;
    if (!P.periodic[axis]) return;
    int dst, src;
    if (!periodic_ghost_pair(P, axis, blockIdx.x * blockDim.x + threadIdx.x, dst, src)) return;
#pragma unroll
    for (int m = 0; m < 6; ++m) M.S[m][dst] = M.S[m][src];
}

// the boundary.cpp:58-207 ghost fills of k_fill_bc on the FP32 shadow
// (deviations: the wall/symmetry mirror negates the normal velocity, the
// inflow state is converted as km_from_prims converts a cell)
__global__ void km_fill_bc(DevParams P, MFields M, int face, int kind)
{
    const int axis = face >> 1;
    const int na = axis == 0 ? P.sy : P.sx;
    const int nb = axis == 2 ? P.sy : P.sz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.ng * na * nb) return;
    int dst, src, normal;
    bc_cell(P, face, kind, t, dst, src, normal);
    float v[6];
    if (kind == BCK_INFLOW) {
        double st[5];
        inflow_state(st);
        const double T = st[4] / (st[0] * P.Rgas);
        v[0] = (float)(st[0] - P.rho_ref);
        v[1] = (float)st[1];
        v[2] = (float)st[2];
        v[3] = (float)st[3];
        v[4] = (float)(st[4] - P.p_ref);
        v[5] = (float)(T - 1.0);
    } else {
#pragma unroll
        for (int m = 0; m < 6; ++m) v[m] = M.S[m][src];
        if (kind != BCK_OUTFLOW) v[normal] = -v[normal];
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) M.S[m][dst] = v[m];
}

// Characteristic-wise reconstruction (du_inviscous_character,
// InviscousFlux_new.cpp:165-437) in mixed mode: the shadow widened to FP64
// primitives (every cell) for the FP64 face kernels, then the face
// differences into the FP32 accumulator.
__global__ void km_shadow_to_prims64(DevParams P, MFields M, Prims64 out, int ncell)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncell) return;
    double rho, u, v, w, p;
    shadow_prims(P, M, id, rho, u, v, w, p);
    out.p[0][id] = rho;
    out.p[1][id] = u;
    out.p[2][id] = v;
    out.p[3][id] = w;
    out.p[4][id] = p;
    out.p[5][id] = 1.0 + (double)M.S[5][id];
}

// (F(f) - F(f-1)) / d per direction, summed as k_face_diff does, into M.A
__global__ void km_face_diff(DevParams P, MFields M, FaceArrays ff)
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
        const double fx = ff.f[m][fxr] - ff.f[m][fxr - 1];
        const double fy = ff.f[5 + m][fyt] - ff.f[5 + m][fyt - P.sx];
        const double fz = ff.f[10 + m][fzf] - ff.f[10 + m][fzf - P.sx * P.sy];
        M.A[m][id] = (float)(fx * P.idx + fy * P.idy + fz * P.idz);
    }
}

// CFL minimum (timestep.cpp:5-33) from the shadow
__global__ __launch_bounds__(256) void km_cfl(DevParams P, MFields M, unsigned long long *out)
{
    __shared__ double red[4];
    const int nxy = P.nx * P.ny;
    const int total = nxy * P.nz;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int k = t / nxy;
        const int r = t - k * nxy;
        const int j = r / P.nx;
        const int i = r - j * P.nx;
        const int id = ((P.ng + k) * P.sy + (P.ng + j)) * P.sx + (P.ng + i);
        double rho, u, v, w, p;
        shadow_prims(P, M, id, rho, u, v, w, p);
        cfl_accum(P, rho, u, v, w, p, cx, cy, cz);
    }
    const double m = wave_min(cfl_bound(P, cx, cy, cz));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
        atomicMin(out, (unsigned long long)__double_as_longlong(b));
    }
}

// ------------------------------------------------------------- launches

hipError_t launch_m_from_prims(hipStream_t s, const DevParams &P, const MFields &M, const double *const *src,
                               int write_u)
{
    const int n = P.sx * P.sy * P.sz;
    hipLaunchKernelGGL(km_from_prims, dim3((n + 255) / 256), dim3(256), 0, s, P, M, src, n, write_u);
    return hipGetLastError();
}

hipError_t launch_m_to_prims(hipStream_t s, const DevParams &P, const MFields &M, double *const *dst)
{
    const int n = P.sx * P.sy * P.sz;
    hipLaunchKernelGGL(km_to_prims, dim3((n + 255) / 256), dim3(256), 0, s, P, M, dst, n);
    return hipGetLastError();
}

hipError_t launch_m_fill_ghosts(hipStream_t s, const DevParams &P, const MFields &M)
{
    int m = P.ny * P.nz;
    if (P.nx * P.nz > m) m = P.nx * P.nz;
    if (P.nx * P.ny > m) m = P.nx * P.ny;
    const int total = 2 * P.ng * m;
    hipLaunchKernelGGL(km_fill_periodic, dim3((total + 255) / 256, 3), dim3(256), 0, s, P, M);
    for (int face = 0; face < 6; ++face) {
        const int axis = face >> 1;
        if (P.periodic[axis] || P.bc[face] == BCK_PERIODIC) continue;
        const int n = P.ng * (axis == 0 ? P.sy : P.sx) * (axis == 2 ? P.sy : P.sz);
        hipLaunchKernelGGL(km_fill_bc, dim3((n + 255) / 256), dim3(256), 0, s, P, M, face, P.bc[face]);
    }
    return hipGetLastError();
}

hipError_t launch_m_state_prims64(hipStream_t s, const DevParams &P, const MFields &M, const Prims64 &out, int add_du)
{
    hipLaunchKernelGGL(km_state_prims64, dim3((P.nx + 255) / 256, P.ny, P.nz), dim3(256), 0, s, P, M, out, add_du);
    return hipGetLastError();
}

hipError_t launch_m_shadow_to_prims64(hipStream_t s, const DevParams &P, const MFields &M, const Prims64 &out)
{
    const int n = P.sx * P.sy * P.sz;
    hipLaunchKernelGGL(km_shadow_to_prims64, dim3((n + 255) / 256), dim3(256), 0, s, P, M, out, n);
    return hipGetLastError();
}

hipError_t launch_m_face_diff(hipStream_t s, const DevParams &P, const MFields &M, const FaceArrays &ff)
{
    hipLaunchKernelGGL(km_face_diff, dim3((P.nx + 63) / 64, (P.ny + 3) / 4, P.nz), dim3(64, 4, 1), 0, s, P, M, ff);
    return hipGetLastError();
}

hipError_t launch_m_cfl(hipStream_t s, const DevParams &P, const MFields &M, unsigned long long *out)
{
    const long total = (long)P.nx * P.ny * P.nz;
    long blocks = (total + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(km_cfl, dim3((unsigned)blocks), dim3(256), 0, s, P, M, out);
    return hipGetLastError();
}

using MXrowFn = void (*)(DevParams, MFields);
using MSweepFn = void (*)(DevParams, MFields, int);

template <int FVS, int PER>
static MXrowFn m_pick_xrow(int recon)
{
    switch (recon) {
    case R_WENO5: return km_sweep_xrow<R_WENO5, FVS, PER>;
    case R_LINEAR: return km_sweep_xrow<R_LINEAR, FVS, PER>;
    case R_MDCD: return km_sweep_xrow<R_MDCD, FVS, PER>;
    case R_MDCD_HYBRID: return km_sweep_xrow64<R_MDCD_HYBRID, FVS, PER>;
    case R_UD7: return km_sweep_xrow<R_UD7, FVS, PER>;
    case R_UD7L: return km_sweep_xrow<R_UD7L, FVS, PER>;
    default: return nullptr;
    }
}

template <int DIR, int FVS>
static MSweepFn m_pick_sweep(int recon)
{
    switch (recon) {
    case R_WENO5: return km_sweep<DIR, R_WENO5, FVS>;
    case R_LINEAR: return km_sweep<DIR, R_LINEAR, FVS>;
    case R_MDCD: return km_sweep<DIR, R_MDCD, FVS>;
    case R_MDCD_HYBRID: return km_sweep64<DIR, R_MDCD_HYBRID, FVS>;
    case R_UD7: return km_sweep<DIR, R_UD7, FVS>;
    case R_UD7L: return km_sweep<DIR, R_UD7L, FVS>;
    default: return nullptr;
    }
}

// x-row launch shape of the mixed path: threads and LDS bytes
static void m_xrow_shape(const DevParams &P, int &threads, size_t &lds)
{
    const int S = P.recon == R_WENO5 || P.recon == R_MDCD || P.recon == R_MDCD_HYBRID ? 6
                  : P.recon == R_LINEAR ? 2 : 8;
    const int sh = P.periodic[0] ? 0 : 1;
    if (P.recon == R_MDCD_HYBRID) {
        threads = ((P.nx + 63) / 64) * 64;
        lds = (size_t)(10 * (P.nx + S - 1 + sh) + 5 * (P.nx + sh)) * sizeof(double);
    } else {
        threads = 64 * ((P.nx + S + 62) / 63);
        lds = (size_t)(10 * (P.nx + S) + 5 * (P.nx + sh)) * sizeof(float);
    }
}

// the mixed x sweep holds whole rows (1024 threads, the LDS of one
// workgroup at most)
bool m_xrow_supported(const DevParams &P)
{
    int threads;
    size_t lds;
    m_xrow_shape(P, threads, lds);
    return threads <= 1024 && lds <= (size_t)P.lds_max;
}

hipError_t launch_m_sweep(hipStream_t s, const DevParams &P, const MFields &M, int dir, int seg_len)
{
    if (dir == 0) {
        MXrowFn fn;
        const int fk = fvs_kind(P);
        if (P.periodic[0])
            fn = fk == FVS_SW ? m_pick_xrow<FVS_SW, 1>(P.recon)
                 : fk == FVS_SWE ? m_pick_xrow<FVS_SWE, 1>(P.recon) : m_pick_xrow<FVS_LF, 1>(P.recon);
        else
            fn = fk == FVS_SW ? m_pick_xrow<FVS_SW, 0>(P.recon)
                 : fk == FVS_SWE ? m_pick_xrow<FVS_SWE, 0>(P.recon) : m_pick_xrow<FVS_LF, 0>(P.recon);
        int threads;
        size_t lds;
        m_xrow_shape(P, threads, lds);
        if (!fn || threads > 1024 || lds > (size_t)P.lds_max) return hipErrorInvalidValue;
        if (lds > 65536) (void)hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(fn, dim3(P.ny, P.nz), dim3(threads), lds, s, P, M);
        return hipGetLastError();
    }
    MSweepFn fn;
    const int fk = fvs_kind(P);
    if (dir == 1)
        fn = fk == FVS_SW ? m_pick_sweep<1, FVS_SW>(P.recon)
             : fk == FVS_SWE ? m_pick_sweep<1, FVS_SWE>(P.recon) : m_pick_sweep<1, FVS_LF>(P.recon);
    else
        fn = fk == FVS_SW ? m_pick_sweep<2, FVS_SW>(P.recon)
             : fk == FVS_SWE ? m_pick_sweep<2, FVS_SWE>(P.recon) : m_pick_sweep<2, FVS_LF>(P.recon);
    if (!fn) return hipErrorInvalidValue;
    const int nb = dir == 2 ? P.ny : P.nz;
    const int nd = dir == 1 ? P.ny : P.nz;
    hipLaunchKernelGGL(fn, dim3((nd + seg_len - 1) / seg_len, (P.nx + 63) / 64, nb), dim3(64), 0, s, P, M, seg_len);
    return hipGetLastError();
}

template <int BC>
static void m_gradflux_vis(hipStream_t s, const DevParams &P, const MFields &M, int seg_len, int nwg, dim3 b, int gx,
                           int gy)
{
    switch (P.vis) {
    case VIS_C8: hipLaunchKernelGGL((km_gradflux<VIS_C8, BC>), dim3(nwg), b, 0, s, P, M, seg_len, gx, gy); break;
    case VIS_C6: hipLaunchKernelGGL((km_gradflux<VIS_C6, BC>), dim3(nwg), b, 0, s, P, M, seg_len, gx, gy); break;
    default: hipLaunchKernelGGL((km_gradflux<VIS_C4, BC>), dim3(nwg), b, 0, s, P, M, seg_len, gx, gy); break;
    }
}

// rows 3: the LDS-staged two-row form where the box is periodic, the block
// is 64 x 4 and it tiles x and y exactly; rows 2 (or 3 where that does not
// hold): two rows per thread where the box is periodic and ny is a multiple
// of 2 by; else one row per thread
hipError_t launch_m_gradflux(hipStream_t s, const DevParams &P, const MFields &M, int seg_len, int bx, int by, int rows)
{
    if (bx * by > 256 || bx % 32) return hipErrorInvalidValue;
    if (P.vis != VIS_C8 && P.vis != VIS_C6 && P.vis != VIS_C4) return hipErrorInvalidValue;
    const bool per = P.periodic[0] && P.periodic[1] && P.periodic[2];
    dim3 b(bx, by, 1);
    const int gx = (P.nx + bx - 1) / bx, gz = (P.nz + seg_len - 1) / seg_len;
    if (rows == 3 && per && bx == 64 && by == 4 && P.nx % 64 == 0 && P.ny % 8 == 0) {
        // 64 x 8 threads (16 rows) where that tiles y and P.m_ty asks for it
        if (P.m_ty == 8 && P.ny % 16 == 0) {
            const int gy = P.ny / 16, n = gx * gy * gz;
            const dim3 b8(64, 8, 1);
            switch (P.vis) {
            case VIS_C8: hipLaunchKernelGGL((km_gradflux_lds<VIS_C8, 8>), dim3(n), b8, 0, s, P, M, seg_len, gx, gy); break;
            case VIS_C6: hipLaunchKernelGGL((km_gradflux_lds<VIS_C6, 8>), dim3(n), b8, 0, s, P, M, seg_len, gx, gy); break;
            default: hipLaunchKernelGGL((km_gradflux_lds<VIS_C4, 8>), dim3(n), b8, 0, s, P, M, seg_len, gx, gy); break;
            }
            return hipGetLastError();
        }
        const int gy = P.ny / 8;
        const int n = gx * gy * gz;
        switch (P.vis) {
        case VIS_C8: hipLaunchKernelGGL((km_gradflux_lds<VIS_C8, 4>), dim3(n), b, 0, s, P, M, seg_len, gx, gy); break;
        case VIS_C6: hipLaunchKernelGGL((km_gradflux_lds<VIS_C6, 4>), dim3(n), b, 0, s, P, M, seg_len, gx, gy); break;
        default: hipLaunchKernelGGL((km_gradflux_lds<VIS_C4, 4>), dim3(n), b, 0, s, P, M, seg_len, gx, gy); break;
        }
        return hipGetLastError();
    }
    if (rows >= 2 && per && P.ny % (2 * by) == 0) {
        const int gy = P.ny / (2 * by);
        const int n = gx * gy * gz;
        switch (P.vis) {
        case VIS_C8: hipLaunchKernelGGL((km_gradflux_rows2<VIS_C8>), dim3(n), b, 0, s, P, M, seg_len, gx, gy); break;
        case VIS_C6: hipLaunchKernelGGL((km_gradflux_rows2<VIS_C6>), dim3(n), b, 0, s, P, M, seg_len, gx, gy); break;
        default: hipLaunchKernelGGL((km_gradflux_rows2<VIS_C4>), dim3(n), b, 0, s, P, M, seg_len, gx, gy); break;
        }
        return hipGetLastError();
    }
    const int gy = (P.ny + by - 1) / by;
    if (per) m_gradflux_vis<0>(s, P, M, seg_len, gx * gy * gz, b, gx, gy);
    else m_gradflux_vis<1>(s, P, M, seg_len, gx * gy * gz, b, gx, gy);
    return hipGetLastError();
}

template <int VIS, int BC>
static void m_vdiv_mode(hipStream_t s, const DevParams &P, const MFields &M, int mode, double dt, int seg,
                        unsigned long long *flag, int dev_dt, int bx, int by)
{
    dim3 g((P.nx + bx - 1) / bx, (P.ny + by - 1) / by, (P.nz + seg - 1) / seg);
    dim3 b(bx, by, 1);
    switch (mode) {
    case 0: hipLaunchKernelGGL((km_vdiv_z<VIS, 0, BC>), g, b, 0, s, P, M, dt, seg, flag, dev_dt); break;
    case 1: hipLaunchKernelGGL((km_vdiv_z<VIS, 1, BC>), g, b, 0, s, P, M, dt, seg, flag, dev_dt); break;
    case 2: hipLaunchKernelGGL((km_vdiv_z<VIS, 2, BC>), g, b, 0, s, P, M, dt, seg, flag, dev_dt); break;
    default: hipLaunchKernelGGL((km_vdiv_z<VIS, 3, BC>), g, b, 0, s, P, M, dt, seg, flag, dev_dt); break;
    }
}

template <int BC>
static hipError_t m_vdiv_vis(hipStream_t s, const DevParams &P, const MFields &M, int mode, double dt, int seg,
                             unsigned long long *flag, int dev_dt, int bx, int by)
{
    switch (P.vis) {
    case VIS_C8: m_vdiv_mode<VIS_C8, BC>(s, P, M, mode, dt, seg, flag, dev_dt, bx, by); break;
    case VIS_C6: m_vdiv_mode<VIS_C6, BC>(s, P, M, mode, dt, seg, flag, dev_dt, bx, by); break;
    case VIS_C4: m_vdiv_mode<VIS_C4, BC>(s, P, M, mode, dt, seg, flag, dev_dt, bx, by); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool m_vdiv_cfl_next(const DevParams &P, int bx, int by)
{
    return (bx * by) % 64 == 0 && P.nx % bx == 0 && P.ny % by == 0;
}

// thread rows of the LDS-staged stage pass: 8 where P.m_ty asks for it and
// 8 divides ny, else 4
static int m_vdiv_ty(const DevParams &P) { return P.m_ty == 8 && P.ny % 8 == 0 ? 8 : 4; }

template <int VIS, bool IMG, int TY>
static void m_vdiv_lds_ty(hipStream_t s, const DevParams &P, const MFields &M, int mode, double dt, int seg,
                          unsigned long long *flag, int dev_dt)
{
    const int gx = P.nx / 64, gy = P.ny / TY, n = gx * gy * ((P.nz + seg - 1) / seg);
    dim3 b(64, TY, 1);
    switch (mode) {
    case 0: hipLaunchKernelGGL((km_vdiv_lds<VIS, 0, false, TY>), dim3(n), b, 0, s, P, M, dt, seg, gx, gy, flag, dev_dt); break;
    case 1: hipLaunchKernelGGL((km_vdiv_lds<VIS, 1, IMG, TY>), dim3(n), b, 0, s, P, M, dt, seg, gx, gy, flag, dev_dt); break;
    case 2: hipLaunchKernelGGL((km_vdiv_lds<VIS, 2, IMG, TY>), dim3(n), b, 0, s, P, M, dt, seg, gx, gy, flag, dev_dt); break;
    default: hipLaunchKernelGGL((km_vdiv_lds<VIS, 3, IMG, TY>), dim3(n), b, 0, s, P, M, dt, seg, gx, gy, flag, dev_dt); break;
    }
}

template <int VIS, bool IMG>
static void m_vdiv_lds(hipStream_t s, const DevParams &P, const MFields &M, int mode, double dt, int seg,
                       unsigned long long *flag, int dev_dt)
{
    if (m_vdiv_ty(P) == 8) m_vdiv_lds_ty<VIS, IMG, 8>(s, P, M, mode, dt, seg, flag, dev_dt);
    else m_vdiv_lds_ty<VIS, IMG, 4>(s, P, M, mode, dt, seg, flag, dev_dt);
}

// whether launch_m_vdiv takes the LDS-staged form for these arguments
bool m_vdiv_lds_form(const DevParams &P, int bx, int by, int lds)
{
    const bool per = P.periodic[0] && P.periodic[1] && P.periodic[2];
    return lds && per && bx == 64 && by == 4 && P.nx % 64 == 0 && P.ny % 4 == 0;
}

// whether a stage can run as km_vdiv_zs (the z sweep folded into the
// LDS-staged stage kernel with ghost images): the LDS form's conditions and
// a reconstruction built into it
bool m_vdiv_zs_form(const DevParams &P, int bx, int by, int lds, int img)
{
    return img && m_vdiv_lds_form(P, bx, by, lds) && (P.recon == R_WENO5 || P.recon == R_UD7) &&
           (P.vis == VIS_C8 || P.vis == VIS_C6 || P.vis == VIS_C4);
}

template <int VIS, int RECON, int FVS, int TY>
static void m_vdiv_zs_ty(hipStream_t s, const DevParams &P, const MFields &M, const Shadow &So, int mode, double dt,
                         int seg, unsigned long long *flag, int dev_dt)
{
    const int gx = P.nx / 64, gy = P.ny / TY, n = gx * gy * ((P.nz + seg - 1) / seg);
    const dim3 b(64, TY, 1);
    if (mode == 1) hipLaunchKernelGGL((km_vdiv_zs<VIS, 1, RECON, FVS, TY>), dim3(n), b, 0, s, P, M, So, dt, seg, gx, gy, flag, dev_dt);
    else if (mode == 2) hipLaunchKernelGGL((km_vdiv_zs<VIS, 2, RECON, FVS, TY>), dim3(n), b, 0, s, P, M, So, dt, seg, gx, gy, flag, dev_dt);
    else hipLaunchKernelGGL((km_vdiv_zs<VIS, 3, RECON, FVS, TY>), dim3(n), b, 0, s, P, M, So, dt, seg, gx, gy, flag, dev_dt);
}

template <int VIS, int RECON, int FVS>
static void m_vdiv_zs(hipStream_t s, const DevParams &P, const MFields &M, const Shadow &So, int mode, double dt,
                      int seg, unsigned long long *flag, int dev_dt)
{
    if (m_vdiv_ty(P) == 8) m_vdiv_zs_ty<VIS, RECON, FVS, 8>(s, P, M, So, mode, dt, seg, flag, dev_dt);
    else m_vdiv_zs_ty<VIS, RECON, FVS, 4>(s, P, M, So, mode, dt, seg, flag, dev_dt);
}

template <int VIS, int RECON>
static void m_vdiv_zs_f(hipStream_t s, const DevParams &P, const MFields &M, const Shadow &So, int mode, double dt,
                        int seg, unsigned long long *flag, int dev_dt)
{
    const int fk = fvs_kind(P);
    if (fk == FVS_SW) m_vdiv_zs<VIS, RECON, FVS_SW>(s, P, M, So, mode, dt, seg, flag, dev_dt);
    else if (fk == FVS_SWE) m_vdiv_zs<VIS, RECON, FVS_SWE>(s, P, M, So, mode, dt, seg, flag, dev_dt);
    else m_vdiv_zs<VIS, RECON, FVS_LF>(s, P, M, So, mode, dt, seg, flag, dev_dt);
}

template <int VIS>
static void m_vdiv_zs_r(hipStream_t s, const DevParams &P, const MFields &M, const Shadow &So, int mode, double dt,
                        int seg, unsigned long long *flag, int dev_dt)
{
    if (P.recon == R_WENO5) m_vdiv_zs_f<VIS, R_WENO5>(s, P, M, So, mode, dt, seg, flag, dev_dt);
    else m_vdiv_zs_f<VIS, R_UD7>(s, P, M, So, mode, dt, seg, flag, dev_dt);
}

// a stage (mode 1..3) as km_vdiv_zs: the z sweep and the stage in one
// kernel, the new shadow (with its ghost images) into So
hipError_t launch_m_vdiv_zs(hipStream_t s, const DevParams &P, const MFields &M, const Shadow &So, int mode,
                            double dt, int seg, unsigned long long *flag, int dev_dt)
{
    if (mode < 1 || mode > 3 || !m_vdiv_zs_form(P, 64, 4, 1, 1)) return hipErrorInvalidValue;
    if (dev_dt == 2 && !m_vdiv_cfl_next(P, 64, 4)) return hipErrorInvalidValue;
    switch (P.vis) {
    case VIS_C8: m_vdiv_zs_r<VIS_C8>(s, P, M, So, mode, dt, seg, flag, dev_dt); break;
    case VIS_C6: m_vdiv_zs_r<VIS_C6>(s, P, M, So, mode, dt, seg, flag, dev_dt); break;
    default: m_vdiv_zs_r<VIS_C4>(s, P, M, So, mode, dt, seg, flag, dev_dt); break;
    }
    return hipGetLastError();
}

// lds: the LDS-staged form where the box is periodic, the block is 64 x 4
// and it tiles x and y exactly; with img the stages (mode 1..3) of that
// form also write the shadow's periodic ghost images
hipError_t launch_m_vdiv(hipStream_t s, const DevParams &P, const MFields &M, int mode, double dt, int seg,
                         unsigned long long *flag, int dev_dt, int bx, int by, int lds, int img)
{
    if (bx * by > 256 || bx % 32) return hipErrorInvalidValue;
    if (dev_dt == 2 && !m_vdiv_cfl_next(P, bx, by)) return hipErrorInvalidValue;   // partial waves
    const bool per = P.periodic[0] && P.periodic[1] && P.periodic[2];
    if (m_vdiv_lds_form(P, bx, by, lds)) {
        switch (P.vis) {
        case VIS_C8: img ? m_vdiv_lds<VIS_C8, true>(s, P, M, mode, dt, seg, flag, dev_dt)
                         : m_vdiv_lds<VIS_C8, false>(s, P, M, mode, dt, seg, flag, dev_dt); break;
        case VIS_C6: img ? m_vdiv_lds<VIS_C6, true>(s, P, M, mode, dt, seg, flag, dev_dt)
                         : m_vdiv_lds<VIS_C6, false>(s, P, M, mode, dt, seg, flag, dev_dt); break;
        case VIS_C4: img ? m_vdiv_lds<VIS_C4, true>(s, P, M, mode, dt, seg, flag, dev_dt)
                         : m_vdiv_lds<VIS_C4, false>(s, P, M, mode, dt, seg, flag, dev_dt); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (per)
        return m_vdiv_vis<0>(s, P, M, mode, dt, seg, flag, dev_dt, bx, by);
    return m_vdiv_vis<1>(s, P, M, mode, dt, seg, flag, dev_dt, bx, by);
}

}  // namespace ns3d
