// ns3d_device.hpp -- device functions of the FP64 RHS path (gfx950).
//
// Every expression keeps the reference's evaluation order; the library is
// built with -ffp-contract=off so the FP64 path forms no FMA, matching the
// x86-64 baseline build of the reference (reference Makefile:6).
#pragma once

#include <hip/hip_runtime.h>

namespace ns3d {

// A double constant as the float pair h + l (h the nearest float, l the
// nearest float to the remainder), for FP32 arithmetic: s * h + s * l scales
// by the constant with the rounding of one product, where a float rounded
// once would scale every use in a run by the same wrong factor.
struct F2 {
    float h, l;
};

// Scalars every kernel needs, passed by value.
struct DevParams {
    double gamma, Rgas, Cp, Pr, mu, sutherland_s, mu_power, sw_eps;
    double mdcd_disp, mdcd_diss, ud7l_alpha;
    double dx, dy, dz, idx, idy, idz;
    int fvs, recon, vis, mu_law;
    int nx, ny, nz, ng, sx, sy, sz;
    int periodic[3];         // axis has both sides periodic (initialization.cpp:315-323)
    int bc[6];               // ns3d_bc_type of xmin, xmax, ymin, ymax, zmin, zmax
    double rho_ref, p_ref;   // reference state of the mixed-precision shadow
    int adj_lines;           // 1: adjoint sweeps may hold whole lines in LDS
    int adj_zmarch;          // 1: the viscous adjoint kernels march z segments (periodic boxes)
    int m_ty;                // 4 or 8: thread rows of the mixed LDS-staged viscous passes
    int lds_max;             // LDS bytes one workgroup may use on this device
    double cfl;              // CFL number of the device time step
    int check_mask;          // bit s: stage s scans for non-physical cells (the third only)
    // constants of the mixed-precision path (set_fp32_pairs): id2[a] =
    // 1 / d_a; dsc[a][f] = 1 / (K_f d_a) for the integer-weight derivative sums,
    // K = 840 (C8), 60 (C6), 12 (C4), 2 (second-order closures); mu0 the
    // mu law's constant (mu, or mu (1 + S) for Sutherland), mk0 = mu0 Cp / Pr,
    // suth = S
    F2 id2[3];
    F2 dsc[3][4];
    F2 mu0, mk0, suth;
};

// float pair of x (host and device)
__host__ __device__ inline F2 f2_of(double x)
{
    F2 r;
    r.h = (float)x;
    r.l = (float)(x - (double)r.h);
    return r;
}

// s times the constant c (one FMA; contraction-independent)
__device__ __forceinline__ float mul2(float s, F2 c) { return fmaf(s, c.h, s * c.l); }

// the F2 members of d from its FP64 members
inline void set_fp32_pairs(DevParams &d)
{
    const double id[3] = {d.idx, d.idy, d.idz};
    const double K[4] = {840.0, 60.0, 12.0, 2.0};
    for (int a = 0; a < 3; ++a) {
        d.id2[a] = f2_of(id[a]);
        for (int f = 0; f < 4; ++f) d.dsc[a][f] = f2_of(id[a] / K[f]);
    }
    const double mu0 = d.mu_law == 0 ? d.mu * (1.0 + d.sutherland_s) : d.mu;
    d.mu0 = f2_of(mu0);
    d.mk0 = f2_of(mu0 * (d.Cp / d.Pr));
    d.suth = f2_of(d.sutherland_s);
}

// Slots of a context's device flag words: [0] key of the first non-physical
// cell, [1] the CFL minimum (bits of a positive double), [2] the RK3 step
// counter, [3] the simulated time of the device-time-step loop, [4] the CFL
// minimum of the state the running step produces, where its third stage
// reduces it (the step's end moves it to [1] for the next step).
enum { FLAG_BAD = 0, FLAG_CFL = 1, FLAG_STEP = 2, FLAG_TIME = 3, FLAG_CFL_NEXT = 4, FLAG_WORDS = 5 };

// dt = cfl * min over cells (timestep.cpp:32), from the CFL reduction
__device__ __forceinline__ double cfl_dt(const DevParams &P, const unsigned long long *flag)
{
    return P.cfl * __longlong_as_double((long long)flag[FLAG_CFL]);
}

// one cell's bound of compute_timestep (timestep.cpp:14-27): the smallest
// d / (|velocity| + a) over the three axes
__device__ __forceinline__ double cfl_cell(const DevParams &P, double rho, double u, double v, double w, double p)
{
    const double a = sqrt(P.gamma * p / fmax(rho, 1e-14));
    return fmin(P.dx / (fabs(u) + a), fmin(P.dy / (fabs(v) + a), P.dz / (fabs(w) + a)));
}

// The same minimum over a set of cells, taken per axis as d / max(|vel| + a):
// a correctly rounded division is monotone in its divisor, so this equals
// the minimum of the per-cell bounds bit for bit, with one division and one
// square root per cell instead of four and one.  cfl_accum folds a cell in
// (maxima start at 0; NaN bounds drop out as fmin drops them), cfl_bound
// gives the set's minimum (1e9 for an empty set, compute_timestep's start).
__device__ __forceinline__ void cfl_accum(const DevParams &P, double rho, double u, double v, double w, double p,
                                          double &cx, double &cy, double &cz)
{
    const double a = sqrt(P.gamma * p / fmax(rho, 1e-14));
    cx = fmax(cx, fabs(u) + a);
    cy = fmax(cy, fabs(v) + a);
    cz = fmax(cz, fabs(w) + a);
}

__device__ __forceinline__ double cfl_bound(const DevParams &P, double cx, double cy, double cz)
{
    return fmin(1e9, fmin(P.dx / cx, fmin(P.dy / cy, P.dz / cz)));
}

// minimum over the 64 lanes of a wave (every lane active)
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// key of a non-physical cell: the step counter in the high bits, so the
// atomicMin keeps the earliest step's lowest cell index
__device__ __forceinline__ unsigned long long nonphysical_key(const unsigned long long *flag, int id)
{
    return (flag[FLAG_STEP] << 40) | (unsigned long long)id;
}

// FVS_SWE: Steger-Warming with the eps-smoothed eigenvalue split (sw_eps !=
// 0).  The split kernels take the form as a template value: chosen at run
// time, the compiler evaluated both forms of every eigenvalue (three FP64
// square roots per split) and selected one.  P.fvs itself stays SW or LF.
enum { FVS_SW = 0, FVS_LF = 1, FVS_SWE = 2 };

// the split kernels' template value for the parameters
__host__ __device__ inline int fvs_kind(const DevParams &P)
{
    return P.fvs == FVS_LF ? FVS_LF : (P.sw_eps != 0.0 ? FVS_SWE : FVS_SW);
}
enum { R_WENO5 = 0, R_LINEAR = 2, R_MDCD = 3, R_MDCD_HYBRID = 4, R_UD7 = 5, R_UD7L = 6 };
enum { VIS_C8 = 0, VIS_C6 = 1, VIS_C4 = 2 };

// stencil width per reconstruction (initialization.cpp:329-357)
template <int RECON> struct Stencil;
template <> struct Stencil<R_WENO5> { static constexpr int S = 6; };
template <> struct Stencil<R_LINEAR> { static constexpr int S = 2; };
template <> struct Stencil<R_MDCD> { static constexpr int S = 6; };
template <> struct Stencil<R_MDCD_HYBRID> { static constexpr int S = 6; };
template <> struct Stencil<R_UD7> { static constexpr int S = 8; };
template <> struct Stencil<R_UD7L> { static constexpr int S = 8; };

// --------------------------------------------------------- XCD mapping
// Workgroups are dealt round-robin to the 8 XCDs of the device, each with
// its own L2.  Remap the hardware workgroup id so that XCD x receives the
// contiguous logical range [x*n/8, (x+1)*n/8): tiles that share halos then
// share an L2.  Falls back to identity when n is not a multiple of 8.
__device__ __forceinline__ int xcd_logical(int wg, int nwg)
{
    if (nwg & 7) return wg;
    return (wg & 7) * (nwg >> 3) + (wg >> 3);
}

// Thread t of the periodic ghost fill of one axis: the ghost cell dst of
// the 2 ng layers on the axis' two faces (face cells only: the stencils are
// axis-aligned and read no edge or corner ghosts) and its periodic image
// src in the interior.  On the x faces the ghost index runs fastest (the
// ghost cells of a row are adjacent); on the y and z faces x runs fastest,
// so a wave writes whole row segments.  False past the last ghost cell.
__device__ __forceinline__ bool periodic_ghost_pair(const DevParams &P, int axis, int t, int &dst, int &src)
{
    const int ng = P.ng;
    const int nd = axis == 0 ? P.nx : axis == 1 ? P.ny : P.nz;
    const int na = axis == 0 ? P.ny : P.nx;   // fastest cross axis
    const int nb = axis == 2 ? P.ny : P.nz;
    if (t >= 2 * ng * na * nb) return false;
    int g, a, b;
    if (axis == 0) {
        g = t % (2 * ng);
        const int rest = t / (2 * ng);
        a = ng + rest % na;
        b = ng + rest / na;
    } else {
        a = ng + t % na;
        const int rest = t / na;
        g = rest % (2 * ng);
        b = ng + rest / (2 * ng);
    }
    int q, qs;   // ghost layer and its image along the axis
    if (g < ng) { q = g; qs = g + nd; }
    else { q = ng + nd + (g - ng); qs = ng + (g - ng); }
    if (axis == 0) { dst = (b * P.sy + a) * P.sx + q; src = (b * P.sy + a) * P.sx + qs; }
    else if (axis == 1) { dst = (b * P.sy + q) * P.sx + a; src = (b * P.sy + qs) * P.sx + a; }
    else { dst = (q * P.sy + b) * P.sx + a; src = (qs * P.sy + b) * P.sx + a; }
    return true;
}

// logical (x, y, z) block coordinates of a 3D grid after the XCD remap
__device__ __forceinline__ void xcd_block3(int &bx, int &by, int &bz)
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int nwg = gx * gy * gridDim.z;
    const int wg = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int l = xcd_logical(wg, nwg);
    bx = l % gx;
    by = (l / gx) % gy;
    bz = l / (gx * gy);
}

// For kernels whose blocks each take one tile of one plane and read R planes
// above and below it (the non-marching 3D stencils): logical blocks go band
// by band (about 16 tiles of a plane: band = 16 / gx tile rows), within a
// band plane by plane.  The blocks an XCD runs at once are then a few
// consecutive planes of one band, whose z stencil (~3 MB for 8 FP64 arrays
// and R = 4) stays in its L2; plane by plane over whole planes, as
// xcd_block3 orders them, the 2R + 1 planes of a 128^2 plane exceed it.
// Falls back to xcd_block3's order where the band does not divide gy.
__device__ __forceinline__ void xcd_block3_band(int &bx, int &by, int &bz)
{
    const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const int nwg = gx * gy * gz;
    const int wg = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int l = xcd_logical(wg, nwg);
    int band = 16 / gx;
    if (band < 1) band = 1;
    if (gy % band) band = gy;
    const int per_band = gx * band * gz;
    const int b = l / per_band, r = l % per_band;
    bx = r % gx;
    by = b * band + (r / gx) % band;
    bz = r / (gx * band);
}

// 1/x: the hardware estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r) (quadratic convergence; x > 0 and normal where used).
// For the mixed-precision and derivative kernels, which need the value and
// not the reference's rounding; the FP64 parity path divides.
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// ----------------------------------------------------------- conversions

// E and c of primitiveToConserved (field_structures.h:445-449) from a
// primitive state whose density is already clamped.
__device__ __forceinline__ void p2c_energy_sound(double gamma, double rr, double u, double v,
                                                 double w, double p, double &E, double &c)
{
    double kinetic = 0.5 * (u * u + v * v + w * w);
    double e_internal = p / ((gamma - 1.0) * rr);
    E = rr * (e_internal + kinetic);
    c = sqrt(gamma * p / rr);
}

// T^1.5 for the Sutherland law: T * sqrt(T) evaluated with an exact FMA
// residual for the product and one Newton correction of the square root
// (Dekker-style compensation), i.e. correctly rounded except within a tiny
// fraction of an ulp of a rounding midpoint.  The reference calls the host
// libm pow(T, 1.5), which is correctly rounded in all but ~0.1% of inputs;
// the device library's pow is faithfully rounded only (off by one ulp for
// ~15% of T in [0.5, 2]), which would make the FP64 path drift from the
// reference at the last bit.
__device__ __forceinline__ double pow15(double T)
{
    const double s = sqrt(T);
    const double es = fma(-s, s, T) / (2.0 * s);   // sqrt(T) - s, first order
    const double p = T * s;
    const double pe = fma(T, s, -p);                // exact rounding error of T*s
    return p + (pe + T * es);
}

// SolverParams::get_mu (field_structures.h:94-105) + selectable laws.
__device__ __forceinline__ double get_mu(const DevParams &P, double T)
{
    if (P.mu_law == 1) return P.mu;
    if (P.mu_law == 2) return P.mu * pow(T, P.mu_power);
    double S_ref = P.sutherland_s;
    return P.mu * pow15(T) * (1.0 + S_ref) / (T + S_ref);
}

// --------------------------------------------------------------- splits

__device__ __forceinline__ void lam_split(double lam, double eps, double &lp, double &lm)
{
    if (eps == 0.0) {
        lp = 0.5 * (lam + fabs(lam));
        lm = 0.5 * (lam - fabs(lam));
    } else {  // README.md:179
        double r = sqrt(lam * lam + eps * eps);
        lp = 0.5 * (lam + r);
        lm = 0.5 * (lam - r);
    }
}

// lam_split with its form fixed by the kernel's FVS value (the same
// expressions)
template <int FVS>
__device__ __forceinline__ void lam_split_fvs(double lam, double eps, double &lp, double &lm)
{
    if constexpr (FVS == FVS_SWE) {
        double r = sqrt(lam * lam + eps * eps);
        lp = 0.5 * (lam + r);
        lm = 0.5 * (lam - r);
    } else {
        lp = 0.5 * (lam + fabs(lam));
        lm = 0.5 * (lam - fabs(lam));
    }
}

// split_StagerWarming / split_LaxFriedrichs (InviscousFlux_new.cpp:439-530)
// for the axis (N1,N2,N3); E and c come from p2c_energy_sound.
template <int FVS, int DIR>
__device__ __forceinline__ void split_state(const DevParams &P, double rho, double u, double v,
                                            double w, double p, double Fp[5], double Fm[5])
{
    constexpr double n1 = DIR == 0 ? 1.0 : 0.0;
    constexpr double n2 = DIR == 1 ? 1.0 : 0.0;
    constexpr double n3 = DIR == 2 ? 1.0 : 0.0;
    const double gamma = P.gamma;
    double E, c;
    p2c_energy_sound(gamma, rho, u, v, w, p, E, c);
    if constexpr (FVS == FVS_SW || FVS == FVS_SWE) {
        double un = u * n1 + v * n2 + w * n3;
        double lambda1 = un - c;
        double lambda2 = un;
        double lambda3 = un + c;
        double l1p, l1m, l2p, l2m, l3p, l3m;
        lam_split_fvs<FVS>(lambda1, P.sw_eps, l1p, l1m);
        lam_split_fvs<FVS>(lambda2, P.sw_eps, l2p, l2m);
        lam_split_fvs<FVS>(lambda3, P.sw_eps, l3p, l3m);
        double S1_p = l1p + l3p;
        double S1_m = l1m + l3m;
        double S2_p = l3p - l1p;
        double S2_m = l3m - l1m;
        double H = (E + p) / rho;
        double ga1 = rho / (2.0 * gamma);
        double ga2 = 2.0 * (gamma - 1.0);
        double V = 0.5 * (u * u + v * v + w * w);
        Fp[0] = ga1 * (S1_p + ga2 * l2p);
        Fp[1] = ga1 * (S1_p * u + ga2 * l2p * u + S2_p * n1 * c);
        Fp[2] = ga1 * (S1_p * v + ga2 * l2p * v + S2_p * n2 * c);
        Fp[3] = ga1 * (S1_p * w + ga2 * l2p * w + S2_p * n3 * c);
        Fp[4] = ga1 * (S1_p * H + ga2 * l2p * V + S2_p * un * c);
        Fm[0] = ga1 * (S1_m + ga2 * l2m);
        Fm[1] = ga1 * (S1_m * u + ga2 * l2m * u + S2_m * n1 * c);
        Fm[2] = ga1 * (S1_m * v + ga2 * l2m * v + S2_m * n2 * c);
        Fm[3] = ga1 * (S1_m * w + ga2 * l2m * w + S2_m * n3 * c);
        Fm[4] = ga1 * (S1_m * H + ga2 * l2m * V + S2_m * un * c);
    } else {
        double un = u * n1 + v * n2 + w * n3;
        double a_max = fabs(un) + c;
        double F1 = rho * un;
        double F2 = rho * u * un + p * n1;
        double F3 = rho * v * un + p * n2;
        double F4 = rho * w * un + p * n3;
        double F5 = (E + p) * un;
        Fp[0] = 0.5 * (F1 + a_max * rho);
        Fp[1] = 0.5 * (F2 + a_max * rho * u);
        Fp[2] = 0.5 * (F3 + a_max * rho * v);
        Fp[3] = 0.5 * (F4 + a_max * rho * w);
        Fp[4] = 0.5 * (F5 + a_max * E);
        Fm[0] = 0.5 * (F1 - a_max * rho);
        Fm[1] = 0.5 * (F2 - a_max * rho * u);
        Fm[2] = 0.5 * (F3 - a_max * rho * v);
        Fm[3] = 0.5 * (F4 - a_max * rho * w);
        Fm[4] = 0.5 * (F5 - a_max * E);
    }
}

// ------------------------------------------------------- reconstructions
// s[] is the stencil already in the orientation reconstruct_select passes
// (reversed for the minus side, reconstruction.cpp:29-72).

__device__ __forceinline__ double recon_weno5(const double *s)
{   // reconstruction.cpp:93-124
    double f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
    double eps = 1e-6;
    double beta0 = 13.0 / 12.0 * (f0 - 2 * f1 + f2) * (f0 - 2 * f1 + f2)
                 + 0.25 * (f0 - 4 * f1 + 3 * f2) * (f0 - 4 * f1 + 3 * f2);
    double beta1 = 13.0 / 12.0 * (f1 - 2 * f2 + f3) * (f1 - 2 * f2 + f3)
                 + 0.25 * (f1 - f3) * (f1 - f3);
    double beta2 = 13.0 / 12.0 * (f2 - 2 * f3 + f4) * (f2 - 2 * f3 + f4)
                 + 0.25 * (3 * f2 - 4 * f3 + f4) * (3 * f2 - 4 * f3 + f4);
    double alpha0 = 0.1 / ((eps + beta0) * (eps + beta0));
    double alpha1 = 0.6 / ((eps + beta1) * (eps + beta1));
    double alpha2 = 0.3 / ((eps + beta2) * (eps + beta2));
    double asum = alpha0 + alpha1 + alpha2;
    double w0 = alpha0 / asum;
    double w1 = alpha1 / asum;
    double w2 = alpha2 / asum;
    double p0 = (2 * f0 - 7 * f1 + 11 * f2) / 6.0;
    double p1 = (-f1 + 5 * f2 + 2 * f3) / 6.0;
    double p2 = (2 * f2 + 5 * f3 - f4) / 6.0;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

// ------------------------------------ quotients with a refined reciprocal
// hipcc takes an FP64 quotient a / b on gfx950 as
//   d = div_scale(b, b, a), n = div_scale(a, b, a), r = rcp(d),
//   r = fma(r, fma(-d, r, 1), r) twice, q = n r, e = fma(-d, q, n),
//   a / b = div_fixup(div_fmas(e, r, q), b, a).
// For finite operands away from the ends of the exponent range (a above
// 2^-969, a / b normal and below 2^768 b, b below 2^1022) div_scale returns
// its operand, div_fmas is fma(e, r, q) and div_fixup returns its first
// operand, so quot_rcp(a, b, rcp_refined(b)) is the same bits as a / b, and
// one refined reciprocal serves every quotient with that denominator.  (A
// numerator of -0 gives +0 here: equal as a number, the sign bit differs.)

__device__ __forceinline__ double rcp_refined(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    return r;
}

__device__ __forceinline__ double quot_rcp(double a, double b, double r)
{
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}

// The refined reciprocal of 6 as the device computes it for x / 6.0 (the
// empty asm keeps the compiler from folding rcp of a constant into a value
// the hardware's rcp need not give).
__device__ __forceinline__ double rcp_refined_six()
{
    double six = 6.0;
    asm volatile("" : "+v"(six));
    return rcp_refined(six);
}

// The range of split fluxes under which recon_weno5_rcp equals recon_weno5
// bit for bit: zero or 2^-900 <= |f| <= 2^200.  Then each smoothness
// indicator is below 2^406, (eps + beta)^2 lies in [2^-40, 2^812], each
// alpha in [2^-816, 2^40], each weight in [2^-857, 1], and a nonzero
// numerator of the 1/6 quotients, an integer combination of stencil values
// that are multiples of 2^-952, is at least 2^-952 and below 2^205: every
// quotient inside the range where the steps above equal a / b.  Non-finite
// values fall outside.
__device__ __forceinline__ bool weno_rcp_range_ok(double f)
{
    const double a = fabs(f);
    return a <= 0x1p200 && (a >= 0x1p-900 || a == 0.0);
}

__device__ __forceinline__ bool weno_rcp_range_ok5(const double f[5])
{
    bool ok = true;
#pragma unroll
    for (int m = 0; m < 5; ++m) ok = ok && weno_rcp_range_ok(f[m]);
    return ok;
}

// recon_weno5 (reconstruction.cpp:93-124) with its nine quotients taken by
// the steps above: three reciprocals for the alphas, one shared by the
// three weights, and r6 (rcp_refined_six) for the 1/6 of the candidate
// polynomials; 47 instead of 99 VALU operations.  The same expressions in
// the same order otherwise.  Only for stencils that weno_rcp_range_ok
// admits.
__device__ __forceinline__ double recon_weno5_rcp(const double *s, double r6)
{
    double f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
    double eps = 1e-6;
    double beta0 = 13.0 / 12.0 * (f0 - 2 * f1 + f2) * (f0 - 2 * f1 + f2)
                 + 0.25 * (f0 - 4 * f1 + 3 * f2) * (f0 - 4 * f1 + 3 * f2);
    double beta1 = 13.0 / 12.0 * (f1 - 2 * f2 + f3) * (f1 - 2 * f2 + f3)
                 + 0.25 * (f1 - f3) * (f1 - f3);
    double beta2 = 13.0 / 12.0 * (f2 - 2 * f3 + f4) * (f2 - 2 * f3 + f4)
                 + 0.25 * (3 * f2 - 4 * f3 + f4) * (3 * f2 - 4 * f3 + f4);
    double d0 = (eps + beta0) * (eps + beta0);
    double d1 = (eps + beta1) * (eps + beta1);
    double d2 = (eps + beta2) * (eps + beta2);
    double alpha0 = quot_rcp(0.1, d0, rcp_refined(d0));
    double alpha1 = quot_rcp(0.6, d1, rcp_refined(d1));
    double alpha2 = quot_rcp(0.3, d2, rcp_refined(d2));
    double asum = alpha0 + alpha1 + alpha2;
    const double ra = rcp_refined(asum);
    double w0 = quot_rcp(alpha0, asum, ra);
    double w1 = quot_rcp(alpha1, asum, ra);
    double w2 = quot_rcp(alpha2, asum, ra);
    double p0 = quot_rcp(2 * f0 - 7 * f1 + 11 * f2, 6.0, r6);
    double p1 = quot_rcp(-f1 + 5 * f2 + 2 * f3, 6.0, r6);
    double p2 = quot_rcp(2 * f2 + 5 * f3 - f4, 6.0, r6);
    return w0 * p0 + w1 * p1 + w2 * p2;
}

__device__ __forceinline__ double recon_mdcd(const double *s, double disp, double diss)
{   // reconstruction.cpp:126-135
    return (1.0 / 2.0 * disp + 1.0 / 2.0 * diss) * s[0] +
           (-3.0 / 2.0 * disp - 5.0 / 2.0 * diss - 1.0 / 12.0) * s[1] +
           (7.0 / 12.0 + disp + 5.0 * diss) * s[2] +
           (7.0 / 12.0 + disp - 5.0 * diss) * s[3] +
           (-3.0 / 2.0 * disp + 5.0 / 2.0 * diss - 1.0 / 12.0) * s[4] +
           (1.0 / 2.0 * disp - 1.0 / 2.0 * diss) * s[5];
}

__device__ __forceinline__ double recon_mdcd_hybrid(const double *s, double disp, double diss)
{   // reconstruction.cpp:151-225 (std::pow(x,2) is x*x after g++ folding)
    double eps_small = 1e-40;
    double f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4], f5 = s[5];
    double eps = 0.9 * 0.4 / (1.0 - 0.9 * 0.4) * 1e-4;
    double a1 = fabs(f2 - f1) + fabs(f2 - 2 * f1 + f0);
    double b1 = fabs(f2 - f3) + fabs(f2 - 2 * f3 + f4);
    double a2 = fabs(f3 - f2) + fabs(f3 - 2 * f2 + f1);
    double b2 = fabs(f3 - f4) + fabs(f3 - 2 * f4 + f5);
    double sai1 = (2 * a1 * b1 + eps) / (a1 * a1 + b1 * b1 + eps);
    double sai2 = (2 * a2 * b2 + eps) / (a2 * a2 + b2 * b2 + eps);
    double sai = (sai2 < sai1) ? sai2 : sai1;
    if (sai > 0.4) return recon_mdcd(s, disp, diss);
    double q0 = (2 * f0 - 7 * f1 + 11 * f2) / 6.0;
    double q1 = (-f1 + 5 * f2 + 2 * f3) / 6.0;
    double q2 = (2 * f2 + 5 * f3 - f4) / 6.0;
    double q3 = (11 * f3 - 7 * f4 + 2 * f5) / 6.0;
    double d0 = 1.5 * (disp + diss);
    double d1 = 0.5 - 1.5 * (disp - 3 * diss);
    double d2 = 0.5 - 1.5 * (disp + 3 * diss);
    double d3 = 1.5 * (disp - diss);
    double t0 = f0 - 2 * f1 + f2, t0b = f0 - 4 * f1 + 3 * f2;
    double t1 = f1 - 2 * f2 + f3, t1b = f1 - f3;
    double t2 = f2 - 2 * f3 + f4, t2b = 3 * f2 - 4 * f3 + f4;
    double beta0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (t0b * t0b);
    double beta1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (t1b * t1b);
    double beta2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (t2b * t2b);
    double beta3 =
        (271799 * (f0 * f0)
         + f0 * (-2380800 * f1 + 4086352 * f2 - 3462252 * f3 + 1458762 * f4 - 245620 * f5)
         + f1 * (5653317 * f1 - 20427884 * f2 + 17905032 * f3 - 7727988 * f4 + 1325006 * f5)
         + f2 * (19510972 * f2 - 35817664 * f3 + 15929912 * f4 - 2792660 * f5)
         + f3 * (17195652 * f3 - 15880404 * f4 + 2863984 * f5)
         + f4 * (3824847 * f4 - 1429976 * f5)
         + 139633 * (f5 * f5)) / 120960.0;
    double tau6 = fabs(beta3 - (beta0 + 4 * beta1 + beta2) / 6.0);
    double g0 = 20 + tau6 / (beta0 + eps_small);
    double g1 = 20 + tau6 / (beta1 + eps_small);
    double g2 = 20 + tau6 / (beta2 + eps_small);
    double g3 = 20 + tau6 / (beta3 + eps_small);
    double aa0 = d0 * (g0 * g0);
    double aa1 = d1 * (g1 * g1);
    double aa2 = d2 * (g2 * g2);
    double aa3 = d3 * (g3 * g3);
    return (aa0 * q0 + aa1 * q1 + aa2 * q2 + aa3 * q3) / (aa0 + aa1 + aa2 + aa3);
}

__device__ __forceinline__ double recon_ud7(const double *s)
{   // reconstruction.cpp:227-232
    return (-1.0 / 140.0) * s[0] + (5.0 / 84.0) * s[1] - (101.0 / 420.0) * s[2]
           + (319.0 / 420.0) * s[3] + (107.0 / 210.0) * s[4]
           - (19.0 / 210.0) * s[5] + (1.0 / 105.0) * s[6];
}

__device__ __forceinline__ double recon_ud7l(const double *s, double alpha)
{   // reconstruction.cpp:234-249
    double a0 = -3.0 / 420.0 * alpha - 3.0 / 840.0 * (1.0 - alpha);
    double a1 = 25.0 / 420.0 * alpha + 29.0 / 840.0 * (1.0 - alpha);
    double a2 = -101.0 / 420.0 * alpha - 139.0 / 840.0 * (1.0 - alpha);
    double a3 = 319.0 / 420.0 * alpha + 533.0 / 840.0 * (1.0 - alpha);
    double a4 = 214.0 / 420.0 * alpha + 533.0 / 840.0 * (1.0 - alpha);
    double a5 = -38.0 / 420.0 * alpha - 139.0 / 840.0 * (1.0 - alpha);
    double a6 = 4.0 / 420.0 * alpha + 29.0 / 840.0 * (1.0 - alpha);
    double a7 = (-3.0 / 840.0) * (1.0 - alpha);
    return a0 * s[0] + a1 * s[1] + a2 * s[2] + a3 * s[3]
           + a4 * s[4] + a5 * s[5] + a6 * s[6] + a7 * s[7];
}

template <int RECON>
__device__ __forceinline__ double recon(const DevParams &P, const double *s)
{
    if constexpr (RECON == R_WENO5) return recon_weno5(s);
    else if constexpr (RECON == R_LINEAR) return s[0];
    else if constexpr (RECON == R_MDCD) return recon_mdcd(s, P.mdcd_disp, P.mdcd_diss);
    else if constexpr (RECON == R_MDCD_HYBRID) return recon_mdcd_hybrid(s, P.mdcd_disp, P.mdcd_diss);
    else if constexpr (RECON == R_UD7) return recon_ud7(s);
    else return recon_ud7l(s, P.ud7l_alpha);
}

// ------------------------------------------------- central differences
// ViscousFlux.cpp:19-40 on values f(-r..r) read at stride `st` from `c`.

template <int VIS>
__device__ __forceinline__ double diff_central(const double *__restrict__ c, long st, double h)
{
    if constexpr (VIS == VIS_C8) {
        double c1 = 0.8;
        double c2 = -0.2;
        double c3 = 0.0380952380952381;
        double c4 = -0.00357142857142857;
        return (c1 * (c[st] - c[-st]) +
                c2 * (c[2 * st] - c[-2 * st]) +
                c3 * (c[3 * st] - c[-3 * st]) +
                c4 * (c[4 * st] - c[-4 * st])) / h;
    } else if constexpr (VIS == VIS_C6) {
        return (c[3 * st] - 9.0 * c[2 * st] + 45.0 * c[st] - 45.0 * c[-st] + 9.0 * c[-2 * st] - c[-3 * st]) /
               (60.0 * h);
    } else {
        return (-c[2 * st] + 8.0 * c[st] - 8.0 * c[-st] + c[-2 * st]) / (12.0 * h);
    }
}

template <int VIS> struct VisReach;
template <> struct VisReach<VIS_C8> { static constexpr int R = 4; };
template <> struct VisReach<VIS_C6> { static constexpr int R = 3; };
template <> struct VisReach<VIS_C4> { static constexpr int R = 2; };

// Same operator on values gathered into v[-R..R] (v points at the centre).
template <int VIS>
__device__ __forceinline__ double diff_window(const double *v, double h)
{
    if constexpr (VIS == VIS_C8) {
        double c1 = 0.8;
        double c2 = -0.2;
        double c3 = 0.0380952380952381;
        double c4 = -0.00357142857142857;
        return (c1 * (v[1] - v[-1]) +
                c2 * (v[2] - v[-2]) +
                c3 * (v[3] - v[-3]) +
                c4 * (v[4] - v[-4])) / h;
    } else if constexpr (VIS == VIS_C6) {
        return (v[3] - 9.0 * v[2] + 45.0 * v[1] - 45.0 * v[-1] + 9.0 * v[-2] - v[-3]) / (60.0 * h);
    } else {
        return (-v[2] + 8.0 * v[1] - 8.0 * v[-1] + v[-2]) / (12.0 * h);
    }
}

// ---------------------------------------- closures on non-periodic axes
// choose_scheme_C4/C6/C8 (ViscousFlux.cpp:43-118) on an axis without a
// neighbour, for a cell dlo cells above the first and dhi cells below the
// last interior cell: min(dlo, dhi) = 3, 2, 1 take the 6th, 4th and 2nd
// order central formulas (ViscousFlux.cpp:19-29), the end cells the one-sided
// second-order formula of ViscousFlux.cpp:15-17 with flag +1 (low end) and
// -1 (high end).  With flag -1 that formula, (-3 f0 + 4 f[-1] - f[-2])/(2h),
// equals minus the derivative; it is the reference's expression and is kept.
// Only interior values are read.  Periodic axes pass dlo = dhi = R.
__device__ __forceinline__ double diff_closure(const double *v, double h, int dlo, int dhi)
{
    if (dlo == 0) return (-3.0 * v[0] + 4.0 * v[1] - v[2]) / (2.0 * h);
    if (dhi == 0) return (-3.0 * v[0] + 4.0 * v[-1] - v[-2]) / (2.0 * h);
    const int d = dlo < dhi ? dlo : dhi;
    if (d == 1) return (v[1] - v[-1]) / (2.0 * h);
    if (d == 2) return diff_window<VIS_C4>(v, h);
    return diff_window<VIS_C6>(v, h);
}

template <int VIS>
__device__ __forceinline__ double diff_window_bc(const double *v, double h, int dlo, int dhi)
{
    constexpr int R = VisReach<VIS>::R;
    if (dlo >= R && dhi >= R) return diff_window<VIS>(v, h);
    return diff_closure(v, h, dlo, dhi);
}

// diff_central with the closures; the window is gathered from memory only
// near a non-periodic end (the ghost layers keep every read in bounds)
template <int VIS>
__device__ __forceinline__ double diff_central_bc(const double *__restrict__ c, long st, double h, int dlo, int dhi)
{
    constexpr int R = VisReach<VIS>::R;
    if (dlo >= R && dhi >= R) return diff_central<VIS>(c, st, h);
    double v[2 * R + 1];
#pragma unroll
    for (int o = -R; o <= R; ++o) v[o + R] = c[o * st];
    return diff_closure(v + R, h, dlo, dhi);
}

// distances of interior index q (padded numbering) to the ends of an axis of
// n cells; R, R on periodic axes
__device__ __forceinline__ void axis_dist(int periodic, int q, int ng, int n, int R, int &dlo, int &dhi)
{
    if (periodic) { dlo = R; dhi = R; }
    else { dlo = q - ng; dhi = ng + n - 1 - q; }
}

// BC kinds of the ghost fill (ns3d_bc_type values)
enum { BCK_PERIODIC = 0, BCK_WALL = 1, BCK_SYMMETRY = 2, BCK_OUTFLOW = 3, BCK_INFLOW = 4 };

// One ghost cell of face `face` (0..5 = xmin..zmax) over the full cross
// ranges, t in [0, ng * cross1 * cross2): the cell id, the id it copies
// (mirror for wall/symmetry, the adjacent interior cell for outflow) and the
// normal-velocity component (1..3).  boundary.cpp:58-182.
__device__ __forceinline__ void bc_cell(const DevParams &P, int face, int kind, int t, int &dst, int &src, int &normal)
{
    const int axis = face >> 1, hi = face & 1, ng = P.ng;
    const int na = axis == 0 ? P.sy : P.sx;     // fastest cross axis
    const int g = t % ng;
    const int rest = t / ng;
    const int a = rest % na;
    const int b = rest / na;
    const int s = axis == 0 ? P.sx : axis == 1 ? P.sy : P.sz;
    int q, qs;
    if (!hi) {
        q = g;
        qs = kind == BCK_OUTFLOW ? ng : 2 * ng - 1 - g;
    } else {
        q = s - ng + g;
        qs = kind == BCK_OUTFLOW ? s - ng - 1 : 2 * (s - ng) - 1 - q;
    }
    if (axis == 0) { dst = (b * P.sy + a) * P.sx + q; src = (b * P.sy + a) * P.sx + qs; }
    else if (axis == 1) { dst = (b * P.sy + q) * P.sx + a; src = (b * P.sy + qs) * P.sx + a; }
    else { dst = (q * P.sy + b) * P.sx + a; src = (qs * P.sy + b) * P.sx + a; }
    normal = 1 + axis;
}

// apply_inflow_bc's hard-wired state (boundary.cpp:188): rho, u, v, w, p
__device__ __forceinline__ void inflow_state(double s[5]) { s[0] = 1.0; s[1] = 1.0; s[2] = 0.0; s[3] = 0.0; s[4] = 1.0; }

}  // namespace ns3d
