// ns3d_ad.hpp -- differentiation helpers of the FP64 RHS: dual numbers for
// forward-mode (tangent) evaluation, scalar-generic copies of the split and
// reconstructions, and hand-written reverse-mode (vector-Jacobian) rules of
// the reconstructions and of the viscous flux.
//
// The generic functions restate split_state / recon_* / the viscous flux of
// ns3d_device.hpp for any scalar type T with the same formulas; with T =
// Dual<N> they carry N directional derivatives.  |x| is differentiated as
// sign(x) (0 at 0): the plain Steger-Warming and Lax-Friedrichs splits have
// a kink where an eigenvalue vanishes, which the epsilon-modified split
// (sw_eps > 0) removes.
#pragma once

#include "ns3d_device.hpp"

namespace ns3d {

template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(double c)
{
    Dual<N> r;
    r.v = c;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a)
{
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double c) { Dual<N> r = a; r.v += c; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator+(double c, const Dual<N> &a) { return a + c; }
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double c) { Dual<N> r = a; r.v -= c; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator-(double c, const Dual<N> &a) { return -a + c; }
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double c)
{
    Dual<N> r;
    r.v = a.v * c;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * c;
    return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(double c, const Dual<N> &a) { return a * c; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double c) { return a * (1.0 / c); }
template <int N>
__device__ __forceinline__ Dual<N> operator/(double c, const Dual<N> &b) { return dconst<N>(c) / b; }

template <int N>
__device__ __forceinline__ Dual<N> sqrt_t(const Dual<N> &a)
{
    Dual<N> r;
    r.v = sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fabs_t(const Dual<N> &a)
{
    const double s = a.v > 0.0 ? 1.0 : (a.v < 0.0 ? -1.0 : 0.0);
    Dual<N> r;
    r.v = fabs(a.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> pow_t(const Dual<N> &a, double e)
{
    Dual<N> r;
    r.v = pow(a.v, e);
    const double g = a.v != 0.0 ? e * r.v / a.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * g;
    return r;
}
__device__ __forceinline__ double sqrt_t(double a) { return sqrt(a); }
__device__ __forceinline__ double fabs_t(double a) { return fabs(a); }
__device__ __forceinline__ double pow_t(double a, double e) { return pow(a, e); }
__device__ __forceinline__ double value_of(double a) { return a; }
template <int N> __device__ __forceinline__ double value_of(const Dual<N> &a) { return a.v; }

// ------------------------------------------------------ generic forward

template <class T>
__device__ __forceinline__ void lam_split_t(T lam, double eps, T &lp, T &lm)
{
    if (eps == 0.0) {
        lp = 0.5 * (lam + fabs_t(lam));
        lm = 0.5 * (lam - fabs_t(lam));
    } else {
        T r = sqrt_t(lam * lam + eps * eps);
        lp = 0.5 * (lam + r);
        lm = 0.5 * (lam - r);
    }
}

// split_state (InviscousFlux_new.cpp:439-530) for a generic scalar
template <int FVS, int DIR, class T>
__device__ __forceinline__ void split_t(const DevParams &P, T rho, T u, T v, T w, T p, T Fp[5], T Fm[5])
{
    const double gamma = P.gamma;
    T kinetic = 0.5 * (u * u + v * v + w * w);
    T e_internal = p / ((gamma - 1.0) * rho);
    T E = rho * (e_internal + kinetic);
    T c = sqrt_t(gamma * p / rho);
    T un = DIR == 0 ? u : (DIR == 1 ? v : w);
    if constexpr (FVS == FVS_SW) {
        T l1p, l1m, l2p, l2m, l3p, l3m;
        lam_split_t(un - c, P.sw_eps, l1p, l1m);
        lam_split_t(un, P.sw_eps, l2p, l2m);
        lam_split_t(un + c, P.sw_eps, l3p, l3m);
        T S1_p = l1p + l3p, S1_m = l1m + l3m;
        T S2_p = l3p - l1p, S2_m = l3m - l1m;
        T H = (E + p) / rho;
        T ga1 = rho / (2.0 * gamma);
        const double ga2 = 2.0 * (gamma - 1.0);
        T V = 0.5 * (u * u + v * v + w * w);
        T cp = S2_p * c, cm = S2_m * c;
        Fp[0] = ga1 * (S1_p + ga2 * l2p);
        Fp[1] = ga1 * (S1_p * u + ga2 * l2p * u + (DIR == 0 ? cp : cp * 0.0));
        Fp[2] = ga1 * (S1_p * v + ga2 * l2p * v + (DIR == 1 ? cp : cp * 0.0));
        Fp[3] = ga1 * (S1_p * w + ga2 * l2p * w + (DIR == 2 ? cp : cp * 0.0));
        Fp[4] = ga1 * (S1_p * H + ga2 * l2p * V + cp * un);
        Fm[0] = ga1 * (S1_m + ga2 * l2m);
        Fm[1] = ga1 * (S1_m * u + ga2 * l2m * u + (DIR == 0 ? cm : cm * 0.0));
        Fm[2] = ga1 * (S1_m * v + ga2 * l2m * v + (DIR == 1 ? cm : cm * 0.0));
        Fm[3] = ga1 * (S1_m * w + ga2 * l2m * w + (DIR == 2 ? cm : cm * 0.0));
        Fm[4] = ga1 * (S1_m * H + ga2 * l2m * V + cm * un);
    } else {
        T a_max = fabs_t(un) + c;
        T F1 = rho * un;
        T F2 = rho * u * un + (DIR == 0 ? p : p * 0.0);
        T F3 = rho * v * un + (DIR == 1 ? p : p * 0.0);
        T F4 = rho * w * un + (DIR == 2 ? p : p * 0.0);
        T F5 = (E + p) * un;
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

// computeRoeAveragedState_new and build_eigen_matrices_new
// (InviscousFlux_new.cpp:532-659) for a generic scalar: the Roe state of the
// primitive states ul, ur of a face with unit normal (nx, ny, nz), the left
// (Lm) and right (Rm) eigenvector matrices of the characteristic
// reconstruction, with the tangential pair of the face kernels
template <class T>
__device__ __forceinline__ void char_matrices_t(double gamma, const T ul[5], const T ur[5], double nx, double ny,
                                                double nz, T Lm[5][5], T Rm[5][5])
{
    const T E_L = ul[0] * (0.5 * (ul[1] * ul[1] + ul[2] * ul[2] + ul[3] * ul[3]) + ul[4] / ((gamma - 1.0) * ul[0]));
    const T E_R = ur[0] * (0.5 * (ur[1] * ur[1] + ur[2] * ur[2] + ur[3] * ur[3]) + ur[4] / ((gamma - 1.0) * ur[0]));
    const T sl = sqrt_t(ul[0]), sr = sqrt_t(ur[0]);
    const T den = 1.0 / (sl + sr);
    const T u0 = (sl * ul[1] + sr * ur[1]) * den;
    const T v0 = (sl * ul[2] + sr * ur[2]) * den;
    const T w0 = (sl * ul[3] + sr * ur[3]) * den;
    const T h0 = (sl * ((E_L + ul[4]) / ul[0]) + sr * ((E_R + ur[4]) / ur[0])) * den;
    const T V = 0.5 * (u0 * u0 + v0 * v0 + w0 * w0);
    const T un = u0 * nx + v0 * ny + w0 * nz;
    const T c = sqrt_t((gamma - 1.0) * (h0 - V));
    const T one = 1.0 + 0.0 * c, zero = 0.0 * c;
    Rm[0][0] = one; Rm[1][0] = u0 - c * nx; Rm[2][0] = v0 - c * ny; Rm[3][0] = w0 - c * nz; Rm[4][0] = h0 - un * c;
    Rm[0][1] = one; Rm[1][1] = u0; Rm[2][1] = v0; Rm[3][1] = w0; Rm[4][1] = V;
    Rm[0][2] = zero; Rm[1][2] = zero + nz; Rm[2][2] = zero + nx; Rm[3][2] = zero + ny;
    Rm[4][2] = u0 * nz + v0 * nx + w0 * ny;
    Rm[0][3] = zero; Rm[1][3] = zero + ny; Rm[2][3] = zero + nz; Rm[3][3] = zero + nx;
    Rm[4][3] = u0 * ny + v0 * nz + w0 * nx;
    Rm[0][4] = one; Rm[1][4] = u0 + c * nx; Rm[2][4] = v0 + c * ny; Rm[3][4] = w0 + c * nz; Rm[4][4] = h0 + un * c;
    const T chi = 0.5 / c;
    const T K = (gamma - 1.0) / (c * c);
    const T K1 = K * 0.5;
    Lm[0][0] = K1 * V + chi * un; Lm[0][1] = -(chi * nx) - K1 * u0; Lm[0][2] = -(chi * ny) - K1 * v0;
    Lm[0][3] = -(chi * nz) - K1 * w0; Lm[0][4] = K1;
    Lm[1][0] = 1.0 - K * V; Lm[1][1] = K * u0; Lm[1][2] = K * v0; Lm[1][3] = K * w0; Lm[1][4] = -K;
    Lm[2][0] = -(u0 * nz + v0 * nx + w0 * ny); Lm[2][1] = zero + nz; Lm[2][2] = zero + nx; Lm[2][3] = zero + ny;
    Lm[2][4] = zero;
    Lm[3][0] = -(u0 * ny + v0 * nz + w0 * nx); Lm[3][1] = zero + ny; Lm[3][2] = zero + nz; Lm[3][3] = zero + nx;
    Lm[3][4] = zero;
    Lm[4][0] = K1 * V - chi * un; Lm[4][1] = chi * nx - K1 * u0; Lm[4][2] = chi * ny - K1 * v0;
    Lm[4][3] = chi * nz - K1 * w0; Lm[4][4] = K1;
}

template <class T>
__device__ __forceinline__ T recon_weno5_t(const T *s)
{
    T f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
    const double eps = 1e-6;
    T d0 = f0 - 2.0 * f1 + f2, e0 = f0 - 4.0 * f1 + 3.0 * f2;
    T d1 = f1 - 2.0 * f2 + f3, e1 = f1 - f3;
    T d2 = f2 - 2.0 * f3 + f4, e2 = 3.0 * f2 - 4.0 * f3 + f4;
    T b0 = (13.0 / 12.0) * d0 * d0 + 0.25 * e0 * e0;
    T b1 = (13.0 / 12.0) * d1 * d1 + 0.25 * e1 * e1;
    T b2 = (13.0 / 12.0) * d2 * d2 + 0.25 * e2 * e2;
    T a0 = 0.1 / ((eps + b0) * (eps + b0));
    T a1 = 0.6 / ((eps + b1) * (eps + b1));
    T a2 = 0.3 / ((eps + b2) * (eps + b2));
    T as = a0 + a1 + a2;
    T p0 = (2.0 * f0 - 7.0 * f1 + 11.0 * f2) / 6.0;
    T p1 = (-f1 + 5.0 * f2 + 2.0 * f3) / 6.0;
    T p2 = (2.0 * f2 + 5.0 * f3 - f4) / 6.0;
    return (a0 * p0 + a1 * p1 + a2 * p2) / as;
}

// linear reconstruction coefficients (reconstruction.cpp:88-90, 126-135,
// 227-249), stencil order as reconstruct_select passes it
template <int RECON>
__device__ __forceinline__ void linear_coeffs(const DevParams &P, double *c)
{
    if constexpr (RECON == R_LINEAR) {
        c[0] = 1.0; c[1] = 0.0;
    } else if constexpr (RECON == R_MDCD) {
        const double disp = P.mdcd_disp, diss = P.mdcd_diss;
        c[0] = 1.0 / 2.0 * disp + 1.0 / 2.0 * diss;
        c[1] = -3.0 / 2.0 * disp - 5.0 / 2.0 * diss - 1.0 / 12.0;
        c[2] = 7.0 / 12.0 + disp + 5.0 * diss;
        c[3] = 7.0 / 12.0 + disp - 5.0 * diss;
        c[4] = -3.0 / 2.0 * disp + 5.0 / 2.0 * diss - 1.0 / 12.0;
        c[5] = 1.0 / 2.0 * disp - 1.0 / 2.0 * diss;
    } else if constexpr (RECON == R_UD7) {
        c[0] = -1.0 / 140.0; c[1] = 5.0 / 84.0; c[2] = -101.0 / 420.0; c[3] = 319.0 / 420.0;
        c[4] = 107.0 / 210.0; c[5] = -19.0 / 210.0; c[6] = 1.0 / 105.0; c[7] = 0.0;
    } else if constexpr (RECON == R_UD7L) {
        const double alpha = P.ud7l_alpha;
        c[0] = -3.0 / 420.0 * alpha - 3.0 / 840.0 * (1.0 - alpha);
        c[1] = 25.0 / 420.0 * alpha + 29.0 / 840.0 * (1.0 - alpha);
        c[2] = -101.0 / 420.0 * alpha - 139.0 / 840.0 * (1.0 - alpha);
        c[3] = 319.0 / 420.0 * alpha + 533.0 / 840.0 * (1.0 - alpha);
        c[4] = 214.0 / 420.0 * alpha + 533.0 / 840.0 * (1.0 - alpha);
        c[5] = -38.0 / 420.0 * alpha - 139.0 / 840.0 * (1.0 - alpha);
        c[6] = 4.0 / 420.0 * alpha + 29.0 / 840.0 * (1.0 - alpha);
        c[7] = (-3.0 / 840.0) * (1.0 - alpha);
    }
}

// mdcd_hybrid_reconstruction (reconstruction.cpp:151-225, as the oracle and
// recon_mdcd_hybrid restate it) for a generic scalar: the shock sensor
// chooses the branch on values (its derivative is zero where it is
// defined), |x| is differentiated as sign(x)
template <class T>
__device__ __forceinline__ T recon_mdcd_hybrid_t(const DevParams &P, const T *s)
{
    const double disp = P.mdcd_disp, diss = P.mdcd_diss;
    const double eps_small = 1e-40;
    const double eps = 0.9 * 0.4 / (1.0 - 0.9 * 0.4) * 1e-4;
    const T f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4], f5 = s[5];
    {
        const double g0 = value_of(f0), g1 = value_of(f1), g2 = value_of(f2);
        const double g3 = value_of(f3), g4 = value_of(f4), g5 = value_of(f5);
        const double a1 = fabs(g2 - g1) + fabs(g2 - 2 * g1 + g0);
        const double b1 = fabs(g2 - g3) + fabs(g2 - 2 * g3 + g4);
        const double a2 = fabs(g3 - g2) + fabs(g3 - 2 * g2 + g1);
        const double b2 = fabs(g3 - g4) + fabs(g3 - 2 * g4 + g5);
        const double sai1 = (2 * a1 * b1 + eps) / (a1 * a1 + b1 * b1 + eps);
        const double sai2 = (2 * a2 * b2 + eps) / (a2 * a2 + b2 * b2 + eps);
        const double sai = (sai2 < sai1) ? sai2 : sai1;
        if (sai > 0.4) {
            double c[6];
            c[0] = 1.0 / 2.0 * disp + 1.0 / 2.0 * diss;
            c[1] = -3.0 / 2.0 * disp - 5.0 / 2.0 * diss - 1.0 / 12.0;
            c[2] = 7.0 / 12.0 + disp + 5.0 * diss;
            c[3] = 7.0 / 12.0 + disp - 5.0 * diss;
            c[4] = -3.0 / 2.0 * disp + 5.0 / 2.0 * diss - 1.0 / 12.0;
            c[5] = 1.0 / 2.0 * disp - 1.0 / 2.0 * diss;
            T r = s[0] * c[0];
#pragma unroll
            for (int q = 1; q < 6; ++q) r = r + s[q] * c[q];
            return r;
        }
    }
    const T q0 = (2.0 * f0 - 7.0 * f1 + 11.0 * f2) / 6.0;
    const T q1 = (-f1 + 5.0 * f2 + 2.0 * f3) / 6.0;
    const T q2 = (2.0 * f2 + 5.0 * f3 - f4) / 6.0;
    const T q3 = (11.0 * f3 - 7.0 * f4 + 2.0 * f5) / 6.0;
    const double d0 = 1.5 * (disp + diss);
    const double d1 = 0.5 - 1.5 * (disp - 3 * diss);
    const double d2 = 0.5 - 1.5 * (disp + 3 * diss);
    const double d3 = 1.5 * (disp - diss);
    const T t0 = f0 - 2.0 * f1 + f2, t0b = f0 - 4.0 * f1 + 3.0 * f2;
    const T t1 = f1 - 2.0 * f2 + f3, t1b = f1 - f3;
    const T t2 = f2 - 2.0 * f3 + f4, t2b = 3.0 * f2 - 4.0 * f3 + f4;
    const T beta0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (t0b * t0b);
    const T beta1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (t1b * t1b);
    const T beta2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (t2b * t2b);
    const T beta3 = (271799.0 * (f0 * f0) +
                     f0 * (-2380800.0 * f1 + 4086352.0 * f2 - 3462252.0 * f3 + 1458762.0 * f4 - 245620.0 * f5) +
                     f1 * (5653317.0 * f1 - 20427884.0 * f2 + 17905032.0 * f3 - 7727988.0 * f4 + 1325006.0 * f5) +
                     f2 * (19510972.0 * f2 - 35817664.0 * f3 + 15929912.0 * f4 - 2792660.0 * f5) +
                     f3 * (17195652.0 * f3 - 15880404.0 * f4 + 2863984.0 * f5) +
                     f4 * (3824847.0 * f4 - 1429976.0 * f5) + 139633.0 * (f5 * f5)) /
                    120960.0;
    const T tau6 = fabs_t(beta3 - (beta0 + 4.0 * beta1 + beta2) / 6.0);
    const T g0 = 20.0 + tau6 / (beta0 + eps_small);
    const T g1 = 20.0 + tau6 / (beta1 + eps_small);
    const T g2 = 20.0 + tau6 / (beta2 + eps_small);
    const T g3 = 20.0 + tau6 / (beta3 + eps_small);
    const T aa0 = d0 * (g0 * g0), aa1 = d1 * (g1 * g1), aa2 = d2 * (g2 * g2), aa3 = d3 * (g3 * g3);
    return (aa0 * q0 + aa1 * q1 + aa2 * q2 + aa3 * q3) / (aa0 + aa1 + aa2 + aa3);
}

template <int RECON, class T>
__device__ __forceinline__ T recon_t(const DevParams &P, const T *s)
{
    if constexpr (RECON == R_WENO5) {
        return recon_weno5_t(s);
    } else if constexpr (RECON == R_MDCD_HYBRID) {
        return recon_mdcd_hybrid_t(P, s);
    } else {
        constexpr int S = Stencil<RECON>::S;
        double c[S];
        linear_coeffs<RECON>(P, c);
        T r = s[0] * c[0];
#pragma unroll
        for (int q = 1; q < S; ++q) r = r + s[q] * c[q];
        return r;
    }
}

// ------------------------------------------------------------ reverse rules


// sbar[s] = gbar * d recon / d s[s]
template <int RECON>
__device__ __forceinline__ void recon_vjp(const DevParams &P, const double *s, double gbar, double *sbar)
{
    constexpr int S = Stencil<RECON>::S;
    if constexpr (RECON == R_WENO5) {
        const double f0 = s[0], f1 = s[1], f2 = s[2], f3 = s[3], f4 = s[4];
        const double eps = 1e-6;
        const double d0 = f0 - 2.0 * f1 + f2, e0 = f0 - 4.0 * f1 + 3.0 * f2;
        const double d1 = f1 - 2.0 * f2 + f3, e1 = f1 - f3;
        const double d2 = f2 - 2.0 * f3 + f4, e2 = 3.0 * f2 - 4.0 * f3 + f4;
        const double c13 = 13.0 / 12.0, c16 = 1.0 / 6.0;
        const double it0 = rcp64(eps + c13 * d0 * d0 + 0.25 * e0 * e0);
        const double it1 = rcp64(eps + c13 * d1 * d1 + 0.25 * e1 * e1);
        const double it2 = rcp64(eps + c13 * d2 * d2 + 0.25 * e2 * e2);
        const double a0 = 0.1 * it0 * it0, a1 = 0.6 * it1 * it1, a2 = 0.3 * it2 * it2;
        const double ias = rcp64(a0 + a1 + a2);
        const double w0 = a0 * ias, w1 = a1 * ias, w2 = a2 * ias;
        const double p0 = (2.0 * f0 - 7.0 * f1 + 11.0 * f2) * c16;
        const double p1 = (-f1 + 5.0 * f2 + 2.0 * f3) * c16;
        const double p2 = (2.0 * f2 + 5.0 * f3 - f4) * c16;
        const double val = w0 * p0 + w1 * p1 + w2 * p2;
        // d val / d p_k = w_k ; d val / d a_k = (p_k - val) / as ;
        // d a_k / d beta_k = -2 a_k / (eps + beta_k)
        const double m2g = -2.0 * gbar;
        const double ab0 = m2g * w0 * (p0 - val) * it0;
        const double ab1 = m2g * w1 * (p1 - val) * it1;
        const double ab2 = m2g * w2 * (p2 - val) * it2;
        const double gd0 = ab0 * (13.0 / 6.0) * d0, ge0 = ab0 * 0.5 * e0;
        const double gd1 = ab1 * (13.0 / 6.0) * d1, ge1 = ab1 * 0.5 * e1;
        const double gd2 = ab2 * (13.0 / 6.0) * d2, ge2 = ab2 * 0.5 * e2;
        const double g0 = gbar * w0 * c16, g1 = gbar * w1 * c16, g2 = gbar * w2 * c16;
        sbar[0] = 2.0 * g0 + gd0 + ge0;
        sbar[1] = -7.0 * g0 - g1 - 2.0 * gd0 - 4.0 * ge0 + gd1 + ge1;
        sbar[2] = 11.0 * g0 + 5.0 * g1 + 2.0 * g2 + gd0 + 3.0 * ge0 - 2.0 * gd1 + gd2 + 3.0 * ge2;
        sbar[3] = 2.0 * g1 + 5.0 * g2 + gd1 - ge1 - 2.0 * gd2 - 4.0 * ge2;
        sbar[4] = -g2 + gd2 + ge2;
        sbar[5] = 0.0;
    } else if constexpr (RECON == R_MDCD_HYBRID) {
        // the local gradient of the 6-point formula by forward mode, seeded
        // with the identity, then scaled by gbar
        Dual<6> d[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            d[q] = dconst<6>(s[q]);
            d[q].d[q] = 1.0;
        }
        const Dual<6> r = recon_mdcd_hybrid_t(P, d);
#pragma unroll
        for (int q = 0; q < 6; ++q) sbar[q] = gbar * r.d[q];
    } else {
        double c[S];
        linear_coeffs<RECON>(P, c);
#pragma unroll
        for (int q = 0; q < S; ++q) sbar[q] = gbar * c[q];
    }
}


// ------------------------------------------- one flux component per lane

// Component m of the split fluxes F+ and F- (split_state) without branching
// on m: F_m = ga1 (S1 q_m + ga2 l2 r_m + S2 c s_m) for Steger-Warming with
// q = (1,u,v,w,H), r = (1,u,v,w,V), s = (0,n,un); Lax-Friedrichs
// F_m = (Fc_m +- a U_m) / 2 with Fc_m = U_m un + p s_m.
template <int FVS, int DIR>
__device__ __forceinline__ void split_comp(const DevParams &P, double rho, double u, double v, double w, double p,
                                           int m, double &fp, double &fm)
{
    const double gamma = P.gamma;
    const double irho = rcp64(rho);
    const double V = 0.5 * (u * u + v * v + w * w);
    const double c2 = gamma * p * irho;
    const double c = sqrt(c2);
    const double un = DIR == 0 ? u : (DIR == 1 ? v : w);
    const double um = m == 1 ? u : (m == 2 ? v : w);
    const double sm = m == 4 ? un : (m == DIR + 1 ? 1.0 : 0.0);
    const double ig1 = 1.0 / (gamma - 1.0);   // uniform
    if constexpr (FVS == FVS_SW) {
        double l1p, l1m, l2p, l2m, l3p, l3m;
        lam_split(un - c, P.sw_eps, l1p, l1m);
        lam_split(un, P.sw_eps, l2p, l2m);
        lam_split(un + c, P.sw_eps, l3p, l3m);
        const double H = c2 * ig1 + V;
        const double q = m == 0 ? 1.0 : (m == 4 ? H : um);
        const double r = m == 0 ? 1.0 : (m == 4 ? V : um);
        const double ga1 = rho * (0.5 / gamma), ga2 = 2.0 * (gamma - 1.0);
        fp = ga1 * ((l1p + l3p) * q + ga2 * l2p * r + (l3p - l1p) * c * sm);
        fm = ga1 * ((l1m + l3m) * q + ga2 * l2m * r + (l3m - l1m) * c * sm);
    } else {
        const double E = p * ig1 + rho * V;
        const double Um = m == 0 ? rho : (m == 4 ? E : rho * um);
        const double fc = Um * un + p * sm;
        const double a = fabs(un) + c;
        fp = 0.5 * (fc + a * Um);
        fm = 0.5 * (fc - a * Um);
    }
}

// d l+/d lambda and d l-/d lambda of lam_split (|x|' = sign(x), 0 at 0)
__device__ __forceinline__ void lam_split_d(double lam, double eps, double &dp, double &dm)
{
    double t;
    if (eps == 0.0) t = lam > 0.0 ? 1.0 : (lam < 0.0 ? -1.0 : 0.0);
    else t = lam / sqrt(lam * lam + eps * eps);
    dp = 0.5 * (1.0 + t);
    dm = 0.5 * (1.0 - t);
}

// Reverse rule of split_comp: g[0..4] = d(bp F+_m + bm F-_m)/d(rho,u,v,w,p).
template <int FVS, int DIR>
__device__ __forceinline__ void split_comp_vjp(const DevParams &P, double rho, double u, double v, double w,
                                               double p, int m, double bp, double bm, double g[5])
{
    const double gamma = P.gamma;
    const double irho = rcp64(rho);
    const double V = 0.5 * (u * u + v * v + w * w);
    const double c2 = gamma * p * irho;
    const double c = sqrt(c2);
    const double un = DIR == 0 ? u : (DIR == 1 ? v : w);
    const double um = m == 1 ? u : (m == 2 ? v : w);
    const double sm = m == 4 ? un : (m == DIR + 1 ? 1.0 : 0.0);
    const double ig1 = 1.0 / (gamma - 1.0);   // uniform
    double rb = 0.0, pb = 0.0, cb = 0.0, unb = 0.0, Vb = 0.0, umb = 0.0;
    if constexpr (FVS == FVS_SW) {
        double l1p, l1m, l2p, l2m, l3p, l3m;
        lam_split(un - c, P.sw_eps, l1p, l1m);
        lam_split(un, P.sw_eps, l2p, l2m);
        lam_split(un + c, P.sw_eps, l3p, l3m);
        const double H = c2 * ig1 + V;
        const double q = m == 0 ? 1.0 : (m == 4 ? H : um);
        const double r = m == 0 ? 1.0 : (m == 4 ? V : um);
        const double i2g = 0.5 / gamma;   // uniform
        const double ga1 = rho * i2g, ga2 = 2.0 * (gamma - 1.0);
        const double Pp = (l1p + l3p) * q + ga2 * l2p * r + (l3p - l1p) * c * sm;
        const double Pm = (l1m + l3m) * q + ga2 * l2m * r + (l3m - l1m) * c * sm;
        rb = (bp * Pp + bm * Pm) * i2g;
        const double ap = bp * ga1, am = bm * ga1;
        // through P+ and P-
        const double S1b_p = ap * q, S2b_p = ap * c * sm, S1b_m = am * q, S2b_m = am * c * sm;
        const double l2b_p = ap * ga2 * r, l2b_m = am * ga2 * r;
        const double qb = ap * (l1p + l3p) + am * (l1m + l3m);
        const double rbar = (ap * l2p + am * l2m) * ga2;
        const double sb = ap * (l3p - l1p) * c + am * (l3m - l1m) * c;
        cb += ap * (l3p - l1p) * sm + am * (l3m - l1m) * sm;
        const double l1b_p = S1b_p - S2b_p, l3b_p = S1b_p + S2b_p;
        const double l1b_m = S1b_m - S2b_m, l3b_m = S1b_m + S2b_m;
        double d1p, d1m, d2p, d2m, d3p, d3m;
        lam_split_d(un - c, P.sw_eps, d1p, d1m);
        lam_split_d(un, P.sw_eps, d2p, d2m);
        lam_split_d(un + c, P.sw_eps, d3p, d3m);
        const double lam1b = l1b_p * d1p + l1b_m * d1m;
        const double lam2b = l2b_p * d2p + l2b_m * d2m;
        const double lam3b = l3b_p * d3p + l3b_m * d3m;
        unb += lam1b + lam2b + lam3b;
        cb += lam3b - lam1b;
        // q = H (m = 4) or u_m (m = 1..3); r = V (m = 4) or u_m
        if (m == 4) {
            cb += qb * 2.0 * c * ig1;
            Vb += qb + rbar;
            unb += sb;
        } else if (m > 0) {
            umb += qb + rbar;
        }
    } else {
        const double E = p * ig1 + rho * V;
        const double Um = m == 0 ? rho : (m == 4 ? E : rho * um);
        const double a = fabs(un) + c;
        const double beta = 0.5 * (bp + bm), delta = 0.5 * (bp - bm);
        const double Ub = beta * un + delta * a;
        unb += beta * Um;
        pb += beta * sm;
        if (m == 4) unb += beta * p;
        const double ab = delta * Um;
        unb += ab * (un > 0.0 ? 1.0 : (un < 0.0 ? -1.0 : 0.0));
        cb += ab;
        if (m == 0) {
            rb += Ub;
        } else if (m == 4) {
            pb += Ub * ig1;
            rb += Ub * V;
            Vb += Ub * rho;
        } else {
            rb += Ub * um;
            umb += Ub * rho;
        }
    }
    // c = sqrt(gamma p / rho): dc/dp = gamma / (2 rho c), dc/drho = -c / (2 rho)
    const double hc = 0.5 * cb * rcp64(c);
    pb += hc * gamma * irho;
    rb -= hc * c2 * irho;
    g[0] = rb;
    g[1] = Vb * u + (m == 1 ? umb : 0.0) + (DIR == 0 ? unb : 0.0);
    g[2] = Vb * v + (m == 2 ? umb : 0.0) + (DIR == 1 ? unb : 0.0);
    g[3] = Vb * w + (m == 3 ? umb : 0.0) + (DIR == 2 ? unb : 0.0);
    g[4] = pb;
}

}  // namespace ns3d
