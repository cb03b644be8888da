// ns3d_monitors.hip -- device-side monitors: the energy spectrum
// (EnergySpectrum.cpp:62-154) through a hipFFT real-to-complex transform with
// the reference's shell binning, the sums of compute_turbulence_statistics
// (ana_isotropic.cpp:226-305) and of compute_diagnostics
// (postprocess.cpp:7-96).  The reductions write per-block partials that the
// host adds in block order; the spectrum's per-block shell sums use LDS
// atomics, so their last bits depend on the order of arrival.
#include "ns3d_internal.hpp"

namespace ns3d {

__device__ __forceinline__ int mwrap(int q, int ng, int n)
{
    return q < ng ? q + n : (q >= ng + n ? q - n : q);
}

// interior of a primitive array -> contiguous (nz, ny, nx)
__global__ void km_pack(DevParams P, const double *src, double *dst)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)P.nx * P.ny * P.nz;
    if (t >= n) return;
    const int i = (int)(t % P.nx);
    const int j = (int)((t / P.nx) % P.ny);
    const int k = (int)(t / ((long long)P.nx * P.ny));
    dst[t] = src[((P.ng + k) * P.sy + (P.ng + j)) * P.sx + (P.ng + i)];
}

// Shell sums of 0.5 |U_hat|^2 / N^2 over the modes of one R2C transform
// (nz, ny, nx/2+1).  Modes 0 < kx < nx/2 stand for themselves and their
// conjugates (weight 2).  Shell = floor(|k|) with k_d = i <= n/2 ? i : i - n
// (EnergySpectrum.cpp:108-125).  Each block writes its nbins partials.
__global__ __launch_bounds__(256) void km_spec_bins(DevParams P, const double2 *U, double *partial, int nbins,
                                                    double scale, int accumulate)
{
    extern __shared__ double hist[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    const int nxh = P.nx / 2 + 1;
    const long long total = (long long)nxh * P.ny * P.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int kx = (int)(t % nxh);
        const int j = (int)((t / nxh) % P.ny);
        const int k = (int)(t / ((long long)nxh * P.ny));
        const int k2 = j <= P.ny / 2 ? j : j - P.ny;
        const int k3 = k <= P.nz / 2 ? k : k - P.nz;
        const int k1 = kx;   // kx <= nx/2 here
        const int shell = (int)sqrt((double)(k1 * k1 + k2 * k2 + k3 * k3));
        const double2 c = U[t];
        double e = 0.5 * (c.x * c.x + c.y * c.y) * scale;
        const bool twice = kx > 0 && 2 * kx != P.nx;
        if (twice) e *= 2.0;
        if (shell < nbins) atomicAdd(&hist[shell], e);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
        double *dst = partial + (long long)blockIdx.x * nbins + b;
        *dst = accumulate ? *dst + hist[b] : hist[b];
    }
}

// Sums of the turbulence statistics per block, in this order:
//   0 sum 0.5|u|^2   1 sum |u|^2   2 sum u   3 sum v   4 sum w
//   5 sum u^2   6 sum v^2   7 sum w^2   8 sum (du/dx)^2
//   9 sum sqrt(gamma R T)   10 sum mu(T)   11 sum rho
// du/dx with the viscous scheme of compute_gradients_dudx
// (ViscousFlux.cpp:389-411): periodic images (index wrap) on a periodic x
// axis, the closures near the ends otherwise.
template <int VIS>
__global__ __launch_bounds__(256) void km_stats(DevParams P, const double *const *W, double *partial)
{
    constexpr int R = VisReach<VIS>::R;
    __shared__ double red[12][4];
    const long long n = (long long)P.nx * P.ny * P.nz;
    double s[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) s[q] = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = P.ng + (int)(t % P.nx);
        const int j = P.ng + (int)((t / P.nx) % P.ny);
        const int k = P.ng + (int)(t / ((long long)P.nx * P.ny));
        const int row = (k * P.sy + j) * P.sx;
        const int id = row + i;
        const double rho = W[0][id], u = W[1][id], v = W[2][id], w = W[3][id], T = W[5][id];
        double win[2 * R + 1];
#pragma unroll
        for (int o = -R; o <= R; ++o) win[o + R] = W[1][row + mwrap(i + o, P.ng, P.nx)];
        int dlo, dhi;
        axis_dist(P.periodic[0], i, P.ng, P.nx, R, dlo, dhi);
        const double dudx = diff_window_bc<VIS>(win + R, P.dx, dlo, dhi);
        const double q2 = u * u + v * v + w * w;
        s[0] += 0.5 * q2;
        s[1] += q2;
        s[2] += u; s[3] += v; s[4] += w;
        s[5] += u * u; s[6] += v * v; s[7] += w * w;
        s[8] += dudx * dudx;
        s[9] += sqrt(P.gamma * P.Rgas * T);
        s[10] += get_mu(P, T);
        s[11] += rho;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        double x = s[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int q = threadIdx.x;
        partial[(long long)blockIdx.x * 12 + q] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
    }
}

// Residual sums per block (postprocess.cpp:7-96 after updateResiduals):
//   0..4  sum (U_m - U0_m)^2 dx^3      5..9  max |U_m|
// FP64 mode: U from P2C of the primitives (field_structures.h:430-451) and
// U0 the state at the start of the last step.  Mixed mode (conserved != 0):
// U from the conserved arrays and U - U0 the last step's FP32 increment dU.
__global__ __launch_bounds__(256) void km_resid(DevParams P, const double *const *Wc, const double *const *U0,
                                                const float *const *dU, int conserved, double *partial)
{
    __shared__ double red[10][4];
    const long long n = (long long)P.nx * P.ny * P.nz;
    const double dx3 = P.dx * P.dx * P.dx;
    double s[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) s[q] = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = P.ng + (int)(t % P.nx);
        const int j = P.ng + (int)((t / P.nx) % P.ny);
        const int k = P.ng + (int)(t / ((long long)P.nx * P.ny));
        const int id = (k * P.sy + j) * P.sx + i;
        double U[5];
        if (conserved) {
#pragma unroll
            for (int m = 0; m < 5; ++m) U[m] = Wc[m][id];
        } else {
            const double rr = Wc[0][id], u = Wc[1][id], v = Wc[2][id], w = Wc[3][id], p = Wc[4][id];
            double E, c;
            p2c_energy_sound(P.gamma, rr, u, v, w, p, E, c);
            U[0] = rr; U[1] = rr * u; U[2] = rr * v; U[3] = rr * w; U[4] = E;
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double r = conserved ? (double)dU[m][id] : U[m] - U0[m][id];
            s[m] += r * r * dx3;
            s[5 + m] = fmax(s[5 + m], fabs(U[m]));
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        double x = s[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_xor(x, o, 64);
            x = q < 5 ? x + y : fmax(x, y);
        }
        if (lane == 0) red[q][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        const int q = threadIdx.x;
        partial[(long long)blockIdx.x * 10 + q] =
            q < 5 ? (red[q][0] + red[q][1]) + (red[q][2] + red[q][3])
                  : fmax(fmax(red[q][0], red[q][1]), fmax(red[q][2], red[q][3]));
    }
}

// ------------------------------------------------------------- launches

int monitor_blocks() { return 1024; }

hipError_t launch_pack(hipStream_t s, const DevParams &P, const double *src, double *dst)
{
    const long long n = (long long)P.nx * P.ny * P.nz;
    hipLaunchKernelGGL(km_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, src, dst);
    return hipGetLastError();
}

hipError_t launch_spec_bins(hipStream_t s, const DevParams &P, const void *U, double *partial, int nbins,
                            double scale, int accumulate)
{
    hipLaunchKernelGGL(km_spec_bins, dim3(monitor_blocks()), dim3(256), (size_t)nbins * sizeof(double), s, P,
                       (const double2 *)U, partial, nbins, scale, accumulate);
    return hipGetLastError();
}

hipError_t launch_stats(hipStream_t s, const DevParams &P, const double *const *W, double *partial)
{
    switch (P.vis) {
    case VIS_C8: hipLaunchKernelGGL(km_stats<VIS_C8>, dim3(monitor_blocks()), dim3(256), 0, s, P, W, partial); break;
    case VIS_C6: hipLaunchKernelGGL(km_stats<VIS_C6>, dim3(monitor_blocks()), dim3(256), 0, s, P, W, partial); break;
    default: hipLaunchKernelGGL(km_stats<VIS_C4>, dim3(monitor_blocks()), dim3(256), 0, s, P, W, partial); break;
    }
    return hipGetLastError();
}

hipError_t launch_resid(hipStream_t s, const DevParams &P, const double *const *Wc, const double *const *U0,
                        const float *const *dU, int conserved, double *partial)
{
    hipLaunchKernelGGL(km_resid, dim3(monitor_blocks()), dim3(256), 0, s, P, Wc, U0, dU, conserved, partial);
    return hipGetLastError();
}

}  // namespace ns3d
