// ns3d_noalloc.hip -- compute_timestep (timestep.cpp:5-33) of caller-owned
// device arrays with no device allocation: the variant of
// ns3d_compute_timestep_local for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The reduction's scratch is the file-scope device array `scratch` below:
// one partial minimum per block of the first kernel, NOALLOC_MAX_BLOCKS
// doubles (16 KiB) per device, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (on the host it names a stand-in, not
// device memory): its device address is taken once per device with
// hipGetSymbolAddress, with that device current, and kept in scratch_dev[].
// Every block of the first kernel writes its slot and the second kernel
// reads only the slots written, so nothing needs clearing between calls.
// Calls that overlap on one device share the array and race.
#include <mutex>

#include "../../include/ns3d_gpu.h"
#include "ns3d_internal.hpp"

namespace ns3d {

constexpr int NOALLOC_MAX_BLOCKS = 2048;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_MAX_BLOCKS];

// per-block minimum of the CFL bound of every interior cell (the same cell
// bound and the same grid-stride traversal as k_cfl)
__global__ __launch_bounds__(256) void kn_cfl_partial(DevParams P, const double *rho, const double *u,
                                                       const double *v, const double *w, const double *p,
                                                       double *partial)
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
        m = fmin(m, cfl_cell(P, rho[id], u[id], v[id], w[id], p[id]));
    }
    m = wave_min(m);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}

// dt = cfl * (minimum of the nb partials): one block
__global__ __launch_bounds__(256) void kn_cfl_finish(const double *partial, int nb, double cfl, double *dt)
{
    __shared__ double red[4];
    double m = 1e9;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmin(m, partial[b]);
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) dt[0] = cfl * fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));   // timestep.cpp:32
}

namespace {

std::mutex scratch_mu;
double *scratch_dev[NOALLOC_MAX_DEVICES];

// the device address of `scratch` on the current device, looked up once per
// device; nullptr on failure
double *scratch_on(int dev)
{
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        double *ptr = nullptr;
        if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        scratch_dev[dev] = ptr;
    }
    return scratch_dev[dev];
}

}  // namespace
}  // namespace ns3d

using namespace ns3d;

extern "C" int ns3d_cfl_dt_noalloc(const ns3d_params *p, const ns3d_grid *g, const double *rho, const double *u,
                                   const double *v, const double *w, const double *pr, double *dt_dev, void *stream)
{
    if (!p || !g || !rho || !u || !v || !w || !pr || !dt_dev) return NS3D_ERR_INVALID_ARG;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->ng < 0) return NS3D_ERR_INVALID_ARG;
    if (!(g->Lx > 0.0) || !(g->Ly > 0.0) || !(g->Lz > 0.0) || !(p->gamma > 0.0)) return NS3D_ERR_INVALID_ARG;
    const long long sx = g->nx + 2LL * g->ng, sy = g->ny + 2LL * g->ng, sz = g->nz + 2LL * g->ng;
    // 32-bit cell indices, as in every other kernel of the library
    if (sx * sy * sz > 2147483647LL) return NS3D_ERR_INVALID_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return NS3D_ERR_HIP;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return NS3D_ERR_UNSUPPORTED;
    double *part = scratch_on(dev);
    if (!part) return NS3D_ERR_HIP;
    DevParams d{};
    d.gamma = p->gamma;
    d.cfl = p->cfl;
    d.dx = g->Lx / g->nx;   // initialization.cpp:324-326
    d.dy = g->Ly / g->ny;
    d.dz = g->Lz / g->nz;
    d.nx = g->nx; d.ny = g->ny; d.nz = g->nz; d.ng = g->ng;
    d.sx = (int)sx; d.sy = (int)sy; d.sz = (int)sz;
    const long long total = (long long)g->nx * g->ny * g->nz;
    long long nb = (total + 255) / 256;
    if (nb > NOALLOC_MAX_BLOCKS) nb = NOALLOC_MAX_BLOCKS;   // grid-stride: one slot per block
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(kn_cfl_partial, dim3((unsigned)nb), dim3(256), 0, s, d, rho, u, v, w, pr, part);
    if (hipGetLastError() != hipSuccess) return NS3D_ERR_HIP;
    hipLaunchKernelGGL(kn_cfl_finish, dim3(1), dim3(256), 0, s, (const double *)part, (int)nb, p->cfl, dt_dev);
    if (hipGetLastError() != hipSuccess) return NS3D_ERR_HIP;
    return NS3D_OK;
}
