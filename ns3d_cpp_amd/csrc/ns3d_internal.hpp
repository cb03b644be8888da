// ns3d_internal.hpp -- types shared by the kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include "ns3d_device.hpp"

namespace ns3d {

// Device arrays of one context (all sx*sy*sz doubles, idx3 layout).
struct Fields {
    double *W[6];    // rho, u, v, w, p, T
    double *U0[5];   // conserved state at the start of the step
    double *A[5];    // inviscid accumulator, then RHS
    double *G[12];   // du_dx du_dy du_dz dv_* dw_* dT_*
    double *VF[12];  // Fv(momx,momy,momz,E) Hv(...) Gv(...)
};

// Face-flux arrays of du_inviscous: f[5*dir + m] (idx_fx / idx_fy / idx_fz).
struct FaceArrays {
    double *f[15];
};

hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n);
hipError_t launch_prim_setup(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_c2p_all(hipStream_t s, const DevParams &P, const Fields &F, const double *const *src);
hipError_t launch_p2c_all(hipStream_t s, const DevParams &P, const Fields &F, double *const *dst);
hipError_t launch_fill_ghosts(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_sweep(hipStream_t s, const DevParams &P, const Fields &F, int dir, int acc, int seg_len);
hipError_t launch_faces(hipStream_t s, const DevParams &P, const Fields &F, const FaceArrays &out,
                        int dir, int char_recon);
hipError_t launch_face_diff(hipStream_t s, const DevParams &P, const Fields &F, const FaceArrays &ff);
hipError_t launch_grad(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_vflux(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_vdiv(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt,
                       unsigned long long *flag, unsigned long long key);
hipError_t launch_cfl(hipStream_t s, const DevParams &P, const Fields &F, unsigned long long *out);
hipError_t launch_copy(hipStream_t s, const double *src, double *dst, long n);
int ke_blocks(const DevParams &P);
hipError_t launch_ke(hipStream_t s, const DevParams &P, const Fields &F, double *partial);

}  // namespace ns3d
