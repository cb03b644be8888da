// ns3d_internal.hpp -- types shared by the kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include "ns3d_device.hpp"

namespace ns3d {

// Device arrays of one context (all sx*sy*sz doubles, idx3 layout).
struct Fields {
    double *W[6];    // rho, u, v, w, p, T
    double *U0[5];   // conserved state at the start of the step (the mixed
                     // mode keeps its FP32 stage increments MFields::dU here)
    double *A[5];    // inviscid accumulator, then RHS
    double *G[12];   // du_dx du_dy du_dz dv_* dw_* dT_*
    double *VF[12];  // Fv(momx,momy,momz,E) Hv(...) Gv(...)
};

// Device arrays of the mixed-precision path.
struct MFields {
    double *U[5];    // conserved rho, rho u, rho v, rho w, E (FP64) at the
                     // start of the step; the third stage writes the new state
    float *dU[5];    // RK3 increment of the stage state over U (FP32); after
                     // the third stage, the whole step's increment
    double *out[5];  // RHS of ns3d_compute_rhs (FP64)
    float *S[6];     // shadow: rho - rho_ref, u, v, w, p - p_ref, T - 1
    float *A[5];     // inviscid accumulator
    float *VF[12];   // viscous fluxes
};

// the second shadow buffer: the output of the stage kernel that folds in the
// z sweep (km_vdiv_zs), which reads the input shadow across its segments'
// ends and so cannot overwrite it
struct Shadow {
    float *S[6];
};

// Arrays of the tangent-linear and adjoint RHS (FP64, interior cells used).
//   TLM:     X = dU (in), Y = dW, T = face-flux tangents / dT, V = dVF, Z = dR (out)
//   adjoint: X = lambda (in), Y = lambda_W, G = lambda of the gradients,
//            C = RK3 accumulator of lambda_U0, Z = combined result (out)
struct AdjArrays {
    double *X[5];
    double *Y[5];
    double *Z[5];
    double *T[5];
    double *C[5];
    double *V[12];
    double *G[12];
};

// Per-face buffer of the face-wise inviscid adjoint: char_buf_arrays(P)
// double arrays of n cells each, array a at base + a * n.
struct CharBuf {
    double *base;
    long long n;
};

__device__ __forceinline__ int wrapi_m(int q, int ng, int n)
{
    return q < ng ? q + n : (q >= ng + n ? q - n : q);
}

// Face-flux arrays of du_inviscous: f[5*dir + m] (idx_fx / idx_fy / idx_fz).
struct FaceArrays {
    double *f[15];
};

// arrays of the batched split: states in, split fluxes out (n each)
struct SplitBatch {
    const double *w[5];
    double *fp[5];
    double *fm[5];
};

// six FP64 cell arrays (rho, u, v, w, p, T) passed by value
struct Prims64 {
    double *p[6];
};

hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n);
hipError_t launch_step_end(hipStream_t s, unsigned long long *flag, double cfl, int dev_dt);
hipError_t launch_prim_setup(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_c2p_all(hipStream_t s, const DevParams &P, const Fields &F, const double *const *src);
hipError_t launch_p2c_all(hipStream_t s, const DevParams &P, const Fields &F, double *const *dst);
hipError_t launch_fill_ghosts(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_sweep(hipStream_t s, const DevParams &P, const Fields &F, int dir, int seg_len, int minw);
bool xrow_supported(const DevParams &P);
hipError_t launch_sweep_xrow(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_gradflux(hipStream_t s, const DevParams &P, const Fields &F, int seg_len, int minb, int lds);
hipError_t launch_vdiv_z(hipStream_t s, const DevParams &P, const Fields &F, int mode, double dt, int seg,
                         unsigned long long *flag, int dev_dt, int lds);
hipError_t launch_faces(hipStream_t s, const DevParams &P, const Fields &F, const FaceArrays &out,
                        int dir, int char_recon);
hipError_t launch_face_diff(hipStream_t s, const DevParams &P, const Fields &F, const FaceArrays &ff);
hipError_t launch_grad(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_vflux(hipStream_t s, const DevParams &P, const Fields &F);
hipError_t launch_cfl(hipStream_t s, const DevParams &P, const Fields &F, unsigned long long *out);
hipError_t launch_split_batch(hipStream_t s, const DevParams &P, long long n, const SplitBatch &b, int dir);
hipError_t launch_recon_batch(hipStream_t s, const DevParams &P, long long n, int width, const double *st,
                              double flag, double *out);
hipError_t launch_copy(hipStream_t s, const double *src, double *dst, long n);
hipError_t launch_copy_nt(hipStream_t s, const double *src, double *dst, long n);
int ke_blocks(const DevParams &P);
hipError_t launch_m_from_prims(hipStream_t s, const DevParams &P, const MFields &M, const double *const *src,
                               int write_u);
hipError_t launch_m_to_prims(hipStream_t s, const DevParams &P, const MFields &M, double *const *dst);
hipError_t launch_m_fill_ghosts(hipStream_t s, const DevParams &P, const MFields &M);
hipError_t launch_m_cfl(hipStream_t s, const DevParams &P, const MFields &M, unsigned long long *out);
hipError_t launch_m_shadow_to_prims64(hipStream_t s, const DevParams &P, const MFields &M, const Prims64 &out);
hipError_t launch_m_state_prims64(hipStream_t s, const DevParams &P, const MFields &M, const Prims64 &out, int add_du);
hipError_t launch_m_face_diff(hipStream_t s, const DevParams &P, const MFields &M, const FaceArrays &ff);
bool m_xrow_supported(const DevParams &P);
hipError_t launch_m_sweep(hipStream_t s, const DevParams &P, const MFields &M, int dir, int seg_len);
hipError_t launch_m_gradflux(hipStream_t s, const DevParams &P, const MFields &M, int seg_len, int bx, int by,
                             int rows);
hipError_t launch_m_vdiv(hipStream_t s, const DevParams &P, const MFields &M, int mode, double dt, int seg,
                         unsigned long long *flag, int dev_dt, int bx = 64, int by = 4, int lds = 0, int img = 0);
bool m_vdiv_lds_form(const DevParams &P, int bx, int by, int lds);
bool m_vdiv_zs_form(const DevParams &P, int bx, int by, int lds, int img);
hipError_t launch_m_vdiv_zs(hipStream_t s, const DevParams &P, const MFields &M, const Shadow &So, int mode,
                            double dt, int seg, unsigned long long *flag, int dev_dt);
// the workgroup shape tiles the interior with whole waves: km_vdiv_z can
// reduce the next step's CFL minimum (dev_dt 2)
bool m_vdiv_cfl_next(const DevParams &P, int bx, int by);
hipError_t launch_ke(hipStream_t s, const DevParams &P, const Fields &F, double *partial);

// tangent-linear / adjoint RHS (ns3d_adjoint.hip)
bool ad_supported(const DevParams &P);
int ad_blocks(const DevParams &P);
int ad_seg(const DevParams &P, int dir);
hipError_t launch_tlm_rhs(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int char_recon);
hipError_t launch_adj_sweep(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, int dir,
                            int seg, const CharBuf *cb, int char_recon);
int char_buf_arrays(const DevParams &P);
bool adj_by_faces(const DevParams &P, int dir, int char_recon);
bool adj_x_stores(const DevParams &P, int char_recon);
hipError_t launch_adj_vflux(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double w_obj,
                            double *partial);
hipError_t launch_adj_grad(hipStream_t s, const DevParams &P, const Fields &F, const AdjArrays &A, double a,
                           double bx, double bl, int init, int add_acc);
bool adj_xrow(const DevParams &P);
hipError_t launch_dissip(hipStream_t s, const DevParams &P, const Fields &F, double *partial);
hipError_t launch_sum_partials(hipStream_t s, const double *partial, int n, double *out);

// monitors (ns3d_monitors.hip)
int monitor_blocks();
hipError_t launch_pack(hipStream_t s, const DevParams &P, const double *src, double *dst);
hipError_t launch_spec_bins(hipStream_t s, const DevParams &P, const void *U, double *partial, int nbins,
                            double scale, int accumulate);
hipError_t launch_stats(hipStream_t s, const DevParams &P, const double *const *W, double *partial);
hipError_t launch_resid(hipStream_t s, const DevParams &P, const double *const *Wc, const double *const *U0,
                        const float *const *dU, int conserved, double *partial);

}  // namespace ns3d
