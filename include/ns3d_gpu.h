/*
 * ns3d_gpu.h -- C ABI of the MI355X RHS + RK3 engine for the 3D compressible
 * Navier-Stokes finite-difference solver.
 *
 * One context owns one subdomain (the whole periodic box on one GPU) resident
 * in HBM as structure-of-arrays FP64 fields with ng ghost layers, indexed like
 * the reference's idx3 (i fastest):  id = (k*sy + j)*sx + i,  sx = nx + 2*ng.
 * Host buffers passed to upload/download use exactly that layout.
 *
 * Conventions: every entry point returns an int status (NS3D_OK == 0); no C++
 * exception crosses this boundary; the caller owns host buffers; one context
 * is bound to the device that was current at ns3d_create and to its own HIP
 * stream; a context is not thread-safe.
 *
 * Reference interface replaced (wylx-2/ns3d_cpp, file:line):
 *   SolverParams          field_structures.h:85-166
 *   GridDesc / LocalDesc  field_structures.h:32-59
 *   Field3D storage       field_structures.h:176-334
 *   free functions        ns3d_func.h:56-163, timestep.cpp:5
 */
#ifndef NS3D_GPU_H
#define NS3D_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference aborts / throws / prints instead) ---- */
enum {
    NS3D_OK = 0,
    NS3D_ERR_INVALID_ARG = 1,   /* null pointer, bad enum value, bad size       */
    NS3D_ERR_CONFIG = 2,        /* scheme combination the reference rejects
                                   (reconstruction.cpp:18-22) or not built      */
    NS3D_ERR_GRID_TOO_SMALL = 3,/* interior smaller than the stencil / ghosts   */
    NS3D_ERR_NONPHYSICAL = 4,   /* rho, p or T <= 0 (or NaN) after a step
                                   (range_kutta.cpp:150-161)                    */
    NS3D_ERR_HIP = 5,           /* HIP runtime error                            */
    NS3D_ERR_OOM = 6,           /* device allocation failed                     */
    NS3D_ERR_UNSUPPORTED = 7,   /* option valid in the reference, not on GPU yet*/
    NS3D_ERR_IO = 8             /* solver.in could not be opened                */
};

/* ---- enums: same members and order as the reference's enum classes ---- */
enum { NS3D_FVS_STEGER_WARMING = 0, NS3D_FVS_LAX_FRIEDRICHS = 1 };
/* SolverParams::Reconstruction, field_structures.h:118-126 */
enum {
    NS3D_RECON_WENO5 = 0, NS3D_RECON_WCNS = 1, NS3D_RECON_LINEAR = 2,
    NS3D_RECON_MDCD = 3, NS3D_RECON_MDCD_HYBRID = 4, NS3D_RECON_UPWIND_7TH = 5,
    NS3D_RECON_UD7L = 6
};
/* SolverParams::ViscousScheme, field_structures.h:127-131 */
enum { NS3D_VIS_C8TH = 0, NS3D_VIS_C6TH = 1, NS3D_VIS_C4TH = 2 };
/* SolverParams::BCType, field_structures.h:145 */
enum { NS3D_BC_PERIODIC = 0, NS3D_BC_WALL = 1, NS3D_BC_SYMMETRY = 2,
       NS3D_BC_OUTFLOW = 3, NS3D_BC_INFLOW = 4 };
/* viscosity law; the reference hard-wires Sutherland (field_structures.h:94-105) */
enum { NS3D_MU_SUTHERLAND = 0, NS3D_MU_CONSTANT = 1, NS3D_MU_POWER = 2 };
/* arithmetic mode of the GPU path */
enum { NS3D_PREC_FP64 = 0, NS3D_PREC_MIXED = 1 };

/* Mirror of SolverParams (field_structures.h:85-166) plus the options the
 * reference cannot express without a source edit (mu law, Sutherland S/T_ref,
 * epsilon-modified Steger-Warming of README.md:179, precision mode).
 * Derived members (Cv, Cp, Rgas, mu, stencil, ghost_layers) are filled by
 * ns3d_params_finalize exactly as read_solver_params_from_file does
 * (initialization.cpp:324-377). */
typedef struct ns3d_params {
    double gamma, Pr, Ma, Re, cfl, dt_fixed;
    int fvs_type, recon, vis_scheme, char_recon;
    double mdcd_diss, mdcd_disp, ud7l_alpha;
    int bc[6];                 /* xmin, xmax, ymin, ymax, zmin, zmax */
    int max_steps, monitor_stepfreq;
    double output_timefreq, total_time;
    /* extensions (defaults reproduce the reference bit for bit) */
    int mu_law;                /* NS3D_MU_*                                   */
    double sutherland_s;       /* S/T_ref; reference value 110.4/288.15       */
    double mu_power;           /* exponent of the power law                   */
    double sw_eps;             /* 0 => plain |lambda| (InviscousFlux_new.cpp:461) */
    int precision;             /* NS3D_PREC_*                                 */
    /* post-processing switches (field_structures.h:160-165) */
    int post_basicfield, isotropic_analyse, monitor_res, monitor_energy;
    /* derived */
    double Cv, Cp, Rgas, mu;
    int stencil, ghost_layers;
} ns3d_params;

/* Mirror of GridDesc + the single-rank LocalDesc (field_structures.h:32-59). */
typedef struct ns3d_grid {
    int nx, ny, nz;            /* interior points                              */
    int ng;                    /* ghost layers per side (all three axes)       */
    double x0, y0, z0, Lx, Ly, Lz;
    double dx, dy, dz;         /* Lx/nx ... (initialization.cpp:324-326)       */
} ns3d_grid;

/* SolverParams() defaults (field_structures.h:85-166) and the parser's grid
 * defaults (initialization.cpp:212-215). */
void ns3d_params_default(ns3d_params *p, ns3d_grid *g);
/* Derived members: stencil/ghost_layers (initialization.cpp:329-372),
 * Cv, Cp, Rgas, mu (:374-377), dx/dy/dz (:324-326); g->ng = ghost_layers. */
int ns3d_params_finalize(ns3d_params *p, ns3d_grid *g);
/* read_solver_params_from_file (initialization.cpp:199-380): key = value,
 * case-insensitive, '#' comments, unknown keys ignored; then finalize. */
int ns3d_params_from_file(const char *path, ns3d_params *p, ns3d_grid *g);

typedef struct ns3d_ctx ns3d_ctx;

/* Field3D::allocate (field_structures.h:225-334) on the current device. */
int ns3d_create(const ns3d_params *p, const ns3d_grid *g, ns3d_ctx **out);
int ns3d_destroy(ns3d_ctx *ctx);
/* Number of cells per array including ghosts, sx*sy*sz. */
long long ns3d_cells(const ns3d_ctx *ctx);
int ns3d_sizes(const ns3d_ctx *ctx, int *sx, int *sy, int *sz, int *ng);
const char *ns3d_status_string(int status);
/* Last HIP error text recorded by the context ("" if none). */
const char *ns3d_last_error(const ns3d_ctx *ctx);

/* Upload primitives rho,u,v,w,p (each sx*sy*sz doubles, idx3 layout; ghost
 * values are ignored) then do what main.cpp:70-71 does: apply_boundary and
 * primitiveToConserved (rho clamp, T). */
int ns3d_upload_primitive(ns3d_ctx *ctx, const double *rho, const double *u,
                          const double *v, const double *w, const double *p);
/* Upload conserved rho, rho*u, rho*v, rho*w, E and convert with
 * conservedToPrimitive (field_structures.h:405-427), then apply_boundary. */
int ns3d_upload_conserved(ns3d_ctx *ctx, const double *rho, const double *ru,
                          const double *rv, const double *rw, const double *E);
/* Download rho,u,v,w,p,T (any pointer may be NULL to skip). */
int ns3d_download_primitive(ns3d_ctx *ctx, double *rho, double *u, double *v,
                            double *w, double *p, double *T);
/* Download conserved = primitiveToConserved(current primitives)
 * (field_structures.h:430-451), i.e. the reference's rho..E arrays. */
int ns3d_download_conserved(ns3d_ctx *ctx, double *rho, double *ru, double *rv,
                            double *rw, double *E);
/* Download the 5 RHS arrays written by ns3d_compute_rhs (interior only). */
int ns3d_download_rhs(ns3d_ctx *ctx, double *r0, double *r1, double *r2,
                      double *r3, double *r4);

/* apply_boundary (boundary.cpp:14-54): periodic ghost fill of rho,u,v,w,p
 * (faces only, like exchange_halos_physical field_structures.h:795-853), then
 * the wall / symmetry / outflow / inflow fills of the non-periodic sides
 * (boundary.cpp:58-207) in the order xmin, xmax, ..., zmax. */
int ns3d_apply_boundary(ns3d_ctx *ctx);

/* du_inviscous (InviscousFlux_new.cpp:4-163): face fluxes into device face
 * arrays laid out as idx_fx/idx_fy/idx_fz (field_structures.h:72-81). */
int ns3d_du_inviscous(ns3d_ctx *ctx);
/* du_inviscous_character (InviscousFlux_new.cpp:165-437, ns3d_func.h:155)
 * into the same face arrays, whatever p->char_recon says. */
int ns3d_du_inviscous_character(ns3d_ctx *ctx);
/* Download the 5 face-flux arrays of direction dir (0=x,1=y,2=z); each has
 * (sx-1)*sy*sz, sx*(sy-1)*sz or sx*sy*(sz-1) doubles. */
int ns3d_download_face_flux(ns3d_ctx *ctx, int dir, double *f0, double *f1,
                            double *f2, double *f3, double *f4);
/* compute_gradients (ViscousFlux.cpp:240-279): 12 arrays, order
 * du_dx du_dy du_dz dv_dx dv_dy dv_dz dw_dx dw_dy dw_dz dT_dx dT_dy dT_dz. */
int ns3d_compute_gradients(ns3d_ctx *ctx);
int ns3d_download_gradients(ns3d_ctx *ctx, double *const grads[12]);
/* compute_viscous_flux (ViscousFlux.cpp:282-339) from the stored gradients:
 * 12 arrays, order Fv(momx,momy,momz,E) Hv(...) Gv(...). */
int ns3d_compute_viscous_flux(ns3d_ctx *ctx);
int ns3d_download_viscous_flux(ns3d_ctx *ctx, double *const flux[12]);

/* compute_vis_flux (ViscousFlux.cpp:341-386) from the fluxes stored by
 * ns3d_compute_viscous_flux, on rhs = 0: the 5 RHS arrays (download with
 * ns3d_download_rhs) receive ((0 + Dx Fv) + Dy Hv) + Dz Gv, flux ghosts read
 * as periodic images on periodic axes (exchange_halos_viscous_flux,
 * field_structures.h:736-791), the choose_scheme closures on the others. */
int ns3d_compute_vis_flux(ns3d_ctx *ctx);
/* compute_rhs (range_kutta.cpp:11-99) with the fused sweep kernels:
 * rhs = viscous divergence - inviscid flux difference, interior cells. */
int ns3d_compute_rhs(ns3d_ctx *ctx);
/* runge_kutta_3 (range_kutta.cpp:102-163) preceded by recordConservedTo0
 * (maincycle.cpp:22).  Returns NS3D_ERR_NONPHYSICAL if the post-step scan
 * finds rho, p or T <= 0 or NaN; ns3d_nonphysical_cell gives the cell. */
int ns3d_rk3_step(ns3d_ctx *ctx, double dt);
int ns3d_nonphysical_cell(const ns3d_ctx *ctx, int *i, int *j, int *k);
/* Advance nsteps RK3 steps with fixed dt (no host round trip per step).
 * With NS3D_GRAPHS=1 in the environment at ns3d_create, the steps after the
 * first replay a captured HIP graph of one step. */
int ns3d_advance(ns3d_ctx *ctx, int nsteps, double dt);
/* The time loop of main/maincycle (maincycle.cpp:18-107) without output
 * stops or dt clipping: nsteps times dt = compute_timestep (timestep.cpp:
 * 5-33, single rank), runge_kutta_3, t += dt, with dt kept on the device (no
 * host round trip per step; graph replay as ns3d_advance).  *time: t in, t
 * after the steps out (the same sums as the host loop).  The same dt and
 * states as ns3d_compute_timestep_local + ns3d_rk3_step per step, bit for
 * bit; NS3D_ERR_NONPHYSICAL as ns3d_rk3_step, reported after the last step. */
int ns3d_advance_cfl(ns3d_ctx *ctx, int nsteps, double *time);
/* compute_timestep (timestep.cpp:5-33) without the MPI_Allreduce: returns
 * cfl * min over interior cells; the caller reduces across ranks. */
int ns3d_compute_timestep_local(ns3d_ctx *ctx, double *dt);
/* compute_total_energy (postprocess.cpp:103-128): sum 0.5 rho |u|^2 dV. */
int ns3d_total_energy(ns3d_ctx *ctx, double *ek);

/* ---- monitors (post-processing of the reference, computed on the device) */
/* Shell spectrum E(k), k = 0 .. nbins-1 (the reference writes Kmax + 1 =
 * 2*nx + 1 bins), of compute_energy_spectrum (EnergySpectrum.cpp:62-154):
 * E(k) = sum over modes with floor(|k|) = k of (|u^|^2 + |v^|^2 + |w^|^2)/2,
 * coefficients normalised by 1/(nx*ny*nz); shells >= nbins are dropped. */
int ns3d_energy_spectrum(ns3d_ctx *ctx, double *Ek, int nbins);
/* compute_turbulence_statistics (ana_isotropic.cpp:226-330) with the spectrum
 * at full precision: out[0..12] = kinetic_energy, tke, u_rms, dissipation
 * (2 nu sum k^2 E(k)), taylor, eta, u_eta, tau_eta, taylor_Li, Re_lambda,
 * Mach_t, mean_mu, mean_rho.  n must be >= 13. */
int ns3d_turbulence_statistics(ns3d_ctx *ctx, double *out, int n);
/* compute_diagnostics (postprocess.cpp:7-96): out[m] = sqrt(sum (U_m -
 * U0_m)^2 dx^3) / max |U_m| for rho, rho u, rho v, rho w, E, with U0 the
 * state at the start of the last RK3 step. */
int ns3d_residuals(ns3d_ctx *ctx, double out[5]);

/* ---- point functions of the reference surface, batched on the device.
 * Host arrays in and out; no context; the call returns when the results are
 * in place.  The arithmetic is the FP64 path's (bit-identical to the
 * reference's expressions). */
/* split_StagerWarming / split_LaxFriedrichs (InviscousFlux_new.cpp:439-530,
 * ns3d_func.h:136-145) of n states W = rho, u, v, w, p (5 arrays of n), with
 * E and c formed as primitiveToConserved forms them (field_structures.h:
 * 445-449); p->fvs_type selects the split, p->sw_eps the SW variant.
 * (n1, n2, n3) must be a unit axis.  Fp, Fm: 5 arrays of n each. */
int ns3d_split_batch(const ns3d_params *p, long long n, const double *const W[5], int n1, int n2, int n3,
                     double *const Fp[5], double *const Fm[5]);
/* reconstruct_select (reconstruction.cpp:13-85, ns3d_func.h:96) of n
 * stencils of `width` values each (row-major), reversed for flag < 0.
 * NS3D_ERR_CONFIG when width is not the scheme's stencil (the reference
 * throws std::runtime_error); other schemes take its averaging fallback,
 * except with width 2 and flag < 0, where the reference reads past the
 * stencil: NS3D_ERR_CONFIG. */
int ns3d_reconstruct_batch(const ns3d_params *p, long long n, int width, const double *stencils, double flag,
                           double *out);

/* Switch the arithmetic of an existing context between NS3D_PREC_FP64 and
 * NS3D_PREC_MIXED, keeping its state (SolverParams has no counterpart: the
 * reference is FP64 only). */
int ns3d_set_precision(ns3d_ctx *ctx, int precision);

/* Synchronise the context's stream. */
int ns3d_sync(ns3d_ctx *ctx);
/* Device pointer of the context's HIP stream (hipStream_t). */
void *ns3d_stream(ns3d_ctx *ctx);

/* Per-kernel timing with HIP events recorded around every launch on the
 * context's stream (the reference only times whole steps with MPI_Wtime,
 * maincycle.cpp:70-77).  Enabling (or disabling) resets the totals. */
int ns3d_profile_enable(ns3d_ctx *ctx, int on);
/* Number of kernel kinds timed so far (negative status on error). */
int ns3d_profile_count(ns3d_ctx *ctx);
int ns3d_profile_entry(ns3d_ctx *ctx, int idx, char *name, int name_len,
                       double *total_ms, long long *launches);
/* Streaming-copy probe over the context's scratch arrays (12 cell arrays,
 * read + write): the better achieved GB/s of `reps` copies by a grid-stride
 * 16-byte copy and by a flat nontemporal one.  No reference counterpart; it
 * measures the achievable HBM bandwidth of the device the context is on.
 * Overwrites the stored gradient / viscous-flux scratch. */
int ns3d_copy_probe(ns3d_ctx *ctx, int reps, double *gbps);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  compute_timestep (timestep.cpp:5-33) of caller-owned
 * device arrays rho, u, v, w, p in the idx3 layout of (p, g) (sx*sy*sz
 * doubles each; ghost values unused; p->gamma, p->cfl and the grid's
 * extents are read, nothing needs finalizing), enqueued on `stream` (a
 * hipStream_t; NULL for the null stream): *dt_dev (device memory) receives
 * cfl * min over interior cells, bit for bit the value
 * ns3d_compute_timestep_local gives for the same state.  Allocates nothing:
 * its scratch is a static 16 KiB device array per device (2048 per-block
 * minima), shared by every call on that device, so calls that overlap on
 * one device race.  NS3D_ERR_INVALID_ARG for null pointers, non-positive
 * extents or more than 2^31 - 1 cells per array. */
int ns3d_cfl_dt_noalloc(const ns3d_params *p, const ns3d_grid *g, const double *rho, const double *u,
                        const double *v, const double *w, const double *pr, double *dt_dev, void *stream);

/* ---- Differentiation (no counterpart in the reference, which has no
 * derivative code).  Every reconstruction a context accepts (WENO5 / linear /
 * MDCD / MDCD_HYBRID / UD7 / UD7L, component-wise or characteristic), with
 * periodic or wall / symmetry / outflow / inflow sides; every grid extent
 * must be >= 8 (else NS3D_ERR_GRID_TOO_SMALL).  Host arrays are full sx*sy*sz
 * layouts; only interior cells are read or written.  |x| in the plain SW/LF
 * splits is differentiated as sign(x); sw_eps > 0 makes the split smooth.
 * Mixed-precision contexts: the forward and recomputed steps are the mixed
 * steps (checkpoints hold the FP64 state and its FP32 shadow, so the
 * recomputation is bitwise), and the derivatives are those of the FP64
 * scheme taken at the FP64 primitives of each stage's state (U, or U plus
 * the stage increment). */

/* Tangent-linear RHS at the current state: out = dR/dU * dU. */
int ns3d_tlm_rhs(ns3d_ctx *ctx, const double *const dU[5], double *const out[5]);
/* Adjoint RHS at the current state: out = (dR/dU)^T lam + w_obj * d(sum of
 * Phi over interior cells)/dU, with Phi = tau_ij du_i/dx_j the viscous
 * dissipation function; *dtheta (may be NULL) = d/d(1/Re) of
 * lam . R + w_obj * sum Phi. */
int ns3d_adj_rhs(ns3d_ctx *ctx, const double *const lam[5], double w_obj,
                 double *const out[5], double *dtheta);
/* Volume-mean dissipation eps = mean over interior cells of Phi (same
 * gradient scheme and mu(T) as the viscous flux). */
int ns3d_dissipation(ns3d_ctx *ctx, double *eps);
/* Run nsteps RK3 steps of fixed dt from the current state, storing a
 * checkpoint every ckpt_every steps, and return
 *     J = sum_{n=0}^{nsteps-1} dt * eps(U^n).
 * The context then holds U^nsteps. */
int ns3d_adj_forward(ns3d_ctx *ctx, int nsteps, double dt, int ckpt_every, double *J);
/* Reverse pass of the last ns3d_adj_forward (checkpoint and recompute):
 * lam0 = dJ/dU^0 + (dU^N/dU^0)^T lamN (lamN may be NULL for zero), conserved
 * variables, and *dJ_dinvRe = the derivative with respect to 1/Re (may be
 * NULL).  The context is left holding U^nsteps again. */
int ns3d_adj_backward(ns3d_ctx *ctx, const double *const lamN[5],
                      double *const lam0[5], double *dJ_dinvRe);
/* Device bytes held for differentiation (adjoint arrays + checkpoints). */
int ns3d_adj_memory(const ns3d_ctx *ctx, long long *bytes);

#ifdef __cplusplus
}
#endif
#endif /* NS3D_GPU_H */
