/*
 * ns3d_oracle.h -- CPU FP64 oracle: a clean-room C restatement of the
 * reference solver's hot path (wylx-2/ns3d_cpp), used ONLY as the checker by
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  It is
 * test infrastructure, never part of the shipped GPU path.
 *
 * Every function follows the operation order of the reference file:line it
 * cites, so that FP64 results can be compared bit for bit or to ~1e-15.
 * The reference itself cannot be built in this image (each hot-path file
 * includes <fftw3-mpi.h>, which is absent), so the restatement is pinned by
 * closed forms, invariants and convergence tests under tests/.
 *
 * WENO5's constants (eps, 13/12, 1/4, the linear weights and the sub-stencil
 * faces) are pinned by an exact rational derivation from the scheme's
 * definition, and the characteristic path's tangential basis (the
 * coordinate axes other than the normal, InviscousFlux_new.cpp:568-578) by
 * its structure, by the exact zero face flux of a tangential velocity that
 * is zero everywhere, and by a bitwise recomposition of the face from the
 * matrices (tests/test_oracle_pins_exact.py).
 */
#ifndef NS3D_ORACLE_H
#define NS3D_ORACLE_H

#include "../include/ns3d_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* array ids for orc_array() */
enum {
    ORC_RHO = 0, ORC_RHOU, ORC_RHOV, ORC_RHOW, ORC_E,            /* 0..4   */
    ORC_RHO0, ORC_RHOU0, ORC_RHOV0, ORC_RHOW0, ORC_E0,           /* 5..9   */
    ORC_U, ORC_V, ORC_W, ORC_P, ORC_T, ORC_C,                    /* 10..15 */
    ORC_RHS0, ORC_RHS1, ORC_RHS2, ORC_RHS3, ORC_RHS4,            /* 16..20 */
    ORC_GRAD0,                                                   /* 21..32 */
    ORC_VF0 = ORC_GRAD0 + 12,                                    /* 33..44 */
    ORC_NCELL_ARRAYS = ORC_VF0 + 12,                             /* 45     */
    ORC_FX0 = ORC_NCELL_ARRAYS,                                  /* 45..49 */
    ORC_FY0 = ORC_FX0 + 5,                                       /* 50..54 */
    ORC_FZ0 = ORC_FY0 + 5,                                       /* 55..59 */
    ORC_NARRAYS = ORC_FZ0 + 5
};

typedef struct orc_field orc_field;

/* derived members, initialization.cpp:324-377 */
void orc_params_finalize(ns3d_params *p, ns3d_grid *g);
/* Field3D::allocate, field_structures.h:225-334 (rho=1, everything else 0) */
orc_field *orc_create(const ns3d_params *p, const ns3d_grid *g);
void orc_destroy(orc_field *f);
double *orc_array(orc_field *f, int id);
long long orc_array_len(const orc_field *f, int id);
void orc_sizes(const orc_field *f, int *sx, int *sy, int *sz, int *ng);

void orc_c2p(orc_field *f);             /* field_structures.h:405-427 */
void orc_p2c(orc_field *f);             /* field_structures.h:430-451 */
void orc_record_u0(orc_field *f);       /* field_structures.h:454-460 */
void orc_apply_boundary(orc_field *f);  /* boundary.cpp:14-54 */
void orc_exchange_vflux(orc_field *f);  /* field_structures.h:736-791 */

/* split_StagerWarming / split_LaxFriedrichs (InviscousFlux_new.cpp:439-530)
 * on one state; fvs = NS3D_FVS_*, eps used by the epsilon-modified SW. */
void orc_split_state(int fvs, double gamma, double eps,
                     double rho, double u, double v, double w, double E,
                     double p, double c, int n1, int n2, int n3,
                     double Fp[5], double Fm[5]);
/* reconstruct_select (reconstruction.cpp:13-85); returns NaN and sets *err
 * if the stencil size does not match the scheme (the reference throws). */
double orc_reconstruct(const double *st, int n, double flag,
                       const ns3d_params *p, int *err);
/* computeRoeAveragedState_new + build_eigen_matrices_new
 * (InviscousFlux_new.cpp:532-659) for the face between states ul and ur
 * (rho, u, v, w, p), normal (n1, n2, n3): L and R row-major 5x5. */
void orc_char_matrices(const double ul[5], const double ur[5], int n1, int n2, int n3, double gamma,
                       double L[25], double R[25]);
/* the central / one-sided derivative of ViscousFlux.cpp:15-40 on the 9-point
 * window f[0..8] centred at f[4]; order in {1,-1,2,4,6,8} */
double orc_diff(const double *f9, int order, double dx);
/* choose_scheme (ViscousFlux.cpp:43-118) */
int orc_choose_scheme(int idx, int nstart, int nend, int periodic, int vis);
/* SolverParams::get_mu (field_structures.h:94-105) + mu-law extensions */
double orc_get_mu(const ns3d_params *p, double T);

void orc_du_inviscous(orc_field *f);            /* InviscousFlux_new.cpp:4-163   */
void orc_du_inviscous_character(orc_field *f);  /* InviscousFlux_new.cpp:165-437 */
void orc_compute_gradients(orc_field *f);       /* ViscousFlux.cpp:240-279 */
void orc_compute_viscous_flux(orc_field *f);    /* ViscousFlux.cpp:282-339 */
void orc_compute_vis_flux(orc_field *f);        /* ViscousFlux.cpp:341-386 */
void orc_compute_rhs(orc_field *f);             /* range_kutta.cpp:11-99   */
/* runge_kutta_3 (range_kutta.cpp:102-163); returns 0, or 1 if the post-step
 * scan finds rho, p or T <= 0 (first such cell in *bad_id, -1 otherwise). */
int orc_rk3_step(orc_field *f, double dt, long long *bad_id);
/* one time_advance step body (maincycle.cpp:22-56) with fixed dt > 0 or
 * CFL dt when dt <= 0; returns the status of orc_rk3_step. */
int orc_step(orc_field *f, double dt, double *dt_used, long long *bad_id);
double orc_compute_timestep(orc_field *f);      /* timestep.cpp:5-33 */
double orc_total_energy(orc_field *f);          /* postprocess.cpp:103-128 */

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
