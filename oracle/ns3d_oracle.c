/*
 * ns3d_oracle.c -- CPU FP64 oracle (test infrastructure only; see header).
 *
 * Clean-room C restatement of the reference's RHS + RK3 path.  Expressions
 * keep the reference's evaluation order (C and C++ evaluate the same
 * left-to-right chains identically under FLT_EVAL_METHOD == 0); build with
 * -ffp-contract=off so no FMA is formed, as in the reference's x86-64
 * baseline build (Makefile:6, no -march).
 *
 * Parallelism: per-cell loops are split over OpenMP threads; every result is
 * independent of the thread count (reductions are serial).
 */
#include "ns3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_field {
    ns3d_params P;
    ns3d_grid G;
    int nx, ny, nz, ng, sx, sy, sz;
    long long ncell, nfx, nfy, nfz;
    double *a[ORC_NARRAYS];
};

#define IDX(F, i, j, k) (((long long)(k) * (F)->sy + (j)) * (F)->sx + (i))
/* idx_fx / idx_fy / idx_fz, field_structures.h:72-81 */
#define IDXFX(F, i, j, k) (((long long)(k) * (F)->sy + (j)) * ((F)->sx - 1) + (i))
#define IDXFY(F, i, j, k) (((long long)(k) * ((F)->sy - 1) + (j)) * (F)->sx + (i))
#define IDXFZ(F, i, j, k) (((long long)(k) * (F)->sy + (j)) * (F)->sx + (i))

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ params */

void orc_params_finalize(ns3d_params *P, ns3d_grid *G)
{
    /* initialization.cpp:324-326 */
    G->dx = G->Lx / (G->nx);
    G->dy = G->Ly / (G->ny);
    G->dz = G->Lz / (G->nz);
    /* initialization.cpp:329-357 */
    switch (P->recon) {
    case NS3D_RECON_WENO5: P->ghost_layers = 3; P->stencil = 6; break;
    case NS3D_RECON_LINEAR: P->ghost_layers = 2; P->stencil = 2; break;
    case NS3D_RECON_MDCD: P->ghost_layers = 3; P->stencil = 6; break;
    case NS3D_RECON_MDCD_HYBRID: P->ghost_layers = 3; P->stencil = 6; break;
    case NS3D_RECON_UPWIND_7TH: P->ghost_layers = 4; P->stencil = 8; break;
    case NS3D_RECON_UD7L: P->ghost_layers = 4; P->stencil = 8; break;
    default: break; /* WCNS: ghost_layers/stencil keep their prior values */
    }
    /* initialization.cpp:358-372 */
    switch (P->vis_scheme) {
    case NS3D_VIS_C4TH: if (P->ghost_layers < 2) P->ghost_layers = 2; break;
    case NS3D_VIS_C6TH: if (P->ghost_layers < 3) P->ghost_layers = 3; break;
    case NS3D_VIS_C8TH: if (P->ghost_layers < 4) P->ghost_layers = 4; break;
    default: break;
    }
    /* initialization.cpp:374-377 */
    P->Cv = 1.0 / (P->gamma * (P->gamma - 1.0) * P->Ma * P->Ma);
    P->Cp = P->Cv * P->gamma;
    P->Rgas = 1.0 / (P->Ma * P->Ma * P->gamma);
    P->mu = 1.0 / P->Re;
    G->ng = P->ghost_layers;
}

/* SolverParams::get_mu, field_structures.h:94-105.  Sutherland is the
 * reference's law; S/T_ref = 110.4/288.15 reproduces it.  Constant and power
 * laws are the commented-out alternatives, made selectable. */
double orc_get_mu(const ns3d_params *P, double T)
{
    if (P->mu_law == NS3D_MU_CONSTANT) return P->mu;
    if (P->mu_law == NS3D_MU_POWER) return P->mu * pow(T, P->mu_power);
    double S_ref = P->sutherland_s;
    return P->mu * pow(T, 1.5) * (1.0 + S_ref) / (T + S_ref);
}

/* ------------------------------------------------------------------ field */

orc_field *orc_create(const ns3d_params *p, const ns3d_grid *g)
{
    orc_field *F = (orc_field *)calloc(1, sizeof(orc_field));
    if (!F) return NULL;
    F->P = *p;
    F->G = *g;
    orc_params_finalize(&F->P, &F->G);
    F->nx = g->nx; F->ny = g->ny; F->nz = g->nz;
    F->ng = F->G.ng;
    F->sx = F->nx + 2 * F->ng;
    F->sy = F->ny + 2 * F->ng;
    F->sz = F->nz + 2 * F->ng;
    F->ncell = (long long)F->sx * F->sy * F->sz;
    F->nfx = (long long)(F->sx - 1) * F->sy * F->sz;
    F->nfy = (long long)F->sx * (F->sy - 1) * F->sz;
    F->nfz = (long long)F->sx * F->sy * (F->sz - 1);
    for (int a = 0; a < ORC_NARRAYS; ++a) {
        long long n = orc_array_len(F, a);
        F->a[a] = (double *)calloc((size_t)n, sizeof(double));
        if (!F->a[a]) { orc_destroy(F); return NULL; }
    }
    /* rho.assign(tot, 1.0), field_structures.h:232 */
    for (long long n = 0; n < F->ncell; ++n) F->a[ORC_RHO][n] = 1.0;
    return F;
}

void orc_destroy(orc_field *F)
{
    if (!F) return;
    for (int a = 0; a < ORC_NARRAYS; ++a) free(F->a[a]);
    free(F);
}

double *orc_array(orc_field *F, int id)
{
    if (id < 0 || id >= ORC_NARRAYS) return NULL;
    return F->a[id];
}

long long orc_array_len(const orc_field *F, int id)
{
    if (id < ORC_NCELL_ARRAYS) return F->ncell;
    if (id < ORC_FY0) return F->nfx;
    if (id < ORC_FZ0) return F->nfy;
    return F->nfz;
}

void orc_sizes(const orc_field *F, int *sx, int *sy, int *sz, int *ng)
{
    if (sx) *sx = F->sx;
    if (sy) *sy = F->sy;
    if (sz) *sz = F->sz;
    if (ng) *ng = F->ng;
}

/* ------------------------------------------------ conversions (L1 layer) */

/* Field3D::conservedToPrimitive, field_structures.h:405-427 */
void orc_c2p(orc_field *F)
{
    const double gamma = F->P.gamma;
    const double Rgas = F->P.Rgas;
    double *rho = F->a[ORC_RHO], *rhou = F->a[ORC_RHOU], *rhov = F->a[ORC_RHOV];
    double *rhow = F->a[ORC_RHOW], *E = F->a[ORC_E];
    double *u = F->a[ORC_U], *v = F->a[ORC_V], *w = F->a[ORC_W];
    double *p = F->a[ORC_P], *T = F->a[ORC_T], *c = F->a[ORC_C];
    const long long N = F->ncell;
#pragma omp parallel for schedule(static)
    for (long long id = 0; id < N; ++id) {
        double rr = rho[id];
        if (rr <= 0.0) { rr = 1e-12; rho[id] = rr; }
        u[id] = rhou[id] / rr;
        v[id] = rhov[id] / rr;
        w[id] = rhow[id] / rr;
        double kinetic = 0.5 * (u[id] * u[id] + v[id] * v[id] + w[id] * w[id]);
        double e_internal = E[id] / rr - kinetic;
        p[id] = (gamma - 1.0) * rr * e_internal;
        T[id] = p[id] / (rr * Rgas);
        c[id] = sqrt(gamma * p[id] / rr);
    }
}

/* Field3D::primitiveToConserved, field_structures.h:430-451 */
void orc_p2c(orc_field *F)
{
    const double gamma = F->P.gamma;
    const double Rgas = F->P.Rgas;
    double *rho = F->a[ORC_RHO], *rhou = F->a[ORC_RHOU], *rhov = F->a[ORC_RHOV];
    double *rhow = F->a[ORC_RHOW], *E = F->a[ORC_E];
    double *u = F->a[ORC_U], *v = F->a[ORC_V], *w = F->a[ORC_W];
    double *p = F->a[ORC_P], *T = F->a[ORC_T], *c = F->a[ORC_C];
    const long long N = F->ncell;
#pragma omp parallel for schedule(static)
    for (long long id = 0; id < N; ++id) {
        double rr = rho[id];
        if (rr <= 0.0) { rr = 1e-12; rho[id] = rr; }
        rhou[id] = rr * u[id];
        rhov[id] = rr * v[id];
        rhow[id] = rr * w[id];
        double kinetic = 0.5 * (u[id] * u[id] + v[id] * v[id] + w[id] * w[id]);
        double e_internal = p[id] / ((gamma - 1.0) * rr);
        E[id] = rr * (e_internal + kinetic);
        T[id] = p[id] / (rr * Rgas);
        c[id] = sqrt(gamma * p[id] / rr);
    }
}

/* Field3D::recordConservedTo0, field_structures.h:454-460 */
void orc_record_u0(orc_field *F)
{
    for (int m = 0; m < 5; ++m)
        memcpy(F->a[ORC_RHO0 + m], F->a[ORC_RHO + m], (size_t)F->ncell * sizeof(double));
}

/* ------------------------------------------------------ boundary / halo */

static int axis_periodic(const orc_field *F, int axis)
{
    /* a single rank has neighbours (itself) on an axis iff both sides are
     * periodic: initialization.cpp:315-323 sets periods[], MPI_Cart_shift
     * then returns self or MPI_PROC_NULL (field_structures.h:908-910). */
    return F->P.bc[2 * axis] == NS3D_BC_PERIODIC && F->P.bc[2 * axis + 1] == NS3D_BC_PERIODIC;
}

/* periodic image copy of a set of arrays over the six faces; faces only, the
 * cross directions restricted to interior cells, as the pack/unpack pairs of
 * field_structures.h:480-612 (and :615-733 for the viscous flux) do on one
 * rank with periods set. */
static void periodic_fill(orc_field *F, double *const *arr, int narr)
{
    const int ng = F->ng, nx = F->nx, ny = F->ny, nz = F->nz;
    if (axis_periodic(F, 0)) {
#pragma omp parallel for schedule(static)
        for (int k = ng; k < ng + nz; ++k)
            for (int j = ng; j < ng + ny; ++j)
                for (int ii = 0; ii < ng; ++ii) {
                    long long gl = IDX(F, ii, j, k), sl = IDX(F, ii + nx, j, k);
                    long long gr = IDX(F, ng + nx + ii, j, k), sr = IDX(F, ng + ii, j, k);
                    for (int a = 0; a < narr; ++a) { arr[a][gl] = arr[a][sl]; arr[a][gr] = arr[a][sr]; }
                }
    }
    if (axis_periodic(F, 1)) {
#pragma omp parallel for schedule(static)
        for (int k = ng; k < ng + nz; ++k)
            for (int jj = 0; jj < ng; ++jj)
                for (int i = ng; i < ng + nx; ++i) {
                    long long gl = IDX(F, i, jj, k), sl = IDX(F, i, jj + ny, k);
                    long long gr = IDX(F, i, ng + ny + jj, k), sr = IDX(F, i, ng + jj, k);
                    for (int a = 0; a < narr; ++a) { arr[a][gl] = arr[a][sl]; arr[a][gr] = arr[a][sr]; }
                }
    }
    if (axis_periodic(F, 2)) {
#pragma omp parallel for schedule(static)
        for (int kk = 0; kk < ng; ++kk)
            for (int j = ng; j < ng + ny; ++j)
                for (int i = ng; i < ng + nx; ++i) {
                    long long gl = IDX(F, i, j, kk), sl = IDX(F, i, j, kk + nz);
                    long long gr = IDX(F, i, j, ng + nz + kk), sr = IDX(F, i, j, ng + kk);
                    for (int a = 0; a < narr; ++a) { arr[a][gl] = arr[a][sl]; arr[a][gr] = arr[a][sr]; }
                }
    }
}

/* apply_wall_bc, boundary.cpp:58-134 (also apply_symmetry_bc, :137-141) */
static void wall_bc(orc_field *F, int face)
{
    double *rho = F->a[ORC_RHO], *u = F->a[ORC_U], *v = F->a[ORC_V];
    double *w = F->a[ORC_W], *p = F->a[ORC_P];
    const int ng = F->ng, sx = F->sx, sy = F->sy, sz = F->sz;
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                long long idg, idm;
                double *flip;
                if (face == 0) { if (i >= ng) continue; idg = IDX(F, i, j, k); idm = IDX(F, 2 * ng - 1 - i, j, k); flip = u; }
                else if (face == 1) { if (i < sx - ng) continue; idg = IDX(F, i, j, k); idm = IDX(F, 2 * (sx - ng) - 1 - i, j, k); flip = u; }
                else if (face == 2) { if (j >= ng) continue; idg = IDX(F, i, j, k); idm = IDX(F, i, 2 * ng - 1 - j, k); flip = v; }
                else if (face == 3) { if (j < sy - ng) continue; idg = IDX(F, i, j, k); idm = IDX(F, i, 2 * (sy - ng) - 1 - j, k); flip = v; }
                else if (face == 4) { if (k >= ng) continue; idg = IDX(F, i, j, k); idm = IDX(F, i, j, 2 * ng - 1 - k); flip = w; }
                else { if (k < sz - ng) continue; idg = IDX(F, i, j, k); idm = IDX(F, i, j, 2 * (sz - ng) - 1 - k); flip = w; }
                rho[idg] = rho[idm];
                u[idg] = u[idm];
                v[idg] = v[idm];
                w[idg] = w[idm];
                flip[idg] = -flip[idm];
                p[idg] = p[idm];
            }
}

/* apply_outflow_bc, boundary.cpp:144-182 */
static void outflow_bc(orc_field *F, int face)
{
    double *arr[5] = {F->a[ORC_RHO], F->a[ORC_U], F->a[ORC_V], F->a[ORC_W], F->a[ORC_P]};
    const int ng = F->ng, sx = F->sx, sy = F->sy, sz = F->sz;
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                long long id1 = IDX(F, i, j, k), id2;
                if (face == 0) { if (i >= ng) continue; id2 = IDX(F, ng, j, k); }
                else if (face == 1) { if (i < sx - ng) continue; id2 = IDX(F, sx - ng - 1, j, k); }
                else if (face == 2) { if (j >= ng) continue; id2 = IDX(F, i, ng, k); }
                else if (face == 3) { if (j < sy - ng) continue; id2 = IDX(F, i, sy - ng - 1, k); }
                else if (face == 4) { if (k >= ng) continue; id2 = IDX(F, i, j, ng); }
                else { if (k < sz - ng) continue; id2 = IDX(F, i, j, sz - ng - 1); }
                for (int a = 0; a < 5; ++a) arr[a][id1] = arr[a][id2];
            }
}

/* apply_inflow_bc, boundary.cpp:185-207 (hard-wired state, as in the source) */
static void inflow_bc(orc_field *F, int face)
{
    const double rho0 = 1.0, u0 = 1.0, v0 = 0.0, w0 = 0.0, p0 = 1.0, gamma = 1.4;
    const double E0 = p0 / (gamma - 1.0) + 0.5 * rho0 * (u0 * u0 + v0 * v0 + w0 * w0);
    const int ng = F->ng, sx = F->sx, sy = F->sy, sz = F->sz;
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                int in;
                if (face == 0) in = i < ng;
                else if (face == 1) in = i >= sx - ng;
                else if (face == 2) in = j < ng;
                else if (face == 3) in = j >= sy - ng;
                else if (face == 4) in = k < ng;
                else in = k >= sz - ng;
                if (!in) continue;
                long long id = IDX(F, i, j, k);
                F->a[ORC_RHO][id] = rho0;
                F->a[ORC_U][id] = u0; F->a[ORC_V][id] = v0; F->a[ORC_W][id] = w0;
                F->a[ORC_P][id] = p0;
                F->a[ORC_E][id] = E0;
            }
}

/* apply_boundary, boundary.cpp:14-54: periodic halo exchange of
 * (rho,u,v,w,p), then physical BCs on sides without a neighbour. */
void orc_apply_boundary(orc_field *F)
{
    double *prim[5] = {F->a[ORC_RHO], F->a[ORC_U], F->a[ORC_V], F->a[ORC_W], F->a[ORC_P]};
    periodic_fill(F, prim, 5);
    for (int face = 0; face < 6; ++face) {
        if (axis_periodic(F, face / 2)) continue;
        switch (F->P.bc[face]) {
        case NS3D_BC_WALL:
        case NS3D_BC_SYMMETRY: wall_bc(F, face); break;
        case NS3D_BC_OUTFLOW: outflow_bc(F, face); break;
        case NS3D_BC_INFLOW: inflow_bc(F, face); break;
        default: break;
        }
    }
}

/* exchange_halos_viscous_flux, field_structures.h:736-791 (one rank) */
void orc_exchange_vflux(orc_field *F)
{
    double *arr[12];
    for (int a = 0; a < 12; ++a) arr[a] = F->a[ORC_VF0 + a];
    periodic_fill(F, arr, 12);
}

/* ----------------------------------------------------------- FVS splits */

/* epsilon-modified eigenvalue split of README.md:179; eps == 0 keeps the
 * plain |lambda| form of InviscousFlux_new.cpp:461-466 bit for bit. */
static inline void lam_split(double lam, double eps, double *lp, double *lm)
{
    if (eps == 0.0) {
        *lp = 0.5 * (lam + fabs(lam));
        *lm = 0.5 * (lam - fabs(lam));
    } else {
        double r = sqrt(lam * lam + eps * eps);
        *lp = 0.5 * (lam + r);
        *lm = 0.5 * (lam - r);
    }
}

void orc_split_state(int fvs, double gamma, double eps,
                     double rho, double u, double v, double w, double E,
                     double p, double c, int n1, int n2, int n3,
                     double Fp[5], double Fm[5])
{
    if (fvs == NS3D_FVS_STEGER_WARMING) {
        /* split_StagerWarming, InviscousFlux_new.cpp:439-489 */
        double un = u * n1 + v * n2 + w * n3;
        double lambda1 = un - c;
        double lambda2 = un;
        double lambda3 = un + c;
        double l1p, l1m, l2p, l2m, l3p, l3m;
        lam_split(lambda1, eps, &l1p, &l1m);
        lam_split(lambda2, eps, &l2p, &l2m);
        lam_split(lambda3, eps, &l3p, &l3m);
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
        /* split_LaxFriedrichs, InviscousFlux_new.cpp:491-530 */
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
    (void)gamma;
}

/* ------------------------------------------------------- reconstructions */

/* linear_reconstruction, reconstruction.cpp:88-90 */
static double recon_linear(const double *s) { return s[0]; }

/* weno5_reconstruction, reconstruction.cpp:93-124 */
static double recon_weno5(const double *s)
{
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

/* mdcd_reconstruction, reconstruction.cpp:126-135 */
static double recon_mdcd(const double *s, double disp, double diss)
{
    return (1.0 / 2.0 * disp + 1.0 / 2.0 * diss) * s[0] +
           (-3.0 / 2.0 * disp - 5.0 / 2.0 * diss - 1.0 / 12.0) * s[1] +
           (7.0 / 12.0 + disp + 5.0 * diss) * s[2] +
           (7.0 / 12.0 + disp - 5.0 * diss) * s[3] +
           (-3.0 / 2.0 * disp + 5.0 / 2.0 * diss - 1.0 / 12.0) * s[4] +
           (1.0 / 2.0 * disp - 1.0 / 2.0 * diss) * s[5];
}

/* mdcd_hybrid_reconstruction, reconstruction.cpp:151-225.  std::pow(x, 2)
 * there is folded by g++ to x*x, which is what is written here. */
static double recon_mdcd_hybrid(const double *s, double disp, double diss)
{
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

/* UpWind_7th_reconstruction, reconstruction.cpp:227-232 */
static double recon_ud7(const double *s)
{
    return (-1.0 / 140.0) * s[0] + (5.0 / 84.0) * s[1] - (101.0 / 420.0) * s[2]
           + (319.0 / 420.0) * s[3] + (107.0 / 210.0) * s[4]
           - (19.0 / 210.0) * s[5] + (1.0 / 105.0) * s[6];
}

/* ud7l_reconstruction, reconstruction.cpp:234-249 */
static double recon_ud7l(const double *s, double alpha)
{
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

/* reconstruct_select, reconstruction.cpp:13-85 */
double orc_reconstruct(const double *st, int n, double flag, const ns3d_params *P, int *err)
{
    double a[8];
    int need;
    switch (P->recon) {
    case NS3D_RECON_WENO5: need = 6; break;
    case NS3D_RECON_LINEAR: need = 2; break;
    case NS3D_RECON_MDCD: need = 6; break;
    case NS3D_RECON_MDCD_HYBRID: need = 6; break;
    case NS3D_RECON_UPWIND_7TH: need = 8; break;
    case NS3D_RECON_UD7L: need = 8; break;
    default:
        /* fallback of reconstruction.cpp:77-84 (reached for WCNS).  With two
         * values and flag < 0 the reference reads st[2], past the stencil
         * (undefined); that case is reported as an error, as the GPU entry
         * point refuses it. */
        if (n >= 2) {
            int mid = n / 2;
            if (flag >= 0.0) return 0.5 * (st[mid] + st[mid - 1]);
            if (mid + 1 >= n) {
                if (err) *err = 1;
                return NAN;
            }
            return 0.5 * (st[mid] + st[mid + 1]);
        }
        return n == 0 ? 0.0 : st[0];
    }
    if (n != need) {
        if (err) *err = 1;
        return NAN;
    }
    for (int i = 0; i < n; ++i) a[i] = st[i];
    if (flag < 0.0)
        for (int i = 0; i < n / 2; ++i) { double t = a[i]; a[i] = a[n - 1 - i]; a[n - 1 - i] = t; }
    switch (P->recon) {
    case NS3D_RECON_WENO5: return recon_weno5(a);
    case NS3D_RECON_LINEAR: return recon_linear(a);
    case NS3D_RECON_MDCD: return recon_mdcd(a, P->mdcd_disp, P->mdcd_diss);
    case NS3D_RECON_MDCD_HYBRID: return recon_mdcd_hybrid(a, P->mdcd_disp, P->mdcd_diss);
    case NS3D_RECON_UPWIND_7TH: return recon_ud7(a);
    default: return recon_ud7l(a, P->ud7l_alpha);
    }
}

/* --------------------------------------------------------- inviscid sweep */

static void split_cell(const orc_field *F, long long id, int n1, int n2, int n3,
                       double Fp[5], double Fm[5])
{
    orc_split_state(F->P.fvs_type, F->P.gamma, F->P.sw_eps,
                    F->a[ORC_RHO][id], F->a[ORC_U][id], F->a[ORC_V][id], F->a[ORC_W][id],
                    F->a[ORC_E][id], F->a[ORC_P][id], F->a[ORC_C][id], n1, n2, n3, Fp, Fm);
}

/* one pencil of du_inviscous (InviscousFlux_new.cpp:15-61, :65-111, :114-160):
 * split every cell of the pencil (ghosts included), then for each face
 * lo-1 .. lo+n-1 reconstruct F+ (flag +1) and F- (flag -1) per component. */
static void pencil_faces(const orc_field *F, int dir, int a, int b,
                         double (*Fp)[5], double (*Fm)[5], double *out[5])
{
    const int S = F->P.stencil, mid = (S - 1) / 2;
    const int ng = F->ng;
    const int len = dir == 0 ? F->sx : dir == 1 ? F->sy : F->sz;
    const int n = dir == 0 ? F->nx : dir == 1 ? F->ny : F->nz;
    const int n1 = dir == 0, n2 = dir == 1, n3 = dir == 2;
    for (int q = 0; q < len; ++q) {
        long long id = dir == 0 ? IDX(F, q, a, b) : dir == 1 ? IDX(F, a, q, b) : IDX(F, a, b, q);
        split_cell(F, id, n1, n2, n3, Fp[q], Fm[q]);
    }
    double wplus[8], wminus[8];
    for (int q = ng - 1; q < ng + n; ++q) {
        double wface[5];
        for (int m5 = 0; m5 < 5; ++m5) {
            for (int m = 0; m < S; ++m) {
                wplus[m] = Fp[q + (m - mid)][m5];
                wminus[m] = Fm[q + (m - mid)][m5];
            }
            double plus_face = orc_reconstruct(wplus, S, +1.0, &F->P, NULL);
            double minus_face = orc_reconstruct(wminus, S, -1.0, &F->P, NULL);
            wface[m5] = plus_face + minus_face;
        }
        long long fid = dir == 0 ? IDXFX(F, q, a, b) : dir == 1 ? IDXFY(F, a, q, b) : IDXFZ(F, a, b, q);
        for (int m5 = 0; m5 < 5; ++m5) out[m5][fid] = wface[m5];
    }
}

/* du_inviscous, InviscousFlux_new.cpp:4-163 */
void orc_du_inviscous(orc_field *F)
{
    const int ng = F->ng;
    int smax = F->sx > F->sy ? F->sx : F->sy;
    if (F->sz > smax) smax = F->sz;
    for (int dir = 0; dir < 3; ++dir) {
        double *out[5];
        for (int m = 0; m < 5; ++m) out[m] = F->a[(dir == 0 ? ORC_FX0 : dir == 1 ? ORC_FY0 : ORC_FZ0) + m];
        /* outer/inner pencil loops: x: (k,j), y: (k,i), z: (j,i) */
        const int nb = dir == 2 ? F->ny : F->nz;
        const int na = dir == 0 ? F->ny : F->nx;
#pragma omp parallel
        {
            double (*Fp)[5] = (double (*)[5])malloc(sizeof(double[5]) * (size_t)smax);
            double (*Fm)[5] = (double (*)[5])malloc(sizeof(double[5]) * (size_t)smax);
#pragma omp for schedule(static)
            for (int bb = ng; bb < ng + nb; ++bb)
                for (int aa = ng; aa < ng + na; ++aa)
                    pencil_faces(F, dir, aa, bb, Fp, Fm, out);
            free(Fp);
            free(Fm);
        }
    }
}

/* computeRoeAveragedState_new, InviscousFlux_new.cpp:628-659 */
static void roe_average(double u_roe[5], const double ul[5], const double ur[5], double gamma)
{
    double rho_L = ul[0], u_L = ul[1], v_L = ul[2], w_L = ul[3], p_L = ul[4];
    double E_L = rho_L * (0.5 * (u_L * u_L + v_L * v_L + w_L * w_L) + p_L / ((gamma - 1.0) * rho_L));
    double rho_R = ur[0], u_R = ur[1], v_R = ur[2], w_R = ur[3], p_R = ur[4];
    double E_R = rho_R * (0.5 * (u_R * u_R + v_R * v_R + w_R * w_R) + p_R / ((gamma - 1.0) * rho_R));
    double sqrt_rho_L = sqrt(rho_L);
    double sqrt_rho_R = sqrt(rho_R);
    u_roe[0] = sqrt_rho_L * sqrt_rho_R;
    double denom = 1.0 / (sqrt_rho_L + sqrt_rho_R);
    u_roe[1] = (sqrt_rho_L * u_L + sqrt_rho_R * u_R) * denom;
    u_roe[2] = (sqrt_rho_L * v_L + sqrt_rho_R * v_R) * denom;
    u_roe[3] = (sqrt_rho_L * w_L + sqrt_rho_R * w_R) * denom;
    double H_L = (E_L + p_L) / rho_L;
    double H_R = (E_R + p_R) / rho_R;
    u_roe[4] = (sqrt_rho_L * H_L + sqrt_rho_R * H_R) * denom;
}

/* build_eigen_matrices_new, InviscousFlux_new.cpp:532-625 (its permuted
 * tangential basis included) */
static void eigen_matrices(const double u[5], double nx, double ny, double nz, double gamma,
                           double Lm[5][5], double Rm[5][5])
{
    double u0 = u[1], v0 = u[2], w0 = u[3], h0 = u[4];
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

/* the Roe average and eigen-matrices above, exported for their tests:
 * ul, ur = rho, u, v, w, p; L and R row-major 5x5 */
void orc_char_matrices(const double ul[5], const double ur[5], int n1, int n2, int n3, double gamma,
                       double L[25], double R[25])
{
    double uref[5], Lm[5][5], Rm[5][5];
    roe_average(uref, ul, ur, gamma);
    eigen_matrices(uref, n1, n2, n3, gamma, Lm, Rm);
    for (int r = 0; r < 5; ++r)
        for (int q = 0; q < 5; ++q) { L[5 * r + q] = Lm[r][q]; R[5 * r + q] = Rm[r][q]; }
}

static void pencil_faces_char(const orc_field *F, int dir, int a, int b,
                              double (*Fp)[5], double (*Fm)[5], double *out[5])
{
    const int S = F->P.stencil, mid = (S - 1) / 2;
    const int ng = F->ng;
    const int len = dir == 0 ? F->sx : dir == 1 ? F->sy : F->sz;
    const int n = dir == 0 ? F->nx : dir == 1 ? F->ny : F->nz;
    const int n1 = dir == 0, n2 = dir == 1, n3 = dir == 2;
    for (int q = 0; q < len; ++q) {
        long long id = dir == 0 ? IDX(F, q, a, b) : dir == 1 ? IDX(F, a, q, b) : IDX(F, a, b, q);
        split_cell(F, id, n1, n2, n3, Fp[q], Fm[q]);
    }
    for (int q = ng - 1; q < ng + n; ++q) {
        double LFp[5][8], LFm[5][8], wchar[5], wface[5], Lm[5][5], Rm[5][5], ul[5], ur[5], uref[5];
        long long il = dir == 0 ? IDX(F, q, a, b) : dir == 1 ? IDX(F, a, q, b) : IDX(F, a, b, q);
        long long ir = dir == 0 ? IDX(F, q + 1, a, b) : dir == 1 ? IDX(F, a, q + 1, b) : IDX(F, a, b, q + 1);
        const int arr[5] = {ORC_RHO, ORC_U, ORC_V, ORC_W, ORC_P};
        for (int m = 0; m < 5; ++m) { ul[m] = F->a[arr[m]][il]; ur[m] = F->a[arr[m]][ir]; }
        roe_average(uref, ul, ur, F->P.gamma);
        eigen_matrices(uref, n1, n2, n3, F->P.gamma, Lm, Rm);
        for (int m = 0; m < S; ++m) {
            int ii = q + (m - mid);
            for (int nn = 0; nn < 5; ++nn) {
                LFp[nn][m] = 0.0;
                LFm[nn][m] = 0.0;
                for (int r = 0; r < 5; ++r) {
                    LFp[nn][m] += Lm[nn][r] * Fp[ii][r];
                    LFm[nn][m] += Lm[nn][r] * Fm[ii][r];
                }
            }
        }
        for (int nn = 0; nn < 5; ++nn) {
            double pf = orc_reconstruct(LFp[nn], S, +1.0, &F->P, NULL);
            double mf = orc_reconstruct(LFm[nn], S, -1.0, &F->P, NULL);
            wchar[nn] = pf + mf;
        }
        for (int nn = 0; nn < 5; ++nn) {
            double sum = 0.0;
            for (int r = 0; r < 5; ++r) sum += Rm[nn][r] * wchar[r];
            wface[nn] = sum;
        }
        long long fid = dir == 0 ? IDXFX(F, q, a, b) : dir == 1 ? IDXFY(F, a, q, b) : IDXFZ(F, a, b, q);
        for (int m5 = 0; m5 < 5; ++m5) out[m5][fid] = wface[m5];
    }
}

/* du_inviscous_character, InviscousFlux_new.cpp:165-437 */
void orc_du_inviscous_character(orc_field *F)
{
    const int ng = F->ng;
    int smax = F->sx > F->sy ? F->sx : F->sy;
    if (F->sz > smax) smax = F->sz;
    for (int dir = 0; dir < 3; ++dir) {
        double *out[5];
        for (int m = 0; m < 5; ++m) out[m] = F->a[(dir == 0 ? ORC_FX0 : dir == 1 ? ORC_FY0 : ORC_FZ0) + m];
        const int nb = dir == 2 ? F->ny : F->nz;
        const int na = dir == 0 ? F->ny : F->nx;
#pragma omp parallel
        {
            double (*Fp)[5] = (double (*)[5])malloc(sizeof(double[5]) * (size_t)smax);
            double (*Fm)[5] = (double (*)[5])malloc(sizeof(double[5]) * (size_t)smax);
#pragma omp for schedule(static)
            for (int bb = ng; bb < ng + nb; ++bb)
                for (int aa = ng; aa < ng + na; ++aa)
                    pencil_faces_char(F, dir, aa, bb, Fp, Fm, out);
            free(Fp);
            free(Fm);
        }
    }
}

/* ---------------------------------------------------------------- viscous */

/* ViscousFlux.cpp:15-40 on a window centred at f[4] */
double orc_diff(const double *f9, int order, double dx)
{
    const double *f = f9 + 4;
    if (order == 1 || order == -1) {
        int flag = order;
        return (-3.0 * f[0] + 4.0 * f[flag] - f[2 * flag]) / (2.0 * dx);
    }
    if (order == 2) return (f[1] - f[-1]) / (2.0 * dx);
    if (order == 4) return (-f[2] + 8.0 * f[1] - 8.0 * f[-1] + f[-2]) / (12.0 * dx);
    if (order == 6)
        return (f[3] - 9.0 * f[2] + 45.0 * f[1] - 45.0 * f[-1] + 9.0 * f[-2] - f[-3]) / (60.0 * dx);
    {
        double c1 = 0.8;
        double c2 = -0.2;
        double c3 = 0.0380952380952381;
        double c4 = -0.00357142857142857;
        return (c1 * (f[1] - f[-1]) +
                c2 * (f[2] - f[-2]) +
                c3 * (f[3] - f[-3]) +
                c4 * (f[4] - f[-4])) / dx;
    }
}

/* choose_scheme_C4/C6/C8 + choose_scheme, ViscousFlux.cpp:43-118 */
int orc_choose_scheme(int idx, int nstart, int nend, int periodic, int vis)
{
    int full = vis == NS3D_VIS_C8TH ? 8 : vis == NS3D_VIS_C6TH ? 6 : vis == NS3D_VIS_C4TH ? 4 : 0;
    if (full == 0) return 0;
    if (periodic) return full;
    int dist_left = idx - nstart;
    if (dist_left == 0) return 1;
    int dist_right = (nend - 1) - idx;
    if (dist_right == 0) return -1;
    int dmin = dist_left < dist_right ? dist_left : dist_right;
    if (dmin == 1) return 2;
    if (full == 4) return 4;
    if (dmin == 2) return 4;
    if (full == 6) return 6;
    if (dmin == 3) return 6;
    return 8;
}

/* diff_x / diff_y / diff_z, ViscousFlux.cpp:120-235: gather the 9-point window
 * along the axis (only the points the chosen order reads are touched). */
static double diff_axis(const orc_field *F, const double *f, int i, int j, int k,
                        int axis, double h, int order)
{
    double win[9];
    int reach = (order == 1 || order == -1) ? 2 : order == 2 ? 1 : order / 2;
    for (int m = -4; m <= 4; ++m) {
        if (m < -reach || m > reach) { win[m + 4] = 0.0; continue; }
        long long id = axis == 0 ? IDX(F, i + m, j, k) : axis == 1 ? IDX(F, i, j + m, k) : IDX(F, i, j, k + m);
        win[m + 4] = f[id];
    }
    return orc_diff(win, order, h);
}

/* compute_gradients, ViscousFlux.cpp:240-279 */
void orc_compute_gradients(orc_field *F)
{
    const int ng = F->ng, nx = F->nx, ny = F->ny, nz = F->nz;
    const double dx = F->G.dx, dy = F->G.dy, dz = F->G.dz;
    const int px = axis_periodic(F, 0), py = axis_periodic(F, 1), pz = axis_periodic(F, 2);
    const int vis = F->P.vis_scheme;
    const double *src[4] = {F->a[ORC_U], F->a[ORC_V], F->a[ORC_W], F->a[ORC_T]};
#pragma omp parallel for schedule(static)
    for (int k = ng; k < ng + nz; ++k)
        for (int j = ng; j < ng + ny; ++j)
            for (int i = ng; i < ng + nx; ++i) {
                long long id = IDX(F, i, j, k);
                int ox = orc_choose_scheme(i, ng, ng + nx, px, vis);
                int oy = orc_choose_scheme(j, ng, ng + ny, py, vis);
                int oz = orc_choose_scheme(k, ng, ng + nz, pz, vis);
                for (int q = 0; q < 4; ++q) {
                    F->a[ORC_GRAD0 + 3 * q + 0][id] = diff_axis(F, src[q], i, j, k, 0, dx, ox);
                    F->a[ORC_GRAD0 + 3 * q + 1][id] = diff_axis(F, src[q], i, j, k, 1, dy, oy);
                    F->a[ORC_GRAD0 + 3 * q + 2][id] = diff_axis(F, src[q], i, j, k, 2, dz, oz);
                }
            }
}

/* compute_viscous_flux, ViscousFlux.cpp:282-339.  Output order:
 * Fv momx,momy,momz,E | Hv momx,momy,momz,E | Gv momx,momy,momz,E
 * (the *_mass components are identically 0 and not stored). */
void orc_compute_viscous_flux(orc_field *F)
{
    const int ng = F->ng, nx = F->nx, ny = F->ny, nz = F->nz;
    const double Pr = F->P.Pr, Cp = F->P.Cp;
    double *const *g = &F->a[ORC_GRAD0];
    double *const *vf = &F->a[ORC_VF0];
#pragma omp parallel for schedule(static)
    for (int k = ng; k < ng + nz; ++k)
        for (int j = ng; j < ng + ny; ++j)
            for (int i = ng; i < ng + nx; ++i) {
                long long id = IDX(F, i, j, k);
                double mu = orc_get_mu(&F->P, F->a[ORC_T][id]);
                double du_dx = g[0][id], du_dy = g[1][id], du_dz = g[2][id];
                double dv_dx = g[3][id], dv_dy = g[4][id], dv_dz = g[5][id];
                double dw_dx = g[6][id], dw_dy = g[7][id], dw_dz = g[8][id];
                double dT_dx = g[9][id], dT_dy = g[10][id], dT_dz = g[11][id];
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
                double u = F->a[ORC_U][id], v = F->a[ORC_V][id], w = F->a[ORC_W][id];
                vf[0][id] = tau_xx;
                vf[1][id] = tau_xy;
                vf[2][id] = tau_xz;
                vf[3][id] = u * tau_xx + v * tau_xy + w * tau_xz + qx;
                vf[4][id] = tau_xy;
                vf[5][id] = tau_yy;
                vf[6][id] = tau_yz;
                vf[7][id] = u * tau_xy + v * tau_yy + w * tau_yz + qy;
                vf[8][id] = tau_xz;
                vf[9][id] = tau_yz;
                vf[10][id] = tau_zz;
                vf[11][id] = u * tau_xz + v * tau_yz + w * tau_zz + qz;
            }
}

/* compute_vis_flux, ViscousFlux.cpp:341-386 */
void orc_compute_vis_flux(orc_field *F)
{
    const int ng = F->ng, nx = F->nx, ny = F->ny, nz = F->nz;
    const double dx = F->G.dx, dy = F->G.dy, dz = F->G.dz;
    const int px = axis_periodic(F, 0), py = axis_periodic(F, 1), pz = axis_periodic(F, 2);
    const int vis = F->P.vis_scheme;
#pragma omp parallel for schedule(static)
    for (int k = ng; k < ng + nz; ++k)
        for (int j = ng; j < ng + ny; ++j)
            for (int i = ng; i < ng + nx; ++i) {
                long long id = IDX(F, i, j, k);
                int ox = orc_choose_scheme(i, ng, ng + nx, px, vis);
                int oy = orc_choose_scheme(j, ng, ng + ny, py, vis);
                int oz = orc_choose_scheme(k, ng, ng + nz, pz, vis);
                for (int m = 0; m < 4; ++m) {
                    double *r = F->a[ORC_RHS1 + m];
                    r[id] += diff_axis(F, F->a[ORC_VF0 + m], i, j, k, 0, dx, ox);
                    r[id] += diff_axis(F, F->a[ORC_VF0 + 4 + m], i, j, k, 1, dy, oy);
                    r[id] += diff_axis(F, F->a[ORC_VF0 + 8 + m], i, j, k, 2, dz, oz);
                }
            }
}

/* ------------------------------------------------------------- RHS / RK3 */

/* compute_rhs, range_kutta.cpp:11-99 */
void orc_compute_rhs(orc_field *F)
{
    const int ng = F->ng, nx = F->nx, ny = F->ny, nz = F->nz;
    const double idx = 1.0 / F->G.dx;
    const double idy = 1.0 / F->G.dy;
    const double idz = 1.0 / F->G.dz;
    for (int m = 0; m < 5; ++m) memset(F->a[ORC_RHS0 + m], 0, (size_t)F->ncell * sizeof(double));
    if (F->P.char_recon) orc_du_inviscous_character(F);
    else orc_du_inviscous(F);
    orc_compute_gradients(F);
    orc_compute_viscous_flux(F);
    orc_exchange_vflux(F);
    orc_compute_vis_flux(F);
#pragma omp parallel for schedule(static)
    for (int k = ng; k < ng + nz; ++k)
        for (int j = ng; j < ng + ny; ++j)
            for (int i = ng; i < ng + nx; ++i) {
                long long id = IDX(F, i, j, k);
                for (int m = 0; m < 5; ++m) {
                    double fx_r = F->a[ORC_FX0 + m][IDXFX(F, i, j, k)];
                    double fx_l = F->a[ORC_FX0 + m][IDXFX(F, i - 1, j, k)];
                    double fy_t = F->a[ORC_FY0 + m][IDXFY(F, i, j, k)];
                    double fy_b = F->a[ORC_FY0 + m][IDXFY(F, i, j - 1, k)];
                    double fz_f = F->a[ORC_FZ0 + m][IDXFZ(F, i, j, k)];
                    double fz_b = F->a[ORC_FZ0 + m][IDXFZ(F, i, j, k - 1)];
                    F->a[ORC_RHS0 + m][id] += -((fx_r - fx_l) * idx + (fy_t - fy_b) * idy + (fz_f - fz_b) * idz);
                }
            }
}

/* runge_kutta_3, range_kutta.cpp:102-163 */
int orc_rk3_step(orc_field *F, double dt, long long *bad_id)
{
    const long long N = F->ncell;
    double *U[5], *U0[5], *R[5];
    for (int m = 0; m < 5; ++m) {
        U[m] = F->a[ORC_RHO + m];
        U0[m] = F->a[ORC_RHO0 + m];
        R[m] = F->a[ORC_RHS0 + m];
    }
    for (int stage = 1; stage <= 3; ++stage) {
        orc_compute_rhs(F);
#pragma omp parallel for schedule(static)
        for (long long n = 0; n < N; ++n) {
            for (int m = 0; m < 5; ++m) {
                if (stage == 1) U[m][n] = U0[m][n] + dt * R[m][n];
                else if (stage == 2) U[m][n] = 0.75 * U0[m][n] + 0.25 * (U[m][n] + dt * R[m][n]);
                else U[m][n] = (1.0 / 3.0) * U0[m][n] + (2.0 / 3.0) * (U[m][n] + dt * R[m][n]);
            }
        }
        orc_c2p(F);
        orc_apply_boundary(F);
        orc_p2c(F);
    }
    /* range_kutta.cpp:150-162 (first offending cell in loop order) */
    const int ng = F->ng;
    for (int k = ng; k < ng + F->nz; ++k)
        for (int j = ng; j < ng + F->ny; ++j)
            for (int i = ng; i < ng + F->nx; ++i) {
                long long id = IDX(F, i, j, k);
                if (F->a[ORC_RHO][id] <= 0.0 || F->a[ORC_P][id] <= 0.0 || F->a[ORC_T][id] <= 0.0) {
                    if (bad_id) *bad_id = id;
                    return 1;
                }
            }
    if (bad_id) *bad_id = -1;
    return 0;
}

/* compute_timestep, timestep.cpp:5-33 (single rank: the Allreduce is the
 * identity) */
double orc_compute_timestep(orc_field *F)
{
    double dt_local = 1e9;
    const int ng = F->ng;
    const double dx = F->G.dx, dy = F->G.dy, dz = F->G.dz;
    for (int k = ng; k < ng + F->nz; ++k)
        for (int j = ng; j < ng + F->ny; ++j)
            for (int i = ng; i < ng + F->nx; ++i) {
                long long id = IDX(F, i, j, k);
                double rho = F->a[ORC_RHO][id];
                double u = F->a[ORC_U][id], v = F->a[ORC_V][id], w = F->a[ORC_W][id];
                double p = F->a[ORC_P][id];
                double a = sqrt(F->P.gamma * p / (rho > 1e-14 ? rho : 1e-14));
                double dt_x = dx / (fabs(u) + a);
                double dt_y = dy / (fabs(v) + a);
                double dt_z = dz / (fabs(w) + a);
                double m = dt_y < dt_z ? dt_y : dt_z;
                m = dt_x < m ? dt_x : m;
                dt_local = dt_local < m ? dt_local : m;
            }
    return F->P.cfl * dt_local;
}

/* the body of one time_advance iteration, maincycle.cpp:20-56, without the
 * output-time clipping (callers run with TotalTime >> steps*dt) */
int orc_step(orc_field *F, double dt, double *dt_used, long long *bad_id)
{
    orc_record_u0(F);
    if (!(dt > 0.0)) dt = orc_compute_timestep(F);
    if (dt_used) *dt_used = dt;
    return orc_rk3_step(F, dt, bad_id);
}

/* compute_total_energy, postprocess.cpp:103-128 (serial sum in loop order) */
double orc_total_energy(orc_field *F)
{
    const double cell_vol = F->G.dx * F->G.dy * F->G.dz;
    double local_sum = 0.0;
    const int ng = F->ng;
    for (int k = ng; k < ng + F->nz; ++k)
        for (int j = ng; j < ng + F->ny; ++j)
            for (int i = ng; i < ng + F->nx; ++i) {
                long long id = IDX(F, i, j, k);
                double rho = F->a[ORC_RHO][id];
                double u = F->a[ORC_U][id], v = F->a[ORC_V][id], w = F->a[ORC_W][id];
                double kinetic = 0.5 * rho * (u * u + v * v + w * w);
                double eint = 0;
                local_sum += (kinetic + eint) * cell_vol;
            }
    return local_sum;
}
