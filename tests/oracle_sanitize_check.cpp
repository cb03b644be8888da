// Exercises every entry point of the CPU oracle (oracle/ns3d_oracle.c) over
// the option surface on small odd-sized boxes, for a build under
// AddressSanitizer and UndefinedBehaviorSanitizer (test_sanitizers.py).
// Prints "oracle ok" and exits 0 when every call returned; a sanitizer
// report aborts the run.  CPU only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/ns3d_gpu.h"
#include "../oracle/ns3d_oracle.h"

static void fill_state(orc_field *F, int seed)
{
    int sx, sy, sz, ng;
    orc_sizes(F, &sx, &sy, &sz, &ng);
    double *rho = orc_array(F, ORC_RHO), *u = orc_array(F, ORC_U), *v = orc_array(F, ORC_V);
    double *w = orc_array(F, ORC_W), *p = orc_array(F, ORC_P);
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                const long long id = ((long long)k * sy + j) * sx + i;
                const double a = 0.7 * i + 1.3 * j - 0.9 * k + seed;
                rho[id] = 1.0 + 0.1 * std::sin(a);
                u[id] = 0.4 * std::cos(1.1 * a);
                v[id] = -0.3 * std::sin(0.7 * a + 1.0);
                w[id] = 0.2 * std::cos(0.5 * a - 2.0) + ((i + j + k) % 5 == 0 ? 0.3 : 0.0);
                p[id] = 20.0 + std::cos(0.3 * a);
            }
}

static int run_config(int recon, int fvs, int vis, int chr, const int bc[6], double sw_eps, int seed)
{
    ns3d_params P;
    ns3d_grid G;
    ns3d_params_default(&P, &G);
    G.nx = 9;
    G.ny = 8;
    G.nz = 7;
    G.Lx = G.Ly = G.Lz = 2.0 * M_PI;
    P.recon = recon;
    P.fvs_type = fvs;
    P.vis_scheme = vis;
    P.char_recon = chr;
    P.sw_eps = sw_eps;
    P.Ma = 0.3;
    P.Re = 100.0;
    P.cfl = 0.4;
    P.mu_law = seed % 3;
    P.mu_power = 0.76;
    for (int f = 0; f < 6; ++f) P.bc[f] = bc[f];
    if (ns3d_params_finalize(&P, &G) != 0) return 1;
    orc_params_finalize(&P, &G);
    orc_field *F = orc_create(&P, &G);
    if (!F) return 1;
    fill_state(F, seed);
    orc_apply_boundary(F);
    orc_p2c(F);
    orc_c2p(F);
    orc_apply_boundary(F);
    orc_p2c(F);
    if (chr) orc_du_inviscous_character(F);
    else orc_du_inviscous(F);
    orc_compute_gradients(F);
    orc_compute_viscous_flux(F);
    orc_exchange_vflux(F);
    orc_compute_vis_flux(F);
    orc_compute_rhs(F);
    double dt = orc_compute_timestep(F), used = 0.0;
    long long bad = 0;
    int st = orc_step(F, 0.0, &used, &bad);
    st |= orc_step(F, 0.5 * dt, &used, &bad);
    orc_record_u0(F);
    st |= orc_rk3_step(F, 0.5 * dt, &bad);
    const double ke = orc_total_energy(F);
    orc_destroy(F);
    return st != 0 || !(ke > 0.0);
}

int main()
{
    const int periodic[6] = {0, 0, 0, 0, 0, 0};
    const int walls[6] = {NS3D_BC_WALL, NS3D_BC_WALL, NS3D_BC_OUTFLOW, NS3D_BC_OUTFLOW, NS3D_BC_SYMMETRY,
                          NS3D_BC_SYMMETRY};
    const int inout[6] = {NS3D_BC_INFLOW, NS3D_BC_OUTFLOW, 0, 0, NS3D_BC_WALL, NS3D_BC_WALL};
    const int *bcs[3] = {periodic, walls, inout};
    const int recons[6] = {NS3D_RECON_WENO5, NS3D_RECON_MDCD, NS3D_RECON_MDCD_HYBRID, NS3D_RECON_UPWIND_7TH,
                           NS3D_RECON_UD7L, NS3D_RECON_LINEAR};
    int fails = 0, n = 0;
    for (int r = 0; r < 6; ++r)
        for (int f = 0; f < 2; ++f)
            for (int b = 0; b < 3; ++b) {
                const int vis = (r + b) % 3, chr = (r + f) % 2;
                const double eps = f == 0 && r % 2 ? 0.05 : 0.0;
                fails += run_config(recons[r], f, vis, chr, bcs[b], eps, n);
                ++n;
            }
    // point functions, the stencil-size error path included
    ns3d_params P;
    ns3d_grid G;
    ns3d_params_default(&P, &G);
    ns3d_params_finalize(&P, &G);
    double st8[8] = {1.0, 1.2, 0.9, 1.4, 0.2, 0.1, 0.3, 0.25};
    for (int r = 0; r < 6; ++r) {
        P.recon = recons[r];
        for (int width : {2, 6, 8, 5}) {
            int err = 0;
            double val = orc_reconstruct(st8, width, r % 2 ? -1.0 : 1.0, &P, &err);
            if (!err && !std::isfinite(val)) ++fails;
        }
    }
    double ul[5] = {1.0, 0.3, -0.2, 0.1, 2.0}, ur[5] = {1.2, 0.1, 0.2, -0.3, 2.5}, L[25], R[25];
    for (int axis = 0; axis < 3; ++axis) orc_char_matrices(ul, ur, axis == 0, axis == 1, axis == 2, 1.4, L, R);
    double Fp[5], Fm[5];
    for (int fvs = 0; fvs < 2; ++fvs) orc_split_state(fvs, 1.4, 0.1, 1.0, 0.3, 0.1, -0.2, 50.0, 20.0, 5.0, 0, 1, 0, Fp, Fm);
    double f9[9] = {0.0, 0.1, 0.4, 0.9, 1.6, 2.5, 3.6, 4.9, 6.4};
    for (int o : {1, -1, 2, 4, 6, 8}) (void)orc_diff(f9, o, 0.1);
    for (int idx = 0; idx < 12; ++idx) (void)orc_choose_scheme(idx, 0, 11, idx % 2, NS3D_VIS_C8TH);
    for (int law = 0; law < 3; ++law) {
        P.mu_law = law;
        (void)orc_get_mu(&P, 1.3);
    }
    std::printf("oracle ok: %d configurations, %d failures\n", n, fails);
    return fails != 0;
}
