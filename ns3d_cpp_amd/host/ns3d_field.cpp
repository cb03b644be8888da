// ns3d_field.cpp -- the reference call surface over the C ABI (see header).
#include "ns3d_field.hpp"

#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>

namespace ns3d_host {

void check(int st, const char *what, const Field3D *F)
{
    if (st == NS3D_OK) return;
    std::string msg = std::string(what) + ": " + ns3d_status_string(st);
    if (F && F->ctx) {
        const char *e = ns3d_last_error(F->ctx);
        if (e && *e) msg += std::string(" (") + e + ")";
        if (st == NS3D_ERR_NONPHYSICAL) {
            int i, j, k;
            ns3d_nonphysical_cell(F->ctx, &i, &j, &k);
            msg += " at cell (" + std::to_string(i) + "," + std::to_string(j) + "," + std::to_string(k) + ")";
        }
    }
    throw Error(st, msg);
}

Field3D::~Field3D()
{
    if (ctx) ns3d_destroy(ctx);
}

void Field3D::allocate(const SolverParams &P, const GridDesc &G)
{
    if (ctx) { ns3d_destroy(ctx); ctx = nullptr; }
    check(ns3d_create(&P, &G, &ctx), "ns3d_create");
    check(ns3d_sizes(ctx, &sx, &sy, &sz, &ng), "ns3d_sizes", this);
    nx = G.nx; ny = G.ny; nz = G.nz;
    const size_t tot = (size_t)sx * sy * sz;
    rho.assign(tot, 1.0);   // field_structures.h:232
    for (auto *a : {&u, &v, &w, &p, &T, &rhou, &rhov, &rhow, &E}) a->assign(tot, 0.0);
}

void Field3D::upload_primitive()
{
    check(ns3d_upload_primitive(ctx, rho.data(), u.data(), v.data(), w.data(), p.data()), "upload_primitive", this);
}

void Field3D::sync_to_host()
{
    check(ns3d_download_primitive(ctx, rho.data(), u.data(), v.data(), w.data(), p.data(), T.data()),
          "download_primitive", this);
    std::vector<double> r0(rho.size());
    check(ns3d_download_conserved(ctx, r0.data(), rhou.data(), rhov.data(), rhow.data(), E.data()),
          "download_conserved", this);
}

void apply_boundary(Field3D &F, GridDesc &, const SolverParams &)
{
    check(ns3d_apply_boundary(F.ctx), "apply_boundary", &F);
}

void du_inviscous(Field3D &F, const SolverParams &)
{
    check(ns3d_du_inviscous(F.ctx), "du_inviscous", &F);
}

void compute_gradients(Field3D &F, const GridDesc &, const SolverParams &)
{
    check(ns3d_compute_gradients(F.ctx), "compute_gradients", &F);
}

void compute_viscous_flux(Field3D &F, const SolverParams &)
{
    check(ns3d_compute_viscous_flux(F.ctx), "compute_viscous_flux", &F);
}

void compute_rhs(Field3D &F, GridDesc &, SolverParams &)
{
    check(ns3d_compute_rhs(F.ctx), "compute_rhs", &F);
}

void runge_kutta_3(Field3D &F, GridDesc &, SolverParams &, double dt)
{
    // the reference aborts the MPI job on a non-physical cell
    // (range_kutta.cpp:150-161); here the status becomes an exception
    check(ns3d_rk3_step(F.ctx, dt), "runge_kutta_3", &F);
}

double compute_timestep(Field3D &F, const GridDesc &, const SolverParams &)
{
    double dt = 0.0;
    check(ns3d_compute_timestep_local(F.ctx, &dt), "compute_timestep", &F);
    return dt;   // a multi-rank caller takes the MPI_MIN of this value
}

void compute_total_energy(Field3D &F, const GridDesc &, const SolverParams &)
{
    check(ns3d_total_energy(F.ctx, &F.global_Etot), "total_energy", &F);
}

void time_advance(Field3D &F, GridDesc &G, SolverParams &P)
{
    using clk = std::chrono::steady_clock;
    const auto t_start = clk::now();
    auto t_last = t_start;
    double current_time = 0.0;
    double output_Timefreq = P.output_timefreq;
    const double TotalTime = P.total_time;
    for (int step = 1; step <= P.max_steps; ++step) {
        double dt = P.dt_fixed;
        const double eps = 1e-12;
        if (std::abs(TotalTime - current_time) <= eps) break;
        if (P.dt_fixed < 0.0) dt = compute_timestep(F, G, P);
        if (current_time >= output_Timefreq - eps) output_Timefreq += P.output_timefreq;
        if (current_time + dt > TotalTime) dt = TotalTime - current_time;
        if (current_time + dt > output_Timefreq) {
            dt = output_Timefreq - current_time;
            output_Timefreq += P.output_timefreq;
        }
        if (dt <= eps) {
            std::cerr << "dt too small or zero; aborting to avoid stall at Time=" << current_time << "\n";
            break;
        }
        current_time += dt;
        runge_kutta_3(F, G, P, dt);
        if (step % P.monitor_stepfreq == 0 || step == 1) {
            compute_total_energy(F, G, P);
            const auto t_now = clk::now();
            const double t_elapsed = std::chrono::duration<double>(t_now - t_start).count();
            const double t_step = std::chrono::duration<double>(t_now - t_last).count();
            t_last = t_now;
            std::cout << std::fixed << std::setprecision(6) << "[Step " << step << "] "
                      << "dt=" << dt << "  Time/step=" << t_step << "s  Elapsed=" << t_elapsed << "s"
                      << "  Time=" << current_time << "/" << TotalTime << "  Ek_avg=" << F.global_Etot << "\n";
        }
        if (current_time >= TotalTime) break;
    }
    const double total = std::chrono::duration<double>(clk::now() - t_start).count();
    std::cout << "Simulation complete. Total time: " << total << " s\n";
}

}  // namespace ns3d_host
