// ns3d_field.cpp -- the reference call surface over the C ABI (see header).
#include "ns3d_field.hpp"
#include "ns3d_io.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
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

void du_inviscous_character(Field3D &F, const SolverParams &)
{
    check(ns3d_du_inviscous_character(F.ctx), "du_inviscous_character", &F);
}

void compute_vis_flux(Field3D &F, const GridDesc &, const SolverParams &)
{
    check(ns3d_compute_vis_flux(F.ctx), "compute_vis_flux", &F);
}

namespace {

void split_point(const Field3D &F, std::vector<double> &Fp, std::vector<double> &Fm, int id, int n1, int n2, int n3,
                 SolverParams P, int fvs)
{
    P.fvs_type = fvs;
    const double *W[5] = {&F.rho[(size_t)id], &F.u[(size_t)id], &F.v[(size_t)id], &F.w[(size_t)id],
                          &F.p[(size_t)id]};
    Fp.resize(5);
    Fm.resize(5);
    double *fp[5], *fm[5];
    for (int m = 0; m < 5; ++m) { fp[m] = &Fp[(size_t)m]; fm[m] = &Fm[(size_t)m]; }
    check(ns3d_split_batch(&P, 1, W, n1, n2, n3, fp, fm), "split");
}

}  // namespace

void split_StagerWarming(const Field3D &F, std::vector<double> &F_plus, std::vector<double> &F_minus, int id, int n1,
                         int n2, int n3, const SolverParams &P)
{
    split_point(F, F_plus, F_minus, id, n1, n2, n3, P, NS3D_FVS_STEGER_WARMING);
}

void split_LaxFriedrichs(const Field3D &F, std::vector<double> &F_plus, std::vector<double> &F_minus, int id, int n1,
                         int n2, int n3, const SolverParams &P)
{
    split_point(F, F_plus, F_minus, id, n1, n2, n3, P, NS3D_FVS_LAX_FRIEDRICHS);
}

double reconstruct_select(const std::vector<double> &vstencil, double flag, const SolverParams P)
{
    double out = 0.0;
    const int st = ns3d_reconstruct_batch(&P, 1, (int)vstencil.size(), vstencil.data(), flag, &out);
    if (st == NS3D_ERR_CONFIG)
        throw Error(st, "reconstruct_select: stencil of " + std::to_string(vstencil.size()) +
                            " values does not match the reconstruction");
    check(st, "reconstruct_select");
    return out;
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

namespace {

ZoneDesc zone_of(const Field3D &F, const GridDesc &G)
{
    return ZoneDesc{F.nx, F.ny, F.nz, F.ng, F.sx, F.sy, G.x0, G.y0, G.z0, G.dx, G.dy, G.dz};
}

}  // namespace

void compute_diagnostics(Field3D &F, const SolverParams &, const GridDesc &)
{
    check(ns3d_residuals(F.ctx, F.global_res), "residuals", &F);
}

void write_residuals_tecplot(const Field3D &F, int step, const std::string &filename)
{
    write_residual_line("output/" + filename, step, F.global_res, F.global_Etot);
}

void compute_energy_spectrum(Field3D &F, const GridDesc &G, const std::string &filename)
{
    std::vector<double> Ek((size_t)(2 * G.nx + 1));   // Kmax = 2 NX (EnergySpectrum.cpp:108)
    check(ns3d_energy_spectrum(F.ctx, Ek.data(), (int)Ek.size()), "energy_spectrum", &F);
    write_spectrum(filename, Ek);
    std::cout << "[Spectrum] Energy spectrum computed.\n";
}

void compute_turbulence_statistics(Field3D &F, const GridDesc &, const SolverParams &, double current_time)
{
    double o[13];
    check(ns3d_turbulence_statistics(F.ctx, o, 13), "turbulence_statistics", &F);
    TurbStats s{};
    s.kinetic_energy = o[0]; s.tke = o[1]; s.u_rms = o[2];
    s.taylor_Li = o[8]; s.Re_lambda = o[9]; s.Mach_t = o[10]; s.mean_mu = o[11]; s.mean_rho = o[12];
    std::vector<double> k, Ek;
    read_spectrum(time_prefix(current_time) + "_spectrum.dat", k, Ek);
    dissipation_scales(s, k, Ek);
    std::cout << "[Post] mean_mu = " << s.mean_mu << ", mean_rho = " << s.mean_rho << ", u_rms = " << s.u_rms
              << ", Taylor_Li = " << s.taylor_Li << ", Mach_t = " << s.Mach_t << "\n";
    append_turbulence_stats("output/turbulence_stats.dat", current_time, s);
}

void isotropic_post_process(Field3D &F, const GridDesc &G, const SolverParams &P, double current_time)
{
    compute_energy_spectrum(F, G, time_prefix(current_time) + "_spectrum.dat");
    compute_turbulence_statistics(F, G, P, current_time);
}

void write_tecplot_field(Field3D &F, const GridDesc &G, const SolverParams &, double time)
{
    F.sync_to_host();
    const std::string path = tecplot_field_path(time, 0);
    if (write_tecplot(path, time, 0, zone_of(F, G), F.rho.data(), F.u.data(), F.v.data(), F.w.data(), F.E.data(),
                      F.p.data(), F.T.data())) {
        char t[64];
        std::snprintf(t, sizeof(t), "%.8e", time);
        std::cerr << "Wrote Tecplot file: " << path << " (rank 0, time=" << t << ")\n";
    }
}

bool initialize_from_tecplot(Field3D &F, const GridDesc &G, const SolverParams &, const std::string &filename)
{
    return read_tecplot(filename, zone_of(F, G), F.rho.data(), F.u.data(), F.v.data(), F.w.data(), F.p.data());
}

bool initialize_from_tecplot_downsample(Field3D &F, const GridDesc &G, const SolverParams &,
                                        const std::string &filename, int src_nx, int src_ny, int src_nz)
{
    return read_tecplot_downsample(filename, zone_of(F, G), src_nx, src_ny, src_nz, F.rho.data(), F.u.data(),
                                   F.v.data(), F.w.data(), F.p.data());
}

bool write_restart_file(Field3D &F, const GridDesc &G, const SolverParams &P, const std::string &path)
{
    const int kind = P.precision == NS3D_PREC_MIXED ? 1 : 0;
    if (kind == 1) {
        check(ns3d_download_conserved(F.ctx, F.rho.data(), F.rhou.data(), F.rhov.data(), F.rhow.data(), F.E.data()),
              "download_conserved", &F);
    } else {
        check(ns3d_download_primitive(F.ctx, F.rho.data(), F.u.data(), F.v.data(), F.w.data(), F.p.data(),
                                      F.T.data()),
              "download_primitive", &F);
    }
    const double *a[5] = {F.rho.data(), kind ? F.rhou.data() : F.u.data(), kind ? F.rhov.data() : F.v.data(),
                          kind ? F.rhow.data() : F.w.data(), kind ? F.E.data() : F.p.data()};
    return write_restart(path, RestartInfo{F.nx, F.ny, F.nz, kind, F.step, F.time, F.next_output}, zone_of(F, G), a);
}

bool read_restart_file(Field3D &F, const GridDesc &G, const std::string &path)
{
    RestartInfo info{};
    std::vector<double> b[5];
    for (auto &x : b) x.assign(F.rho.size(), 0.0);
    double *a[5] = {b[0].data(), b[1].data(), b[2].data(), b[3].data(), b[4].data()};
    if (!read_restart(path, info, zone_of(F, G), a)) return false;
    if (info.kind == 1)
        check(ns3d_upload_conserved(F.ctx, a[0], a[1], a[2], a[3], a[4]), "upload_conserved", &F);
    else
        check(ns3d_upload_primitive(F.ctx, a[0], a[1], a[2], a[3], a[4]), "upload_primitive", &F);
    F.step = info.step;
    F.time = info.time;
    F.next_output = info.next_output;
    return true;
}

void time_advance(Field3D &F, GridDesc &G, SolverParams &P)
{
    using clk = std::chrono::steady_clock;
    const auto t_start = clk::now();
    auto t_last = t_start;
    double current_time = F.time;
    double output_Timefreq = P.output_timefreq;
    const double TotalTime = P.total_time;
    if (std::isfinite(F.next_output)) {
        // a restarted run takes the schedule where the run that wrote the
        // file left it (an output due at the restart time is still owed)
        output_Timefreq = F.next_output;
    } else {
        // a field without the schedule: resume it by the same repeated
        // additions an unbroken run makes
        while (F.time > 0.0 && output_Timefreq <= current_time + 1e-12 && P.output_timefreq > 0.0)
            output_Timefreq += P.output_timefreq;
    }
    F.next_output = output_Timefreq;
    int step = (int)F.step + 1;
    for (; step <= P.max_steps; ++step) {
        double dt = P.dt_fixed;
        const double eps = 1e-12;
        if (std::abs(TotalTime - current_time) <= eps) break;
        if (P.dt_fixed < 0.0) dt = compute_timestep(F, G, P);
        bool if_output = false;
        if (current_time >= output_Timefreq - eps) {
            if_output = true;
            output_Timefreq += P.output_timefreq;
        }
        if (current_time + dt > TotalTime) dt = TotalTime - current_time;
        if (current_time + dt > output_Timefreq) {
            dt = output_Timefreq - current_time;
            output_Timefreq += P.output_timefreq;
            if_output = true;
        }
        if (dt <= eps) {
            std::cerr << "dt too small or zero; aborting to avoid stall at Time=" << current_time << "\n";
            break;
        }
        current_time += dt;
        // U0 := U is taken inside the step (recordConservedTo0, maincycle.cpp:21)
        runge_kutta_3(F, G, P, dt);
        std::cout << "Completed step " << step << " / " << P.max_steps << ", Time = " << current_time << " / "
                  << TotalTime << "\n";
        if (step % P.monitor_stepfreq == 0 || step == 1) {
            if (P.monitor_res) compute_diagnostics(F, P, G);
            if (P.monitor_energy) compute_total_energy(F, G, P);
            const auto t_now = clk::now();
            const double t_elapsed = std::chrono::duration<double>(t_now - t_start).count();
            const double t_step = std::chrono::duration<double>(t_now - t_last).count();
            t_last = t_now;
            std::cout << std::fixed << std::setprecision(6) << "[Step " << step << "] "
                      << "dt=" << dt << "  Time/step=" << t_step << "s  Elapsed=" << t_elapsed << "s"
                      << "  Time=" << current_time << "/" << TotalTime;
            if (P.monitor_energy) std::cout << "  Ek_avg=" << F.global_Etot;
            std::cout << "\n";
            if (P.monitor_res) write_residuals_tecplot(F, step, "output_residuals.dat");
            if (P.isotropic_analyse) isotropic_post_process(F, G, P, current_time);
        }
        if (if_output) {
            if (P.post_basicfield) write_tecplot_field(F, G, P, current_time);
            if (P.isotropic_analyse) isotropic_post_process(F, G, P, current_time);
        }
        F.step = step;
        F.time = current_time;
        F.next_output = output_Timefreq;
        if (current_time >= TotalTime) break;
    }
    const double total = std::chrono::duration<double>(clk::now() - t_start).count();
    std::cout << "Simulation complete. Total time: " << total << " s\n";
}

}  // namespace ns3d_host
