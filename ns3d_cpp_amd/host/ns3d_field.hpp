// ns3d_field.hpp -- C++ host layer over the C ABI that keeps the reference's
// call surface (ns3d_func.h) for one periodic subdomain resident on one GPU.
//
// The reference's free functions take a host Field3D& and mutate it in place;
// here Field3D owns a device context and the same functions act on the
// device-resident state.  Host mirrors of the arrays are filled on demand
// (sync_to_host) for monitors and output, as the reference's monitors read
// its host arrays (maincycle.cpp:62-102).
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "ns3d_gpu.h"

namespace ns3d_host {

// GridDesc (field_structures.h:32-37) and SolverParams (:85-166) are the C
// structs of the ABI.
using GridDesc = ns3d_grid;
using SolverParams = ns3d_params;

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};

// Field3D (field_structures.h:176-472) with device-resident storage.
struct Field3D {
    ns3d_ctx *ctx = nullptr;
    int nx = 0, ny = 0, nz = 0, ng = 0, sx = 0, sy = 0, sz = 0;
    // host mirrors (idx3 layout), valid after sync_to_host()
    std::vector<double> rho, u, v, w, p, T;
    std::vector<double> rhou, rhov, rhow, E;
    double global_Etot = 0.0;
    // compute_diagnostics results: rho, rho u, rho v, rho w, E
    double global_res[5] = {0, 0, 0, 0, 0};
    // step index and time the state belongs to: time_advance starts after
    // them and leaves its last step and time here (0, 0 for a fresh field)
    long long step = 0;
    double time = 0.0;
    // the output time the schedule waits for after that step (maincycle.cpp:
    // 33-45); NaN when unknown (a fresh field, or a restart file without it)
    double next_output = std::nan("");

    Field3D() = default;
    Field3D(const Field3D &) = delete;
    Field3D &operator=(const Field3D &) = delete;
    ~Field3D();

    // Field3D::allocate (field_structures.h:225-334)
    void allocate(const SolverParams &P, const GridDesc &G);
    inline int I(int i, int j, int k) const { return (k * sy + j) * sx + i; }
    // upload rho,u,v,w,p (host mirrors) then apply_boundary + P2C (main.cpp:70-71)
    void upload_primitive();
    // download primitives and conserved into the host mirrors
    void sync_to_host();
};

void check(int st, const char *what, const Field3D *F = nullptr);

// ns3d_func.h call surface
void apply_boundary(Field3D &F, GridDesc &G, const SolverParams &P);     // :56
void du_inviscous(Field3D &F, const SolverParams &P);                      // :154
void du_inviscous_character(Field3D &F, const SolverParams &P);            // :155
void compute_vis_flux(Field3D &F, const GridDesc &G, const SolverParams &P);   // :163
// the per-point functions (:96, :136-145), on the host mirrors of F (valid
// after sync_to_host) through the batched device entry points
void split_StagerWarming(const Field3D &F, std::vector<double> &F_plus, std::vector<double> &F_minus, int id, int n1,
                         int n2, int n3, const SolverParams &P);
void split_LaxFriedrichs(const Field3D &F, std::vector<double> &F_plus, std::vector<double> &F_minus, int id, int n1,
                         int n2, int n3, const SolverParams &P);
// throws Error (a std::runtime_error) on a stencil of the wrong size, as the
// reference throws
double reconstruct_select(const std::vector<double> &vstencil, double flag, const SolverParams P);
void compute_gradients(Field3D &F, const GridDesc &G, const SolverParams &P);  // :157
void compute_viscous_flux(Field3D &F, const SolverParams &P);              // :160
void compute_rhs(Field3D &F, GridDesc &G, SolverParams &P);                // :99
void runge_kutta_3(Field3D &F, GridDesc &G, SolverParams &P, double dt);   // :102
double compute_timestep(Field3D &F, const GridDesc &G, const SolverParams &P);  // timestep.cpp:5
void compute_total_energy(Field3D &F, const GridDesc &G, const SolverParams &P); // postprocess.cpp:103
void compute_diagnostics(Field3D &F, const SolverParams &P, const GridDesc &G);   // postprocess.cpp:7
void write_residuals_tecplot(const Field3D &F, int step, const std::string &filename);   // output.cpp:96
// spectrum of the device state written as "k Ek" lines (EnergySpectrum.cpp:159)
void compute_energy_spectrum(Field3D &F, const GridDesc &G, const std::string &filename);
// statistics appended to output/turbulence_stats.dat; like the reference,
// the dissipation is taken from the spectrum file that
// isotropic_post_process wrote for this time (ana_isotropic.cpp:207-219)
void compute_turbulence_statistics(Field3D &F, const GridDesc &G, const SolverParams &P, double current_time);
void isotropic_post_process(Field3D &F, const GridDesc &G, const SolverParams &P, double current_time);  // postprocess.cpp:133
void write_tecplot_field(Field3D &F, const GridDesc &G, const SolverParams &P, double time);   // output.cpp:21
// initial field from a Tecplot file of the grid's size, or resampled from one
// of src sizes (initialization.cpp:35, :103); fills the host mirrors of rho,
// u, v, w, p -- upload_primitive() then applies the boundary and P2C
bool initialize_from_tecplot(Field3D &F, const GridDesc &G, const SolverParams &P, const std::string &filename);
bool initialize_from_tecplot_downsample(Field3D &F, const GridDesc &G, const SolverParams &P,
                                        const std::string &filename, int src_nx, int src_ny, int src_nz);
// Binary restart of the context's canonical state (primitives in FP64 mode,
// conserved variables in mixed mode) with F.step and F.time; reading uploads
// the state (boundary fill included) and sets F.step, F.time and
// F.next_output.
bool write_restart_file(Field3D &F, const GridDesc &G, const SolverParams &P, const std::string &path);
bool read_restart_file(Field3D &F, const GridDesc &G, const std::string &path);
// time_advance (maincycle.cpp:7-113): dt fixed or CFL, dt clipping to the
// output time and TotalTime, monitor lines on stdout, residual and
// turbulence histories, field files at output times.
void time_advance(Field3D &F, GridDesc &G, SolverParams &P);

}  // namespace ns3d_host
