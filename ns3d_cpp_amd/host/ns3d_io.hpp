// ns3d_io.hpp -- host file formats of the solver: the Tecplot ASCII field
// (output.cpp:21-87, read back by initialization.cpp:35-100 and resampled by
// :103-190), the shell spectrum file (EnergySpectrum.cpp:127-133), the
// turbulence-statistics history (ana_isotropic.cpp:317-342) and the residual
// history (output.cpp:93-128).
//
// Plain host code over host arrays in the padded idx3 layout
// (cell (i,j,k) of the interior at ((k+ng)*sy + j+ng)*sx + i+ng); no device
// calls, so the formats are testable without a GPU.
#pragma once

#include <string>
#include <vector>

namespace ns3d_host {

// Geometry of the written zone: interior sizes, padding, cell coordinates
// x = x0 + i*dx (output.cpp:62-64).
struct ZoneDesc {
    int nx, ny, nz, ng, sx, sy;
    double x0, y0, z0, dx, dy, dz;
};

// The ten point-packed variables X Y Z rho u v w E p T, i fastest.
// Returns false (and prints why) when the file cannot be written.
bool write_tecplot(const std::string &path, double time, int rank, const ZoneDesc &Z, const double *rho,
                   const double *u, const double *v, const double *w, const double *E, const double *p,
                   const double *T);

// output/time_<int(time*1e5)>/field_rank<rank>.dat, the directory created
std::string tecplot_field_path(double time, int rank);

// Reads a single-zone file of exactly nx*ny*nz points (x fastest) into the
// interior of padded arrays; only rho, u, v, w, p are kept (the caller
// recomputes the conserved state from them).  Header lines up to the first
// line that starts with "ZONE" are skipped.
bool read_tecplot(const std::string &path, const ZoneDesc &Z, double *rho, double *u, double *v, double *w,
                  double *p);

// Nearest-sample resampling of a src_nx*src_ny*src_nz file onto the zone:
// target cell g takes source index clamp(llround(g*srcN/tgtN), 0, srcN-1) per
// axis (initialization.cpp:127-133).
bool read_tecplot_downsample(const std::string &path, const ZoneDesc &Z, int src_nx, int src_ny, int src_nz,
                             double *rho, double *u, double *v, double *w, double *p);

// "output/output_time_<time with 5 decimals>" (postprocess.cpp:136-137)
std::string time_prefix(double time);

// "k Ek" per line, k = 0..Ek.size()-1, Ek at the stream's default six
// significant digits (EnergySpectrum.cpp:128-133)
bool write_spectrum(const std::string &path, const std::vector<double> &Ek);
// the pairs of such a file, in order (ana_isotropic.cpp:215-219)
bool read_spectrum(const std::string &path, std::vector<double> &k, std::vector<double> &Ek);

struct TurbStats {
    double kinetic_energy, tke, u_rms, dissipation, taylor, eta, u_eta, tau_eta, taylor_Li, Re_lambda, Mach_t;
    double mean_mu, mean_rho;
};

// Dissipation and the scales derived from it, eps = sum_i 2 nu k_i^2 E_i with
// nu = mean_mu / mean_rho, in the order of ana_isotropic.cpp:300-313.
void dissipation_scales(TurbStats &s, const std::vector<double> &k, const std::vector<double> &Ek);

// appends one line (header first if the file is new) (ana_isotropic.cpp:317-342)
bool append_turbulence_stats(const std::string &path, double time, const TurbStats &s);

// step 1 truncates and writes the header, later steps append (output.cpp:93-128)
bool write_residual_line(const std::string &path, int step, const double res[5], double Etot);

// Binary restart file (no counterpart in the reference, whose only restart
// path is the 8-digit Tecplot reader).  Little-endian: the 8-byte magic
// "NS3DRST2", int32 nx, ny, nz, kind, int64 step, double time, double
// next_output (the output time the schedule of maincycle.cpp:33-45 waits
// for), then five interior arrays (x fastest) of doubles: kind 0 = rho, u,
// v, w, p (the canonical state of the FP64 mode), kind 1 = rho, rho u,
// rho v, rho w, E (that of the mixed mode), so a run continues bit for bit,
// output schedule included.  Files of the earlier "NS3DRST1" layout (no
// next_output) are read with next_output = NaN (unknown).
struct RestartInfo {
    int nx, ny, nz, kind;
    long long step;
    double time;
    double next_output;
};
bool write_restart(const std::string &path, const RestartInfo &info, const ZoneDesc &Z, const double *const a[5]);
// reads into the interior of padded arrays; false (with a message) on a
// missing file, a wrong magic or sizes that differ from the zone
bool read_restart(const std::string &path, RestartInfo &info, const ZoneDesc &Z, double *const a[5]);

}  // namespace ns3d_host
