// ns3d_gpu_run.cpp -- the main.cpp counterpart (main.cpp:9-86) on one GPU:
// read solver.in, allocate the device field, set the initial field, write
// it (and its spectrum and statistics when isotropic_analyse is on), run
// time_advance, write the final field.
//
// usage: ns3d_gpu_run <solver.in> [--ic <initial field>] [--no-field-files]
//   tgv       Taylor-Green vortex at x_i = x0 + i*dx (rho = 1,
//             p = 1/(gamma Ma^2) + (cos 2x + cos 2y)(cos 2z + 2)/16)
//   uniform   initialize_uniform_field (initialization.cpp:528-546)
//   raw:<f>   five interior arrays rho,u,v,w,p as little-endian doubles,
//             i fastest (the interior of the idx3 layout)
//   tecplot:<f>                  a field file of the grid's size
//                                (initialize_from_tecplot)
//   tecplot_ds:<f>:<nx>:<ny>:<nz> a field file of nx*ny*nz points resampled
//                                onto the grid (initialize_from_tecplot_downsample)
// --no-field-files skips the field files at time 0 and at the end
// (main.cpp:76, :83); the files at output times follow post_basicfield.
// --restart-in <f> starts from a binary restart file (its step and time
// included) instead of --ic; --restart-out <f> writes one after the last step.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "ns3d_field.hpp"

using namespace ns3d_host;

static void set_tgv(Field3D &F, const GridDesc &G, const SolverParams &P)
{
    const int ng = F.ng;
    for (int k = 0; k < F.nz; ++k)
        for (int j = 0; j < F.ny; ++j)
            for (int i = 0; i < F.nx; ++i) {
                const double x = G.x0 + i * G.dx, y = G.y0 + j * G.dy, z = G.z0 + k * G.dz;
                const int id = F.I(i + ng, j + ng, k + ng);
                F.rho[id] = 1.0;
                F.u[id] = std::sin(x) * std::cos(y) * std::cos(z);
                F.v[id] = -std::cos(x) * std::sin(y) * std::cos(z);
                F.w[id] = 0.0;
                F.p[id] = 1.0 / (P.gamma * P.Ma * P.Ma) +
                          (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0) / 16.0;
            }
}

static void set_uniform(Field3D &F)
{
    const int ng = F.ng;
    for (int k = 0; k < F.nz; ++k)
        for (int j = 0; j < F.ny; ++j)
            for (int i = 0; i < F.nx; ++i) {
                const int id = F.I(i + ng, j + ng, k + ng);
                F.rho[id] = 1.0; F.u[id] = 1.0; F.v[id] = 1.0; F.w[id] = 1.0; F.p[id] = 1.0;
            }
}

static bool set_raw(Field3D &F, const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    const int ng = F.ng;
    std::vector<double> *arr[5] = {&F.rho, &F.u, &F.v, &F.w, &F.p};
    std::vector<double> line((size_t)F.nx);
    for (auto *a : arr)
        for (int k = 0; k < F.nz; ++k)
            for (int j = 0; j < F.ny; ++j) {
                if (!in.read(reinterpret_cast<char *>(line.data()), (std::streamsize)(line.size() * sizeof(double))))
                    return false;
                std::memcpy(&(*a)[(size_t)F.I(ng, j + ng, k + ng)], line.data(), line.size() * sizeof(double));
            }
    return true;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <solver.in> [--ic tgv|uniform|raw:<file>]\n";
        return 2;
    }
    std::string ic = "tgv", restart_in, restart_out;
    bool field_files = true;
    for (int a = 2; a < argc; ++a) {
        const std::string arg = argv[a];
        if (arg == "--ic" && a + 1 < argc) ic = argv[++a];
        else if (arg == "--restart-in" && a + 1 < argc) restart_in = argv[++a];
        else if (arg == "--restart-out" && a + 1 < argc) restart_out = argv[++a];
        else if (arg == "--no-field-files") field_files = false;
        else {
            std::cerr << "unknown argument " << arg << "\n";
            return 2;
        }
    }
    SolverParams P;
    GridDesc G;
    int st = ns3d_params_from_file(argv[1], &P, &G);
    if (st) {
        std::cerr << "Error: cannot read solver parameter file " << argv[1] << ": " << ns3d_status_string(st) << "\n";
        return 1;
    }
    std::cout << "Global grid: " << G.nx << " x " << G.ny << " x " << G.nz << "\n"
              << "  dx, dy, dz: [" << G.dx << ", " << G.dy << ", " << G.dz << "]  ghost layers " << G.ng << "\n"
              << "  gamma: " << P.gamma << ", Pr: " << P.Pr << ", Ma: " << P.Ma << ", Re: " << P.Re << "\n";
    try {
        Field3D F;
        F.allocate(P, G);
        if (!restart_in.empty()) {
            if (!read_restart_file(F, G, restart_in)) return 1;
            std::cout << "Restart from " << restart_in << " at step " << F.step << ", time " << F.time << "\n";
        } else if (ic == "tgv") set_tgv(F, G, P);
        else if (ic == "uniform") set_uniform(F);
        else if (ic.rfind("raw:", 0) == 0) {
            if (!set_raw(F, ic.substr(4))) {
                std::cerr << "Error: cannot read " << ic.substr(4) << "\n";
                return 1;
            }
        } else if (ic.rfind("tecplot:", 0) == 0) {
            if (!initialize_from_tecplot(F, G, P, ic.substr(8))) return 1;
        } else if (ic.rfind("tecplot_ds:", 0) == 0) {
            // tecplot_ds:<file>:<nx>:<ny>:<nz>, the sizes after the last three colons
            std::string rest = ic.substr(11);
            int n[3];
            for (int q = 2; q >= 0; --q) {
                const size_t c = rest.rfind(':');
                if (c == std::string::npos) {
                    std::cerr << "--ic tecplot_ds:<file>:<nx>:<ny>:<nz>\n";
                    return 2;
                }
                n[q] = std::atoi(rest.c_str() + c + 1);
                rest.resize(c);
            }
            if (!initialize_from_tecplot_downsample(F, G, P, rest, n[0], n[1], n[2])) return 1;
        } else {
            std::cerr << "unknown --ic " << ic << "\n";
            return 2;
        }
        if (restart_in.empty()) F.upload_primitive();
        std::cout << "Initialization + halo exchange done\n";
        if (field_files) write_tecplot_field(F, G, P, F.time);
        if (P.isotropic_analyse) isotropic_post_process(F, G, P, F.time);
        time_advance(F, G, P);
        if (field_files) write_tecplot_field(F, G, P, P.total_time);
        if (!restart_out.empty() && !write_restart_file(F, G, P, restart_out)) return 1;
    } catch (const Error &e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
