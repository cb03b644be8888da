// ns3d_gpu_run.cpp -- the main.cpp counterpart (main.cpp:9-86) on one GPU:
// read solver.in, allocate the device field, set the initial field, run
// time_advance.
//
// usage: ns3d_gpu_run <solver.in> [--ic tgv|uniform|raw:<file>]
//   tgv       Taylor-Green vortex at x_i = x0 + i*dx (rho = 1,
//             p = 1/(gamma Ma^2) + (cos 2x + cos 2y)(cos 2z + 2)/16)
//   uniform   initialize_uniform_field (initialization.cpp:528-546)
//   raw:<f>   five interior arrays rho,u,v,w,p as little-endian doubles,
//             i fastest (the interior of the idx3 layout)
#include <cmath>
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
    std::string ic = "tgv";
    for (int a = 2; a + 1 < argc; ++a)
        if (std::string(argv[a]) == "--ic") ic = argv[a + 1];
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
        if (ic == "tgv") set_tgv(F, G, P);
        else if (ic == "uniform") set_uniform(F);
        else if (ic.rfind("raw:", 0) == 0) {
            if (!set_raw(F, ic.substr(4))) {
                std::cerr << "Error: cannot read " << ic.substr(4) << "\n";
                return 1;
            }
        } else {
            std::cerr << "unknown --ic " << ic << "\n";
            return 2;
        }
        F.upload_primitive();
        std::cout << "Initialization + halo exchange done\n";
        time_advance(F, G, P);
    } catch (const Error &e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
