// Command-line harness over ns3d_cpp_amd/host/ns3d_io.cpp for the file-format
// tests (test_host_io.py); built with g++ on the host, no device code.
//
//   formats                          snprintf vs iostream output of test values
//   write <out> nx ny nz ng time     pattern field -> write_tecplot
//   read <in> nx ny nz ng <raw>      read_tecplot -> padded rho,u,v,w,p raw
//   downsample <in> snx sny snz nx ny nz ng <raw>
//   spectrum <out> e0 e1 ...         write_spectrum, then read_spectrum echoed
//   stats <spec> <out> time ke tke u_rms taylor_Li Re_lambda Mach_t mu rho
//   resid <out> step r0 r1 r2 r3 r4 Etot
//   paths time                       tecplot_field_path and time_prefix
//   rwrite <out> nx ny nz ng kind step time next_output   pattern field -> write_restart
//   rread <in> nx ny nz ng <raw>     read_restart -> header on stdout, padded arrays raw
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "ns3d_io.hpp"

using namespace ns3d_host;

static double arg(char **argv, int i) { return std::strtod(argv[i], nullptr); }

static double pattern(int q, int i, int j, int k)
{
    // distinct, sign-varying values per variable; exact in binary
    return (q + 1) * 0.5 + i * 0.125 - j * 0.0625 + k * 2.0 + (q == 1 ? -3.75 : 0.0);
}

static int formats()
{
    const double vals[] = {0.0, -0.0, 1.0, -1.0, 0.1, 1.0 / 3.0, 2.0 / 3.0, 123456789.123, 1e-300, 5e-324,
                           1e300, -7.77777777777e-12, 0.999999999999, 9.9999999995e-5, 6.283185307179586,
                           std::numeric_limits<double>::infinity()};
    int bad = 0;
    for (double x : vals) {
        std::ostringstream a, b;
        a << std::scientific << std::setprecision(8) << x;
        b << x;
        char sa[64], sb[64];
        std::snprintf(sa, sizeof(sa), "%.8e", x);
        std::snprintf(sb, sizeof(sb), "%.6g", x);
        if (a.str() != sa || b.str() != sb) {
            std::printf("mismatch %s/%s %s/%s\n", a.str().c_str(), sa, b.str().c_str(), sb);
            ++bad;
        }
    }
    std::printf("formats %s\n", bad ? "differ" : "equal");
    return bad ? 1 : 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if (cmd == "formats") return formats();
    if (cmd == "paths" && argc == 3) {
        std::printf("%s\n%s\n", tecplot_field_path(arg(argv, 2), 0).c_str(), time_prefix(arg(argv, 2)).c_str());
        return 0;
    }
    if ((cmd == "write" && argc == 8) || (cmd == "read" && argc == 8) || (cmd == "downsample" && argc == 11)) {
        const int o = cmd == "downsample" ? 3 : 0;
        ZoneDesc Z{};
        Z.nx = std::atoi(argv[3 + o]); Z.ny = std::atoi(argv[4 + o]); Z.nz = std::atoi(argv[5 + o]);
        Z.ng = std::atoi(argv[6 + o]);
        Z.sx = Z.nx + 2 * Z.ng; Z.sy = Z.ny + 2 * Z.ng;
        const int sz = Z.nz + 2 * Z.ng;
        Z.x0 = -1.0; Z.y0 = 0.5; Z.z0 = 2.0; Z.dx = 0.25; Z.dy = 0.125; Z.dz = 1.0 / 3.0;
        const size_t tot = (size_t)Z.sx * Z.sy * sz;
        std::vector<std::vector<double>> a(7, std::vector<double>(tot, std::nan("")));
        if (cmd == "write") {
            for (int k = 0; k < Z.nz; ++k)
                for (int j = 0; j < Z.ny; ++j)
                    for (int i = 0; i < Z.nx; ++i) {
                        const size_t id = ((size_t)(k + Z.ng) * Z.sy + (j + Z.ng)) * Z.sx + (i + Z.ng);
                        for (int q = 0; q < 7; ++q) a[q][id] = pattern(q, i, j, k);
                    }
            return write_tecplot(argv[2], arg(argv, 7), 0, Z, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                                 a[4].data(), a[5].data(), a[6].data())
                       ? 0
                       : 1;
        }
        bool ok;
        if (cmd == "read")
            ok = read_tecplot(argv[2], Z, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data());
        else
            ok = read_tecplot_downsample(argv[2], Z, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                                         a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data());
        if (!ok) return 1;
        FILE *f = std::fopen(argv[7 + o], "wb");
        if (!f) return 1;
        for (int q = 0; q < 5; ++q) std::fwrite(a[q].data(), sizeof(double), tot, f);
        std::fclose(f);
        return 0;
    }
    if ((cmd == "rwrite" && argc == 11) || (cmd == "rread" && argc == 8)) {
        ZoneDesc Z{};
        Z.nx = std::atoi(argv[3]); Z.ny = std::atoi(argv[4]); Z.nz = std::atoi(argv[5]); Z.ng = std::atoi(argv[6]);
        Z.sx = Z.nx + 2 * Z.ng; Z.sy = Z.ny + 2 * Z.ng;
        const size_t tot = (size_t)Z.sx * Z.sy * (Z.nz + 2 * Z.ng);
        std::vector<std::vector<double>> a(5, std::vector<double>(tot, std::nan("")));
        double *p[5];
        for (int q = 0; q < 5; ++q) p[q] = a[q].data();
        if (cmd == "rwrite") {
            for (int k = 0; k < Z.nz; ++k)
                for (int j = 0; j < Z.ny; ++j)
                    for (int i = 0; i < Z.nx; ++i) {
                        const size_t id = ((size_t)(k + Z.ng) * Z.sy + (j + Z.ng)) * Z.sx + (i + Z.ng);
                        for (int q = 0; q < 5; ++q) a[q][id] = pattern(q, i, j, k) + 1e-13 * q;
                    }
            const RestartInfo info{Z.nx, Z.ny, Z.nz, std::atoi(argv[7]), std::atoll(argv[8]), arg(argv, 9),
                                   arg(argv, 10)};
            const double *c[5] = {p[0], p[1], p[2], p[3], p[4]};
            return write_restart(argv[2], info, Z, c) ? 0 : 1;
        }
        RestartInfo info{};
        if (!read_restart(argv[2], info, Z, p)) return 1;
        std::printf("%d %d %d %d %lld %.17g %.17g\n", info.nx, info.ny, info.nz, info.kind, info.step, info.time,
                    info.next_output);
        FILE *f = std::fopen(argv[7], "wb");
        if (!f) return 1;
        for (int q = 0; q < 5; ++q) std::fwrite(a[q].data(), sizeof(double), tot, f);
        std::fclose(f);
        return 0;
    }
    if (cmd == "spectrum" && argc >= 3) {
        std::vector<double> Ek;
        for (int i = 3; i < argc; ++i) Ek.push_back(arg(argv, i));
        if (!write_spectrum(argv[2], Ek)) return 1;
        std::vector<double> k, e;
        if (!read_spectrum(argv[2], k, e)) return 1;
        for (size_t i = 0; i < k.size(); ++i) std::printf("%.17g %.17g\n", k[i], e[i]);
        return 0;
    }
    if (cmd == "stats" && argc == 13) {
        TurbStats s{};
        s.kinetic_energy = arg(argv, 5); s.tke = arg(argv, 6); s.u_rms = arg(argv, 7);
        s.taylor_Li = arg(argv, 8); s.Re_lambda = arg(argv, 9); s.Mach_t = arg(argv, 10);
        s.mean_mu = arg(argv, 11); s.mean_rho = arg(argv, 12);
        std::vector<double> k, e;
        read_spectrum(argv[2], k, e);
        dissipation_scales(s, k, e);
        return append_turbulence_stats(argv[3], arg(argv, 4), s) ? 0 : 1;
    }
    if (cmd == "resid" && argc == 10) {
        const double r[5] = {arg(argv, 4), arg(argv, 5), arg(argv, 6), arg(argv, 7), arg(argv, 8)};
        return write_residual_line(argv[2], std::atoi(argv[3]), r, arg(argv, 9)) ? 0 : 1;
    }
    return 2;
}
