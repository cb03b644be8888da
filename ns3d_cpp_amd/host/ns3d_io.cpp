// ns3d_io.cpp -- host file formats (see header).  Numbers are printed with
// snprintf conversions equal to the iostream settings of the reference's
// writers: std::scientific + setprecision(8) is "%.8e", the default stream
// format is "%.6g".
#include "ns3d_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cctype>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <memory>

namespace ns3d_host {

namespace {

struct FileCloser {
    void operator()(FILE *f) const { if (f) std::fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

void ensure_parent(const std::string &path)
{
    const std::filesystem::path dir = std::filesystem::path(path).parent_path();
    if (dir.empty()) return;
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) std::cerr << "Warning: could not create directory '" << dir.string() << "': " << ec.message() << "\n";
}

// Whitespace-separated numbers from a file, read through a fixed buffer.
class NumberReader {
  public:
    explicit NumberReader(FILE *f) : f_(f), buf_(1 << 20) {}

    // Skips whole lines until one that starts with `tag` has been consumed.
    bool skip_through_line(const char *tag)
    {
        std::string line;
        for (;;) {
            line.clear();
            int ch;
            while ((ch = get()) != EOF && ch != '\n') line.push_back((char)ch);
            if (line.rfind(tag, 0) == 0) return true;
            if (ch == EOF) return false;
        }
    }

    bool next(double &x)
    {
        int ch;
        do { ch = get(); } while (ch != EOF && std::isspace(ch));
        if (ch == EOF) return false;
        char tok[64];
        size_t n = 0;
        while (ch != EOF && !std::isspace(ch)) {
            if (n + 1 < sizeof(tok)) tok[n++] = (char)ch;
            ch = get();
        }
        tok[n] = '\0';
        char *end = nullptr;
        x = std::strtod(tok, &end);
        return end != tok && *end == '\0';
    }

  private:
    int get()
    {
        if (pos_ == len_) {
            len_ = std::fread(buf_.data(), 1, buf_.size(), f_);
            pos_ = 0;
            if (len_ == 0) return EOF;
        }
        return (unsigned char)buf_[pos_++];
    }
    FILE *f_;
    std::vector<char> buf_;
    size_t pos_ = 0, len_ = 0;
};

inline size_t cell(const ZoneDesc &Z, int i, int j, int k)
{
    return ((size_t)(k + Z.ng) * Z.sy + (size_t)(j + Z.ng)) * Z.sx + (size_t)(i + Z.ng);
}

// one point of a file: X Y Z rho u v w E p T
bool read_point(NumberReader &rd, double v[10])
{
    for (int q = 0; q < 10; ++q)
        if (!rd.next(v[q])) return false;
    return true;
}

}  // namespace

std::string tecplot_field_path(double time, int rank)
{
    // output.cpp:28 names the directory by the truncated time * 1e5
    return "output/time_" + std::to_string(static_cast<int>(time * 100000)) + "/field_rank" + std::to_string(rank) +
           ".dat";
}

bool write_tecplot(const std::string &path, double time, int rank, const ZoneDesc &Z, const double *rho,
                   const double *u, const double *v, const double *w, const double *E, const double *p,
                   const double *T)
{
    ensure_parent(path);
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) {
        std::cerr << "Failed to open output file " << path << " for writing\n";
        return false;
    }
    std::fprintf(f.get(), "TITLE = \"NS3D Field Rank %d  Time=%.8e\"\n", rank, time);
    std::fprintf(f.get(), "VARIABLES = \"X\" \"Y\" \"Z\" \"rho\" \"u\" \"v\" \"w\" \"E\" \"p\" \"T\"\n");
    std::fprintf(f.get(), "ZONE T=\"rank_%d time=%.8e\" I=%d J=%d K=%d DATAPACKING=POINT\n", rank, time, Z.nx, Z.ny,
                 Z.nz);
    std::vector<char> out;
    out.reserve((size_t)Z.nx * 10 * 17 + 64);
    char line[400];
    for (int k = 0; k < Z.nz; ++k)
        for (int j = 0; j < Z.ny; ++j) {
            out.clear();
            for (int i = 0; i < Z.nx; ++i) {
                const size_t id = cell(Z, i, j, k);
                const int n = std::snprintf(line, sizeof(line), "%.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e\n",
                                            Z.x0 + i * Z.dx, Z.y0 + j * Z.dy, Z.z0 + k * Z.dz, rho[id], u[id], v[id],
                                            w[id], E[id], p[id], T[id]);
                out.insert(out.end(), line, line + n);
            }
            if (std::fwrite(out.data(), 1, out.size(), f.get()) != out.size()) {
                std::cerr << "Failed writing " << path << "\n";
                return false;
            }
        }
    return true;
}

bool read_tecplot(const std::string &path, const ZoneDesc &Z, double *rho, double *u, double *v, double *w,
                  double *p)
{
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) {
        std::cerr << "Error: cannot open Tecplot file: " << path << "\n";
        return false;
    }
    NumberReader rd(f.get());
    rd.skip_through_line("ZONE");
    double val[10];
    for (int k = 0; k < Z.nz; ++k)
        for (int j = 0; j < Z.ny; ++j)
            for (int i = 0; i < Z.nx; ++i) {
                if (!read_point(rd, val)) {
                    std::cerr << "Error: Tecplot data lines are fewer than expected when reading " << path << "\n";
                    return false;
                }
                const size_t id = cell(Z, i, j, k);
                rho[id] = val[3]; u[id] = val[4]; v[id] = val[5]; w[id] = val[6]; p[id] = val[8];
            }
    return true;
}

bool read_tecplot_downsample(const std::string &path, const ZoneDesc &Z, int src_nx, int src_ny, int src_nz,
                             double *rho, double *u, double *v, double *w, double *p)
{
    if (src_nx < 1 || src_ny < 1 || src_nz < 1) return false;
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) {
        std::cerr << "Error: cannot open Tecplot file: " << path << "\n";
        return false;
    }
    // per axis, the target indices that sample each source index
    auto targets = [](int tgtN, int srcN) {
        std::vector<std::vector<int>> t((size_t)srcN);
        for (int g = 0; g < tgtN; ++g) {
            int s = 0;
            if (tgtN > 1) {
                const double pos = static_cast<double>(g) * static_cast<double>(srcN) / static_cast<double>(tgtN);
                s = std::clamp(static_cast<int>(std::llround(pos)), 0, srcN - 1);
            }
            t[(size_t)s].push_back(g);
        }
        return t;
    };
    const auto tx = targets(Z.nx, src_nx), ty = targets(Z.ny, src_ny), tz = targets(Z.nz, src_nz);
    NumberReader rd(f.get());
    rd.skip_through_line("ZONE");
    double val[10];
    size_t filled = 0;
    for (int sk = 0; sk < src_nz; ++sk)
        for (int sj = 0; sj < src_ny; ++sj)
            for (int si = 0; si < src_nx; ++si) {
                if (!read_point(rd, val)) {
                    std::cerr << "Error: Tecplot data lines are fewer than expected when reading " << path << "\n";
                    return false;
                }
                for (int k : tz[(size_t)sk])
                    for (int j : ty[(size_t)sj])
                        for (int i : tx[(size_t)si]) {
                            const size_t id = cell(Z, i, j, k);
                            rho[id] = val[3]; u[id] = val[4]; v[id] = val[5]; w[id] = val[6]; p[id] = val[8];
                            ++filled;
                        }
            }
    const size_t want = (size_t)Z.nx * Z.ny * Z.nz;
    if (filled < want)
        std::cerr << "Warning: only filled " << filled << " / " << want << " cells while sampling Tecplot file "
                  << path << "\n";
    return true;
}

std::string time_prefix(double time)
{
    char buf[128];
    std::snprintf(buf, sizeof(buf), "output/output_time_%.5f", time);
    return buf;
}

bool write_spectrum(const std::string &path, const std::vector<double> &Ek)
{
    ensure_parent(path);
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) {
        std::cerr << "Failed to open spectrum file " << path << "\n";
        return false;
    }
    for (size_t k = 0; k < Ek.size(); ++k) std::fprintf(f.get(), "%d %.6g\n", (int)k, Ek[k]);
    return true;
}

bool read_spectrum(const std::string &path, std::vector<double> &k, std::vector<double> &Ek)
{
    k.clear();
    Ek.clear();
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) return false;
    NumberReader rd(f.get());
    double a, b;
    while (rd.next(a) && rd.next(b)) {
        k.push_back(a);
        Ek.push_back(b);
    }
    return true;
}

void dissipation_scales(TurbStats &s, const std::vector<double> &k, const std::vector<double> &Ek)
{
    const double nu = s.mean_mu / s.mean_rho;
    double eps = 0;
    for (size_t i = 0; i < k.size() && i < Ek.size(); i++) eps += 2.0 * nu * k[i] * k[i] * Ek[i];
    s.dissipation = eps;
    s.eta = std::pow(nu * nu * nu / eps, 0.25);
    s.u_eta = std::pow(nu * eps, 0.25);
    s.tau_eta = std::pow(nu / eps, 0.5);
    s.taylor = std::sqrt(15.0 * nu * s.u_rms * s.u_rms / eps);
}

bool append_turbulence_stats(const std::string &path, double time, const TurbStats &s)
{
    ensure_parent(path);
    const bool exists = std::filesystem::exists(path);
    File f(std::fopen(path.c_str(), "ab"));
    if (!f) {
        std::cout << "[Error] Cannot open " << path << " for writing.\n";
        return false;
    }
    if (!exists)
        std::fprintf(f.get(), "current_time Kinetic Energy TKE u_rms Dissipation Taylor Kol_scale Kol_vel Kol_time "
                              "Taylor_Li Re_lambda Mach_t\n");
    std::fprintf(f.get(), "%.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e %.8e\n", time, s.kinetic_energy,
                 s.tke, s.u_rms, s.dissipation, s.taylor, s.eta, s.u_eta, s.tau_eta, s.taylor_Li, s.Re_lambda,
                 s.Mach_t);
    return true;
}

bool write_residual_line(const std::string &path, int step, const double res[5], double Etot)
{
    ensure_parent(path);
    File f(std::fopen(path.c_str(), step == 1 ? "wb" : "ab"));
    if (!f) {
        std::cerr << "Failed to open residuals file " << path << "\n";
        return false;
    }
    if (step == 1) {
        std::fprintf(f.get(), "TITLE = \"Residuals vs Step\"\n");
        std::fprintf(f.get(), "VARIABLES = \"Step\" \"Res_rho\" \"Res_rhou\" \"Res_rhov\" \"Res_rhow\" \"Res_E\" \"Etot\"\n");
    }
    std::fprintf(f.get(), "%d %.8e %.8e %.8e %.8e %.8e %.8e\n", step, res[0], res[1], res[2], res[3], res[4], Etot);
    return true;
}

namespace {
const char kRestartMagic[8] = {'N', 'S', '3', 'D', 'R', 'S', 'T', '2'};
const char kRestartMagicV1[8] = {'N', 'S', '3', 'D', 'R', 'S', 'T', '1'};
}

bool write_restart(const std::string &path, const RestartInfo &info, const ZoneDesc &Z, const double *const a[5])
{
    ensure_parent(path);
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) {
        std::cerr << "Failed to open restart file " << path << " for writing\n";
        return false;
    }
    const int32_t dims[4] = {Z.nx, Z.ny, Z.nz, info.kind};
    const int64_t step = info.step;
    bool ok = std::fwrite(kRestartMagic, 1, 8, f.get()) == 8 && std::fwrite(dims, sizeof(int32_t), 4, f.get()) == 4 &&
              std::fwrite(&step, sizeof(step), 1, f.get()) == 1 &&
              std::fwrite(&info.time, sizeof(double), 1, f.get()) == 1 &&
              std::fwrite(&info.next_output, sizeof(double), 1, f.get()) == 1;
    for (int q = 0; q < 5 && ok; ++q)
        for (int k = 0; k < Z.nz && ok; ++k)
            for (int j = 0; j < Z.ny && ok; ++j)
                ok = std::fwrite(a[q] + cell(Z, 0, j, k), sizeof(double), (size_t)Z.nx, f.get()) == (size_t)Z.nx;
    if (!ok) std::cerr << "Failed writing " << path << "\n";
    return ok;
}

bool read_restart(const std::string &path, RestartInfo &info, const ZoneDesc &Z, double *const a[5])
{
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) {
        std::cerr << "Error: cannot open restart file: " << path << "\n";
        return false;
    }
    char magic[8];
    int32_t dims[4];
    int64_t step;
    double time, next_output = std::nan("");
    const bool head = std::fread(magic, 1, 8, f.get()) == 8;
    const bool v2 = head && std::memcmp(magic, kRestartMagic, 8) == 0;
    const bool v1 = head && std::memcmp(magic, kRestartMagicV1, 8) == 0;
    if (!(v1 || v2) || std::fread(dims, sizeof(int32_t), 4, f.get()) != 4 ||
        std::fread(&step, sizeof(step), 1, f.get()) != 1 || std::fread(&time, sizeof(double), 1, f.get()) != 1 ||
        (v2 && std::fread(&next_output, sizeof(double), 1, f.get()) != 1)) {
        std::cerr << "Error: " << path << " is not a restart file\n";
        return false;
    }
    if (dims[0] != Z.nx || dims[1] != Z.ny || dims[2] != Z.nz || (dims[3] != 0 && dims[3] != 1)) {
        std::cerr << "Error: restart file " << path << " holds a " << dims[0] << "x" << dims[1] << "x" << dims[2]
                  << " grid (kind " << dims[3] << "), the solver has " << Z.nx << "x" << Z.ny << "x" << Z.nz << "\n";
        return false;
    }
    for (int q = 0; q < 5; ++q)
        for (int k = 0; k < Z.nz; ++k)
            for (int j = 0; j < Z.ny; ++j)
                if (std::fread(a[q] + cell(Z, 0, j, k), sizeof(double), (size_t)Z.nx, f.get()) != (size_t)Z.nx) {
                    std::cerr << "Error: restart file " << path << " is truncated\n";
                    return false;
                }
    info = RestartInfo{dims[0], dims[1], dims[2], dims[3], (long long)step, time, next_output};
    return true;
}

}  // namespace ns3d_host
