// ns3d_params.cpp -- SolverParams defaults, derived values and the solver.in
// reader of the C ABI (host C++).
#include <algorithm>
#include <cctype>
#include <fstream>
#include <string>

#include "../../include/ns3d_gpu.h"

namespace {

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(" \t\r\n");
    size_t b = s.find_last_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    return s.substr(a, b - a + 1);
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

// parse_bc of initialization.cpp:288-294 (anything else is Periodic)
// parse_bool (initialization.cpp:28-32), on an already lower-cased value
int parse_bool(const std::string &lv)
{
    return lv == "true" || lv == "yes" || lv == "1" || lv == "on";
}

int parse_bc(const std::string &v)
{
    if (v == "inflow") return NS3D_BC_INFLOW;
    if (v == "wall") return NS3D_BC_WALL;
    if (v == "symmetry") return NS3D_BC_SYMMETRY;
    if (v == "outflow") return NS3D_BC_OUTFLOW;
    return NS3D_BC_PERIODIC;
}

}  // namespace

extern "C" void ns3d_params_default(ns3d_params *p, ns3d_grid *g)
{
    if (p) {
        // SolverParams member initialisers, field_structures.h:86-165
        *p = ns3d_params{};
        p->gamma = 1.4;
        p->Pr = 0.71;
        p->Ma = 0.4;
        p->Re = 500.0;
        p->cfl = 0.1;
        p->dt_fixed = -1.0;
        p->fvs_type = NS3D_FVS_STEGER_WARMING;
        p->recon = NS3D_RECON_WENO5;
        p->vis_scheme = NS3D_VIS_C6TH;
        p->char_recon = 0;
        p->mdcd_diss = 0.01;
        p->mdcd_disp = 0.0463783;
        p->ud7l_alpha = 0.4;
        for (int f = 0; f < 6; ++f) p->bc[f] = NS3D_BC_PERIODIC;
        p->max_steps = 1;
        p->monitor_stepfreq = 1;
        p->output_timefreq = 1.0;
        p->total_time = 1.0;
        p->mu_law = NS3D_MU_SUTHERLAND;
        p->sutherland_s = 110.4 / 288.15;
        p->mu_power = 0.76;
        p->sw_eps = 0.0;
        p->precision = NS3D_PREC_FP64;
        p->post_basicfield = 1;        // field_structures.h:160-165
        p->isotropic_analyse = 1;
        p->monitor_res = 1;
        p->monitor_energy = 1;
        p->mu = 1.0 / p->Re;
        p->Cv = 1.0 / (p->gamma * (p->gamma - 1.0) * p->Ma * p->Ma);
        p->Cp = p->Cv * p->gamma;
        p->Rgas = 1.0 / (p->Ma * p->Ma * p->gamma);
        p->stencil = 6;
        p->ghost_layers = 3;
    }
    if (g) {
        // read_solver_params_from_file grid defaults, initialization.cpp:214-215
        *g = ns3d_grid{};
        g->nx = g->ny = g->nz = 16;
        g->x0 = g->y0 = g->z0 = 0.0;
        g->Lx = g->Ly = g->Lz = 1.0;
        g->dx = g->dy = g->dz = 1.0;
        g->ng = 3;
    }
}

extern "C" int ns3d_params_finalize(ns3d_params *p, ns3d_grid *g)
{
    if (!p || !g) return NS3D_ERR_INVALID_ARG;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return NS3D_ERR_INVALID_ARG;
    // initialization.cpp:324-326
    g->dx = g->Lx / (g->nx);
    g->dy = g->Ly / (g->ny);
    g->dz = g->Lz / (g->nz);
    // initialization.cpp:329-357
    switch (p->recon) {
    case NS3D_RECON_WENO5: p->ghost_layers = 3; p->stencil = 6; break;
    case NS3D_RECON_LINEAR: p->ghost_layers = 2; p->stencil = 2; break;
    case NS3D_RECON_MDCD: p->ghost_layers = 3; p->stencil = 6; break;
    case NS3D_RECON_MDCD_HYBRID: p->ghost_layers = 3; p->stencil = 6; break;
    case NS3D_RECON_UPWIND_7TH: p->ghost_layers = 4; p->stencil = 8; break;
    case NS3D_RECON_UD7L: p->ghost_layers = 4; p->stencil = 8; break;
    case NS3D_RECON_WCNS: break;  // the reference prints an error and keeps going
    default: return NS3D_ERR_INVALID_ARG;
    }
    // initialization.cpp:358-372
    switch (p->vis_scheme) {
    case NS3D_VIS_C4TH: p->ghost_layers = std::max(p->ghost_layers, 2); break;
    case NS3D_VIS_C6TH: p->ghost_layers = std::max(p->ghost_layers, 3); break;
    case NS3D_VIS_C8TH: p->ghost_layers = std::max(p->ghost_layers, 4); break;
    default: return NS3D_ERR_INVALID_ARG;
    }
    // initialization.cpp:374-377
    p->Cv = 1.0 / (p->gamma * (p->gamma - 1.0) * p->Ma * p->Ma);
    p->Cp = p->Cv * p->gamma;
    p->Rgas = 1.0 / (p->Ma * p->Ma * p->gamma);
    p->mu = 1.0 / p->Re;
    g->ng = p->ghost_layers;
    return NS3D_OK;
}

// read_solver_params_from_file, initialization.cpp:199-380
extern "C" int ns3d_params_from_file(const char *path, ns3d_params *p, ns3d_grid *g)
{
    if (!path || !p || !g) return NS3D_ERR_INVALID_ARG;
    std::ifstream fin(path);
    if (!fin.is_open()) return NS3D_ERR_IO;
    ns3d_params_default(p, g);
    std::string line;
    try {
        while (std::getline(fin, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            std::string key = trim(line.substr(0, eq));
            std::string val = trim(line.substr(eq + 1));
            std::string k = lower(key);
            std::string v = lower(val);
            if (k == "gamma") p->gamma = std::stod(val);
            else if (k == "pr") p->Pr = std::stod(val);
            else if (k == "ma") p->Ma = std::stod(val);
            else if (k == "re") p->Re = std::stod(val);
            else if (k == "cfl") p->cfl = std::stod(val);
            else if (k == "dt_fixed") p->dt_fixed = std::stod(val);
            else if (k == "fvs_type") {
                if (v == "stagerwarming") p->fvs_type = NS3D_FVS_STEGER_WARMING;
                else if (v == "laxfriedrichs") p->fvs_type = NS3D_FVS_LAX_FRIEDRICHS;
            } else if (k == "recon") {
                if (v == "mdcd") p->recon = NS3D_RECON_MDCD;
                else if (v == "weno5") p->recon = NS3D_RECON_WENO5;
                else if (v == "linear") p->recon = NS3D_RECON_LINEAR;
                else if (v == "mdcd_hybrid") p->recon = NS3D_RECON_MDCD_HYBRID;
                else if (v == "upwind_7th") p->recon = NS3D_RECON_UPWIND_7TH;
                else if (v == "ud7l") p->recon = NS3D_RECON_UD7L;
            } else if (k == "vis_scheme") {
                if (v == "c4th") p->vis_scheme = NS3D_VIS_C4TH;
                else if (v == "c6th") p->vis_scheme = NS3D_VIS_C6TH;
                else if (v == "c8th") p->vis_scheme = NS3D_VIS_C8TH;
            } else if (k == "char_recon") {
                p->char_recon = (v == "yes" || v == "true");
            } else if (k == "mdcd_diss") p->mdcd_diss = std::stod(val);
            else if (k == "mdcd_disp") p->mdcd_disp = std::stod(val);
            else if (k == "ud7l_alpha") p->ud7l_alpha = std::stod(val);
            else if (k == "global_nx") g->nx = std::stoi(val);
            else if (k == "global_ny") g->ny = std::stoi(val);
            else if (k == "global_nz") g->nz = std::stoi(val);
            else if (k == "x0") g->x0 = std::stod(val);
            else if (k == "y0") g->y0 = std::stod(val);
            else if (k == "z0") g->z0 = std::stod(val);
            else if (k == "lx") g->Lx = std::stod(val);
            else if (k == "ly") g->Ly = std::stod(val);
            else if (k == "lz") g->Lz = std::stod(val);
            else if (k == "max_steps") p->max_steps = std::stoi(val);
            else if (k == "monitor_stepfreq") p->monitor_stepfreq = std::stoi(val);
            else if (k == "output_timefreq") p->output_timefreq = std::stod(val);
            else if (k == "totaltime") p->total_time = std::stod(val);
            // initialization.cpp:304-307 with parse_bool (:28-32)
            else if (k == "post_basicfield") p->post_basicfield = parse_bool(v);
            else if (k == "isotropicanalyse") p->isotropic_analyse = parse_bool(v);
            else if (k == "monitor_res") p->monitor_res = parse_bool(v);
            else if (k == "monitor_energy") p->monitor_energy = parse_bool(v);
            // keys absent from the reference's parser (mu law, SW epsilon, precision)
            else if (k == "mu_law") {
                if (v == "constant") p->mu_law = NS3D_MU_CONSTANT;
                else if (v == "power") p->mu_law = NS3D_MU_POWER;
                else if (v == "sutherland") p->mu_law = NS3D_MU_SUTHERLAND;
            } else if (k == "sutherland_s") p->sutherland_s = std::stod(val);
            else if (k == "mu_power") p->mu_power = std::stod(val);
            else if (k == "sw_eps") p->sw_eps = std::stod(val);
            else if (k == "precision") {
                if (v == "mixed") p->precision = NS3D_PREC_MIXED;
                else if (v == "fp64") p->precision = NS3D_PREC_FP64;
            }
            if (k == "bc_xmin") p->bc[0] = parse_bc(v);
            else if (k == "bc_xmax") p->bc[1] = parse_bc(v);
            else if (k == "bc_ymin") p->bc[2] = parse_bc(v);
            else if (k == "bc_ymax") p->bc[3] = parse_bc(v);
            else if (k == "bc_zmin") p->bc[4] = parse_bc(v);
            else if (k == "bc_zmax") p->bc[5] = parse_bc(v);
        }
    } catch (...) {
        return NS3D_ERR_CONFIG;  // std::stod / std::stoi on a malformed number
    }
    return ns3d_params_finalize(p, g);
}
