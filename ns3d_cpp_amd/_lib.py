"""ctypes bindings of include/ns3d_gpu.h (libns3d_gpu.so)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libns3d_gpu.so")

NS3D_OK = 0
NS3D_ERR_INVALID_ARG = 1
NS3D_ERR_CONFIG = 2
NS3D_ERR_GRID_TOO_SMALL = 3
NS3D_ERR_NONPHYSICAL = 4
NS3D_ERR_HIP = 5
NS3D_ERR_OOM = 6
NS3D_ERR_UNSUPPORTED = 7
NS3D_ERR_IO = 8

FVS = {"StagerWarming": 0, "LaxFriedrichs": 1}
RECON = {"WENO5": 0, "WCNS": 1, "LINEAR": 2, "MDCD": 3, "MDCD_HYBRID": 4, "UPWIND_7TH": 5, "UD7L": 6}
VIS = {"C8th": 0, "C6th": 1, "C4th": 2}
BC = {"Periodic": 0, "Wall": 1, "Symmetry": 2, "Outflow": 3, "Inflow": 4}
MU = {"sutherland": 0, "constant": 1, "power": 2}
PREC = {"fp64": 0, "mixed": 1}


class Params(C.Structure):
    """Mirror of ns3d_params (SolverParams, field_structures.h:85-166)."""
    _fields_ = [
        ("gamma", C.c_double), ("Pr", C.c_double), ("Ma", C.c_double), ("Re", C.c_double),
        ("cfl", C.c_double), ("dt_fixed", C.c_double),
        ("fvs_type", C.c_int), ("recon", C.c_int), ("vis_scheme", C.c_int), ("char_recon", C.c_int),
        ("mdcd_diss", C.c_double), ("mdcd_disp", C.c_double), ("ud7l_alpha", C.c_double),
        ("bc", C.c_int * 6),
        ("max_steps", C.c_int), ("monitor_stepfreq", C.c_int),
        ("output_timefreq", C.c_double), ("total_time", C.c_double),
        ("mu_law", C.c_int), ("sutherland_s", C.c_double), ("mu_power", C.c_double),
        ("sw_eps", C.c_double), ("precision", C.c_int),
        ("post_basicfield", C.c_int), ("isotropic_analyse", C.c_int), ("monitor_res", C.c_int),
        ("monitor_energy", C.c_int),
        ("Cv", C.c_double), ("Cp", C.c_double), ("Rgas", C.c_double), ("mu", C.c_double),
        ("stencil", C.c_int), ("ghost_layers", C.c_int),
    ]


class Grid(C.Structure):
    """Mirror of ns3d_grid (GridDesc + LocalDesc, field_structures.h:32-59)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
    ]


DP = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "ns3d_params_default": (None, [C.POINTER(Params), C.POINTER(Grid)]),
    "ns3d_params_finalize": (C.c_int, [C.POINTER(Params), C.POINTER(Grid)]),
    "ns3d_params_from_file": (C.c_int, [C.c_char_p, C.POINTER(Params), C.POINTER(Grid)]),
    "ns3d_create": (C.c_int, [C.POINTER(Params), C.POINTER(Grid), C.POINTER(C.c_void_p)]),
    "ns3d_destroy": (C.c_int, [C.c_void_p]),
    "ns3d_cells": (C.c_longlong, [C.c_void_p]),
    "ns3d_sizes": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 4),
    "ns3d_status_string": (C.c_char_p, [C.c_int]),
    "ns3d_last_error": (C.c_char_p, [C.c_void_p]),
    "ns3d_upload_primitive": (C.c_int, [C.c_void_p] + [DP] * 5),
    "ns3d_upload_conserved": (C.c_int, [C.c_void_p] + [DP] * 5),
    "ns3d_download_primitive": (C.c_int, [C.c_void_p] + [DP] * 6),
    "ns3d_download_conserved": (C.c_int, [C.c_void_p] + [DP] * 5),
    "ns3d_download_rhs": (C.c_int, [C.c_void_p] + [DP] * 5),
    "ns3d_apply_boundary": (C.c_int, [C.c_void_p]),
    "ns3d_du_inviscous": (C.c_int, [C.c_void_p]),
    "ns3d_du_inviscous_character": (C.c_int, [C.c_void_p]),
    "ns3d_split_batch": (C.c_int, [C.POINTER(Params), C.c_longlong, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.c_void_p, C.c_void_p]),
    "ns3d_reconstruct_batch": (C.c_int, [C.POINTER(Params), C.c_longlong, C.c_int, C.c_void_p, C.c_double,
                                         C.c_void_p]),
    "ns3d_download_face_flux": (C.c_int, [C.c_void_p, C.c_int] + [DP] * 5),
    "ns3d_compute_gradients": (C.c_int, [C.c_void_p]),
    "ns3d_download_gradients": (C.c_int, [C.c_void_p, DP * 12]),
    "ns3d_compute_viscous_flux": (C.c_int, [C.c_void_p]),
    "ns3d_download_viscous_flux": (C.c_int, [C.c_void_p, DP * 12]),
    "ns3d_compute_rhs": (C.c_int, [C.c_void_p]),
    "ns3d_compute_vis_flux": (C.c_int, [C.c_void_p]),
    "ns3d_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "ns3d_energy_spectrum": (C.c_int, [C.c_void_p, DP, C.c_int]),
    "ns3d_turbulence_statistics": (C.c_int, [C.c_void_p, DP, C.c_int]),
    "ns3d_residuals": (C.c_int, [C.c_void_p, DP]),
    "ns3d_rk3_step": (C.c_int, [C.c_void_p, C.c_double]),
    "ns3d_nonphysical_cell": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 3),
    "ns3d_advance": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "ns3d_advance_cfl": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "ns3d_compute_timestep_local": (C.c_int, [C.c_void_p, DP]),
    "ns3d_cfl_dt_noalloc": (C.c_int, [C.POINTER(Params), C.POINTER(Grid)] + [C.c_void_p] * 7),
    "ns3d_total_energy": (C.c_int, [C.c_void_p, DP]),
    "ns3d_sync": (C.c_int, [C.c_void_p]),
    "ns3d_stream": (C.c_void_p, [C.c_void_p]),
    "ns3d_copy_probe": (C.c_int, [C.c_void_p, C.c_int, DP]),
    "ns3d_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "ns3d_profile_count": (C.c_int, [C.c_void_p]),
    "ns3d_profile_entry": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int, DP, C.POINTER(C.c_longlong)]),
    "ns3d_tlm_rhs": (C.c_int, [C.c_void_p, DP * 5, DP * 5]),
    "ns3d_adj_rhs": (C.c_int, [C.c_void_p, DP * 5, C.c_double, DP * 5, DP]),
    "ns3d_dissipation": (C.c_int, [C.c_void_p, DP]),
    "ns3d_adj_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, DP]),
    "ns3d_adj_backward": (C.c_int, [C.c_void_p, C.c_void_p, DP * 5, DP]),
    "ns3d_adj_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
}


def header_symbols():
    """Function names declared in include/ns3d_gpu.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "ns3d_gpu.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(ns3d_[a-z0-9_]+)\s*\(", txt)))


_lib = None


def load(path=LIB_PATH):
    """Load libns3d_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m ns3d_cpp_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
