"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and its host logic (parameter derivation, solver.in reader,
argument checking) matches the reference's semantics.  No kernel runs here."""
import ctypes as C
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from ns3d_cpp_amd import _lib  # noqa: E402
from oracle import load as load_oracle  # noqa: E402


def test_library_exports_header_symbols():
    lib = ns.lib()
    declared = _lib.header_symbols()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(lib, s)
    assert set(declared) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("recon,vis,ng,stencil", [
    ("WENO5", "C8th", 4, 6), ("WENO5", "C4th", 3, 6), ("LINEAR", "C4th", 2, 2), ("LINEAR", "C6th", 3, 2),
    ("MDCD", "C6th", 3, 6), ("UPWIND_7TH", "C4th", 4, 8), ("UD7L", "C8th", 4, 8), ("MDCD_HYBRID", "C8th", 4, 6)])
def test_params_finalize_matches_reference_rules(recon, vis, ng, stencil):
    # initialization.cpp:329-372 ghost/stencil rules and :374-377 constants
    p, g = ns.make_params(32, recon=recon, vis_scheme=vis, Ma=0.3, Re=100.0, gamma=1.4)
    assert (p.ghost_layers, p.stencil, g.ng) == (ng, stencil, ng)
    assert p.Cv == 1.0 / (1.4 * (1.4 - 1.0) * 0.3 * 0.3)
    assert p.Cp == p.Cv * 1.4
    assert p.Rgas == 1.0 / (0.3 * 0.3 * 1.4)
    assert p.mu == 1.0 / 100.0
    assert g.dx == g.Lx / 32
    # the oracle's own restatement agrees
    orc = load_oracle()
    q, h = ns.Params(), ns.Grid()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    C.memmove(C.byref(h), C.byref(g), C.sizeof(g))
    q.ghost_layers = 0
    q.Rgas = 0.0
    orc.orc_params_finalize(C.byref(q), C.byref(h))
    assert (q.ghost_layers, q.stencil, q.Rgas, q.Cp, h.ng) == (p.ghost_layers, p.stencil, p.Rgas, p.Cp, g.ng)


def test_defaults_are_solverparams_defaults():
    lib = ns.lib()
    p, g = ns.Params(), ns.Grid()
    lib.ns3d_params_default(C.byref(p), C.byref(g))
    # field_structures.h:86-165 member initialisers
    assert (p.gamma, p.Pr, p.Ma, p.Re, p.cfl, p.dt_fixed) == (1.4, 0.71, 0.4, 500.0, 0.1, -1.0)
    assert (p.fvs_type, p.recon, p.vis_scheme, p.char_recon) == (0, 0, 1, 0)
    assert (p.mdcd_diss, p.mdcd_disp, p.ud7l_alpha) == (0.01, 0.0463783, 0.4)
    assert list(p.bc) == [0] * 6
    assert p.sutherland_s == 110.4 / 288.15 and p.mu_law == 0 and p.sw_eps == 0.0
    assert (g.nx, g.ny, g.nz) == (16, 16, 16)


def test_solver_in_reader():
    lib = ns.lib()
    p, g = ns.Params(), ns.Grid()
    st = lib.ns3d_params_from_file(os.path.join(ROOT, "tests", "golden", "solver_sample.in").encode(),
                                   C.byref(p), C.byref(g))
    assert st == 0
    assert (p.Re, p.Ma, p.gamma, p.Pr, p.cfl, p.dt_fixed) == (1600.0, 0.1, 1.4, 0.72, 0.5, 0.001)
    assert p.fvs_type == _lib.FVS["StagerWarming"]
    assert p.recon == _lib.RECON["UPWIND_7TH"] and p.stencil == 8
    assert p.vis_scheme == _lib.VIS["C6th"] and p.ghost_layers == 4 and g.ng == 4
    assert p.char_recon == 1
    assert (p.mdcd_disp, p.mdcd_diss, p.ud7l_alpha) == (0.05, 0.02, 0.75)
    assert (g.nx, g.ny, g.nz) == (24, 20, 16)
    # parse_bc: unknown words mean Periodic (initialization.cpp:288-294)
    assert list(p.bc) == [0, 0, 1, 3, 0, 4]
    assert (g.Lx, g.Ly, g.Lz, g.x0) == (6.283185307179586, 3.0, 2.0, 0.5)
    assert g.dx == 6.283185307179586 / 24 and g.dy == 3.0 / 20
    assert (p.max_steps, p.monitor_stepfreq, p.output_timefreq, p.total_time) == (123, 7, 0.25, 2.5)
    assert p.sutherland_s == 0.4041735
    assert p.Rgas == 1.0 / (0.1 * 0.1 * 1.4)


def test_solver_in_missing_file_and_bad_number(tmp_path):
    lib = ns.lib()
    p, g = ns.Params(), ns.Grid()
    assert lib.ns3d_params_from_file(str(tmp_path / "nope.in").encode(), C.byref(p), C.byref(g)) == _lib.NS3D_ERR_IO
    f = tmp_path / "bad.in"
    f.write_text("Re = not-a-number\n")
    assert lib.ns3d_params_from_file(str(f).encode(), C.byref(p), C.byref(g)) == _lib.NS3D_ERR_CONFIG


def test_create_argument_errors():
    lib = ns.lib()
    p, g = ns.make_params(16)
    ctx = C.c_void_p()
    assert lib.ns3d_create(None, C.byref(g), C.byref(ctx)) == _lib.NS3D_ERR_INVALID_ARG
    q, h = ns.make_params(16)
    q.recon = 99
    assert lib.ns3d_create(C.byref(q), C.byref(h), C.byref(ctx)) == _lib.NS3D_ERR_INVALID_ARG
    q, h = ns.make_params(16, recon="WCNS")
    assert lib.ns3d_create(C.byref(q), C.byref(h), C.byref(ctx)) == _lib.NS3D_ERR_CONFIG
    # a periodic side facing a non-periodic one (validated before any device call)
    q, h = ns.make_params(16, bc=["Periodic", "Wall", "Periodic", "Periodic", "Periodic", "Periodic"])
    assert lib.ns3d_create(C.byref(q), C.byref(h), C.byref(ctx)) == _lib.NS3D_ERR_CONFIG
    # 32-bit cell indices: a padded array of 2^31 cells or more is refused
    # (validated before any device call)
    q, h = ns.make_params(1291, recon="WENO5", vis_scheme="C8th")
    assert lib.ns3d_create(C.byref(q), C.byref(h), C.byref(ctx)) == _lib.NS3D_ERR_INVALID_ARG
    q, h = ns.make_params(2, recon="WENO5", vis_scheme="C8th")     # n < ng
    assert lib.ns3d_create(C.byref(q), C.byref(h), C.byref(ctx)) == _lib.NS3D_ERR_GRID_TOO_SMALL
    for s in range(0, 9):
        assert lib.ns3d_status_string(s)


def test_recon_fallback_past_the_stencil_is_refused():
    """reconstruct_select's averaging fallback (reconstruction.cpp:77-84, the
    WCNS value) reads vstencil[mid + 1] on the minus side; with two values
    that is past the stencil.  The entry point refuses it before any device
    work, and the oracle reports it; width 3 and the plus side are served."""
    lib = ns.lib()
    p, _ = ns.make_params(8)
    p.recon = 1   # WCNS: no scheme of its own, the fallback
    st = (C.c_double * 6)(1.0, 2.0, 3.0, 4.0, 5.0, 6.0)
    out = (C.c_double * 3)()
    assert lib.ns3d_reconstruct_batch(C.byref(p), 3, 2, st, -1.0, out) == _lib.NS3D_ERR_CONFIG
    orc = load_oracle()
    err = C.c_int(0)
    v = orc.orc_reconstruct(st, 2, -1.0, C.byref(p), C.byref(err))
    assert err.value == 1 and v != v
    err.value = 0
    assert orc.orc_reconstruct(st, 2, 1.0, C.byref(p), C.byref(err)) == 0.5 * (2.0 + 1.0) and err.value == 0
    assert orc.orc_reconstruct(st, 3, -1.0, C.byref(p), C.byref(err)) == 0.5 * (2.0 + 3.0) and err.value == 0
