"""ns3d_cfl_dt_noalloc, the entry point that allocates nothing, through the
C ABI: on caller-owned device arrays it gives the CFL time step
ns3d_compute_timestep_local gives for the same state, bit for bit, on two
calls in a row (the second on a smaller box, so the static scratch holds
the first call's partials), and rejects bad arguments."""
import ctypes as C

import numpy as np
import pytest

import ns3d_cpp_amd as ns

pytestmark = pytest.mark.gpu


def _state(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    Ma = 0.3
    return [1.0 + 0.1 * rng.standard_normal((nz, ny, nx)), 0.5 * rng.standard_normal((nz, ny, nx)),
            0.5 * rng.standard_normal((nz, ny, nx)), 0.5 * rng.standard_normal((nz, ny, nx)),
            1.0 / (1.4 * Ma * Ma) + rng.standard_normal((nz, ny, nx))], Ma


def _reference_and_noalloc(shape, seed):
    import torch
    st, Ma = _state(shape, seed)
    p, g = ns.make_params(shape, Ma=Ma, cfl=0.45, recon="WENO5", vis_scheme="C8th")
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    dt_ref = s.timestep()
    prim = s.primitive()
    s.close()
    dev = [torch.from_numpy(np.ascontiguousarray(prim[k])).to("cuda") for k in ("rho", "u", "v", "w", "p")]
    dt = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st_ = ns.lib().ns3d_cfl_dt_noalloc(C.byref(p), C.byref(g), *[C.c_void_p(t.data_ptr()) for t in dev],
                                       C.c_void_p(dt.data_ptr()), None)
    assert st_ == 0, ns.lib().ns3d_status_string(st_)
    torch.cuda.synchronize()
    return dt_ref, float(dt.cpu()[0])


def test_noalloc_cfl_equals_context_timestep_twice(gpu):
    for shape, seed in (((40, 36, 33), 1), ((13, 9, 7), 2)):
        ref, got = _reference_and_noalloc(shape, seed)
        assert got == ref, (shape, got, ref)


def test_noalloc_cfl_rejects_bad_arguments(gpu):
    p, g = ns.make_params(8)
    lib = ns.lib()
    z = C.c_void_p(0)
    assert lib.ns3d_cfl_dt_noalloc(C.byref(p), C.byref(g), z, z, z, z, z, z, None) == ns._lib.NS3D_ERR_INVALID_ARG
    g.nx = 70000
    g.ny = 70000
    one = C.c_void_p(8)
    assert lib.ns3d_cfl_dt_noalloc(C.byref(p), C.byref(g), *([one] * 6), None) == ns._lib.NS3D_ERR_INVALID_ARG
