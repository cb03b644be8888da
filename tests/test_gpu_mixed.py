"""Mixed-precision path (FP32 reconstruction / gradients over an FP64 state)
against the FP64 path on the GPU and against the oracle.

Tolerances follow from the design: the FP32 shadow stores deviations from
the reference state (rho-1, p-p_inf, T-1), so the state entering each RHS
carries ~1e-7 relative rounding; per-step RHS errors stay near 1e-6 of the
RHS scale and the accumulated error after tens of steps is ~1e-6 of U.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from ns3d_cpp_amd import fields  # noqa: E402
from oracle import Oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2)) / (np.sqrt(np.sum(b * b)) + 1e-300))


def solvers(n, state, **cfg):
    out = []
    for prec in ("fp64", "mixed"):
        p, g = ns.make_params(n, precision=prec, **cfg)
        s = ns.Solver(p, g)
        s.upload_primitive(*state)
        out.append(s)
    return out


CFG = dict(Re=200.0, Pr=0.72, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th",
           sutherland_s=110.4 / 273.15, cfl=0.5)


def test_mixed_upload_download(gpu):
    n = 16
    st, Ma = fields.hit(n, k0=2, Mt=0.3, seed=5)
    d, m = solvers(n, st, Ma=Ma, **CFG)
    pd, pm = d.primitive(), m.primitive()
    for k in ("rho", "u", "v", "w", "p"):
        assert rel_l2(m.interior(pm[k]), d.interior(pd[k])) < 1e-14, k
    for a, b in zip(m.conserved(), d.conserved()):
        assert rel_l2(m.interior(a), d.interior(b)) < 1e-14


def test_mixed_free_stream_is_exact(gpu):
    n = 16
    st = [np.full((n, n, n), v) for v in (1.1, 0.3, -0.2, 0.25, 30.0)]
    p, g = ns.make_params(n, precision="mixed", Ma=0.17, **CFG)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    for r in s.compute_rhs():
        assert np.all(s.interior(r) == 0.0)


@pytest.mark.parametrize("recon,fvs", [("WENO5", "StagerWarming"), ("MDCD", "LaxFriedrichs"),
                                       ("UPWIND_7TH", "StagerWarming"), ("UD7L", "LaxFriedrichs"),
                                       ("MDCD_HYBRID", "StagerWarming"), ("LINEAR", "LaxFriedrichs")])
def test_mixed_rhs_vs_oracle(gpu, recon, fvs):
    """One RHS of a 32^3 isotropic-turbulence state: per-component relative
    max error against the FP64 oracle <= 2e-5 of the component's scale."""
    n = 32
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=9)
    cfg = dict(CFG, recon=recon, fvs_type=fvs)
    p, g = ns.make_params(n, precision="mixed", Ma=Ma, **cfg)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    pr, gr = ns.make_params(n, Ma=Ma, **cfg)
    o = Oracle(pr, gr)
    o.set_primitive(*st)
    for m, (a, b) in enumerate(zip(s.compute_rhs(), o.compute_rhs())):
        a, b = s.interior(a), o.interior(b)
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err < 2e-5, (m, err)


def test_mixed_vs_fp64_trajectory(gpu):
    """64^3 isotropic turbulence, 50 RK3 steps with the FP64 path's CFL dt:
    rel-L2 of every conserved component <= 1e-5."""
    n = 64
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=21)
    d, m = solvers(n, st, Ma=Ma, **CFG)
    for _ in range(50):
        dt = d.timestep()
        d.rk3_step(dt)
        m.rk3_step(dt)
    errs = [rel_l2(m.interior(a), d.interior(b)) for a, b in zip(m.conserved(), d.conserved())]
    assert max(errs) < 1e-5, errs


def test_mixed_conservation_full_size(gpu):
    """256^3: the FP64 face-flux differences telescope; the FP32 accumulator
    adds rounding only, so sum(RHS) stays ~1e-6 of sum|RHS|."""
    n = 256
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=42)
    p, g = ns.make_params(n, precision="mixed", Ma=Ma, **CFG)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    for m, r in enumerate(s.compute_rhs()):
        r = s.interior(r)
        assert abs(r.sum()) <= 1e-6 * np.abs(r).sum(), m


def test_mixed_rejects_fp64_only_entry_points(gpu):
    n = 16
    st, Ma = fields.hit(n, k0=2, Mt=0.3, seed=1)
    p, g = ns.make_params(n, precision="mixed", Ma=Ma, **CFG)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    with pytest.raises(ns.NS3DError) as e:
        s.du_inviscous()
    assert e.value.status == ns._lib.NS3D_ERR_UNSUPPORTED


@pytest.mark.parametrize("recon,fvs", [("WENO5", "StagerWarming"), ("MDCD", "LaxFriedrichs"),
                                       ("UPWIND_7TH", "StagerWarming")])
def test_mixed_characteristic_rhs_vs_oracle(gpu, recon, fvs):
    """char_recon in mixed mode (du_inviscous_character, InviscousFlux_new.cpp:
    165-437, in FP64 on the widened shadow): one RHS of a 32^3 isotropic
    state within 2e-5 of the FP64 oracle's."""
    n = 32
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=19)
    cfg = dict(CFG, recon=recon, fvs_type=fvs, char_recon=1)
    p, g = ns.make_params(n, precision="mixed", Ma=Ma, **cfg)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    pr, gr = ns.make_params(n, Ma=Ma, **cfg)
    o = Oracle(pr, gr)
    o.set_primitive(*st)
    for m, (a, b) in enumerate(zip(s.compute_rhs(), o.compute_rhs())):
        a, b = s.interior(a), o.interior(b)
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err < 2e-5, (m, err)


def test_mixed_characteristic_trajectory(gpu):
    """30 RK3 steps of char_recon WENO5 at 32^3, mixed against FP64 with the
    FP64 path's dt: rel-L2 of U <= 1e-5; a non-periodic variant (walls in y)
    as well."""
    n = 32
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=23)
    for bc in (None, ["Periodic", "Periodic", "Wall", "Wall", "Periodic", "Periodic"]):
        cfg = dict(CFG, char_recon=1, Ma=Ma)
        if bc:
            cfg["bc"] = bc
        d, m = solvers(n, st, **cfg)
        for _ in range(30):
            dt = d.timestep()
            d.rk3_step(dt)
            m.rk3_step(dt)
        errs = [rel_l2(m.interior(a), d.interior(b)) for a, b in zip(m.conserved(), d.conserved())]
        assert max(errs) < 1e-5, (bc, errs)


@pytest.mark.parametrize("vis", ["C8th", "C6th", "C4th"])
def test_viscous_pass_variants_bitwise(gpu, vis, monkeypatch):
    """The mixed viscous passes have several forms on periodic boxes.  The
    gradient/flux pass has one row per thread with cached stencil loads, two
    rows per thread sharing the y-neighbour loads, and two rows per thread
    with each plane's x / y neighbourhood staged in LDS.  The divergence and
    stage pass has cached stencil loads or an LDS-staged plane, the latter
    writing the new shadow's periodic ghost images itself (no fill pass) or
    not.  The stage pass can also fold in the z sweep, writing the new shadow
    into a second buffer (the two alternate, each with its own step graph).
    The LDS-staged passes run in workgroups of 64 x 8 threads or 64 x 4
    (NS3D_M_TILE_Y).  All of them evaluate the same arithmetic on the same values, so the RHS,
    two RK3 steps at a fixed dt, three graph-replayed steps, four CFL steps
    with the device time step (the third stage reducing the next CFL
    minimum) and the downloaded primitives, ghost layers included, must
    agree bit for bit.  The box is 64 x 32 x 24 (a different spacing per
    axis) in 2 z segments of unequal length: tile shapes and segment ends
    that the 256^3 workload does not exercise."""
    out = _mixed_forms(vis, CFG, monkeypatch, [("1", "0", "0", "0"), ("2", "0", "0", "0"), ("3", "0", "0", "0"),
                                               ("3", "0", "0", "0", "4"),
                                               ("1", "1", "0", "0"), ("3", "1", "0", "0"), ("3", "1", "1", "0"),
                                               ("3", "1", "1", "1"), ("3", "1", "0", "0", "4"),
                                               ("3", "1", "1", "1", "4")])
    _assert_forms_equal(out)


def _mixed_forms(vis, cfg, monkeypatch, forms):
    nx, ny, nz = 64, 32, 24
    rng = np.random.default_rng(7)
    st, Ma = fields.hit(nz, k0=2, Mt=0.3, seed=11)
    shape = (nz, ny, nx)
    x = [np.resize(a, shape) * (1.0 + 0.01 * rng.standard_normal(shape)) for a in st]
    out = {}
    for form in forms:
        rows, lds, img, zf = form[:4]
        monkeypatch.setenv("NS3D_GF_ROWS", rows)
        monkeypatch.setenv("NS3D_VD_LDS", lds)
        monkeypatch.setenv("NS3D_VD_IMG", img)
        monkeypatch.setenv("NS3D_ZFUSE", zf)
        monkeypatch.setenv("NS3D_M_TILE_Y", form[4] if len(form) > 4 else "8")
        monkeypatch.setenv("NS3D_SEG_Z", "16")
        p, g = ns.make_params((nx, ny, nz), precision="mixed", Ma=Ma, **dict(cfg, vis_scheme=vis))
        s = ns.Solver(p, g)
        s.upload_primitive(*x)
        r = [s.interior(a).copy() for a in s.compute_rhs()]
        s.rk3_step(1e-3)
        s.rk3_step(1e-3)
        u = [s.interior(a).copy() for a in s.conserved()]
        s.advance(3, 1e-3)
        u1 = [s.interior(a).copy() for a in s.conserved()]
        t = s.advance_cfl(4)
        u2 = [s.interior(a).copy() for a in s.conserved()]
        w = list(s.primitive().values())
        s.close()
        out[tuple(form)] = (r + u + u1 + u2 + w, t)
    return out


def _assert_forms_equal(out):
    forms = list(out)
    ref = out[forms[0]]
    for f in forms[1:]:
        assert out[f][1] == ref[1], f
        for a, b in zip(out[f][0], ref[0]):
            assert np.array_equal(a, b), f


@pytest.mark.parametrize("recon,fvs,eps", [("WENO5", "LaxFriedrichs", 0.0), ("WENO5", "StagerWarming", 0.2),
                                           ("UPWIND_7TH", "StagerWarming", 0.0),
                                           ("UPWIND_7TH", "LaxFriedrichs", 0.0)])
def test_z_sweep_folded_into_stage_bitwise(gpu, recon, fvs, eps, monkeypatch):
    """The stage kernel with the z sweep folded in against the separate z
    sweep and stage kernels, bit for bit, for the splits and
    reconstructions it is built for."""
    cfg = dict(CFG, recon=recon, fvs_type=fvs, sw_eps=eps)
    out = _mixed_forms("C8th", cfg, monkeypatch, [("3", "1", "1", "0"), ("3", "1", "1", "1"),
                                                  ("3", "1", "1", "1", "4")])
    _assert_forms_equal(out)


@pytest.mark.parametrize("recon,fvs,eps", [("WENO5", "StagerWarming", 0.0), ("WENO5", "LaxFriedrichs", 0.0),
                                           ("WENO5", "StagerWarming", 0.1), ("UPWIND_7TH", "StagerWarming", 0.0)])
def test_one_dimensional_flow_is_the_same_along_each_axis(gpu, recon, fvs, eps):
    """A flow that varies along one axis only, with velocity along that axis
    only, on a periodic 64^3 cube: the x sweep (whole rows in LDS, WENO5's
    smoothness terms computed once per slot), the y and z sweeps (marching
    columns, the terms carried from face to face), the LDS-staged viscous
    passes and the stage update must give the same numbers along each axis,
    bit for bit: the RHS and three CFL steps (the tangential velocities stay
    exactly zero, so every sum meets the same nonzero terms in the same
    order)."""
    n = 64
    x = np.arange(n) * (2 * np.pi / n)
    rng = np.random.default_rng(3)
    r = 1.0 + 0.05 * np.sin(x) + 0.01 * rng.standard_normal(n)
    un = 0.3 * np.cos(2 * x) + 0.02 * rng.standard_normal(n)
    pr = (1.0 + 0.1 * np.sin(3 * x + 0.4)) / (1.4 * 0.3 ** 2)
    out = []
    for axis in range(3):
        shape = [1, 1, 1]
        shape[2 - axis] = n      # arrays are (k, j, i)
        b = lambda a: np.broadcast_to(a.reshape(shape), (n, n, n)).copy()  # noqa: E731
        vel = [np.zeros((n, n, n)) for _ in range(3)]
        vel[axis] = b(un)
        p, g = ns.make_params(n, precision="mixed", Ma=0.3, Re=500.0, Pr=0.72, recon=recon, fvs_type=fvs,
                              vis_scheme="C8th", sutherland_s=110.4 / 273.15, cfl=0.5, sw_eps=eps)
        s = ns.Solver(p, g)
        s.upload_primitive(b(r), vel[0], vel[1], vel[2], b(pr))
        rhs = [s.interior(a).copy() for a in s.compute_rhs()]
        t = s.advance_cfl(3)
        u = [s.interior(a).copy() for a in s.conserved()]
        s.close()
        out.append((rhs, u, t))
    perm = {1: (0, 2, 1), 2: (2, 1, 0)}
    for axis in (1, 2):
        assert out[axis][2] == out[0][2], axis
        for ref, got in ((out[0][0], out[axis][0]), (out[0][1], out[axis][1])):
            # the x run's normal momentum (component 1) is this run's 1 + axis
            order = list(range(5))
            order[1], order[1 + axis] = order[1 + axis], order[1]
            for m in range(5):
                want = np.transpose(ref[order[m]], perm[axis])
                assert np.array_equal(got[m], want), (axis, m)
