"""Non-periodic boundaries on the GPU against the CPU oracle: the ghost fills
of apply_wall_bc / apply_symmetry_bc / apply_outflow_bc / apply_inflow_bc
(boundary.cpp:58-207) and the one-sided derivative closures of
choose_scheme (ViscousFlux.cpp:43-118), through the FP64 path (bit-level
agreement expected, tolerances per test) and the mixed path (against FP64).

The oracle's fills and closures are pinned by their own tests
(test_oracle_pins.py: closure orders and the one-sided formula's truncation
error); here the boxes are non-cubic so that every axis has its own size."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from oracle import Oracle  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPE = (20, 16, 12)          # nx, ny, nz
MA = 1.0 / np.sqrt(1.4)       # p_inf = 1, the inflow state's pressure
BCS = {
    "channel": ["Periodic", "Periodic", "Wall", "Wall", "Outflow", "Outflow"],
    "slab": ["Outflow", "Outflow", "Periodic", "Periodic", "Symmetry", "Symmetry"],
    "box": ["Wall", "Wall", "Symmetry", "Symmetry", "Wall", "Outflow"],
    "inflow": ["Inflow", "Outflow", "Periodic", "Periodic", "Wall", "Wall"],
}


def field(seed=0, u0=0.0):
    nx, ny, nz = SHAPE
    rng = np.random.default_rng(seed)
    X, Y, Z = (2 * np.pi * np.arange(n) / n for n in (nx, ny, nz))
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    out = []
    for m, (base, amp) in enumerate(((1.0, 0.1), (u0, 0.4), (0.0, 0.4), (0.0, 0.4), (1.0, 0.1))):
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(-2, 3, 3)
            f += np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        f += 0.3 * rng.standard_normal(X.shape)
        out.append(base + amp * f)
    return out


def pair(bcs, state, **cfg):
    kw = dict(Re=100.0, Ma=MA, Pr=0.72, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th")
    kw.update(cfg)
    p, g = ns.make_params(SHAPE, bc=bcs, **kw)
    s = ns.Solver(p, g)
    s.upload_primitive(*state)
    o = Oracle(p, g)
    o.set_primitive(*state)
    return s, o


def relmax(a, b):
    return np.abs(a - b).max() / (np.abs(b).max() + 1e-300)


def rel_l2(a, b):
    return np.sqrt(np.sum((a - b) ** 2)) / (np.sqrt(np.sum(b * b)) + 1e-300)


def face_slabs(a, ng):
    """The six ghost slabs with interior cross ranges, (z, y, x) order."""
    i = slice(ng, -ng)
    return [a[i, i, :ng], a[i, i, -ng:], a[i, :ng, i], a[i, -ng:, i], a[:ng, i, i], a[-ng:, i, i]]


@pytest.mark.parametrize("case", list(BCS))
def test_ghost_fill(gpu, case):
    """Interior and face ghosts of rho, u, v, w, p, T after the upload's
    boundary fill: identical to the oracle."""
    s, o = pair(BCS[case], field(1, u0=1.0 if case == "inflow" else 0.0))
    gp, op = s.primitive(), o.primitive()
    for k in ("rho", "u", "v", "w", "p", "T"):
        assert np.array_equal(s.interior(gp[k]), o.interior(op[k])), k
        for f, (a, b) in enumerate(zip(face_slabs(gp[k], s.ng), face_slabs(op[k], s.ng))):
            assert np.array_equal(a, b), (k, f)


def test_wall_mirror_and_outflow_copy(gpu):
    """Closed forms of the fills on the GPU alone: a wall ghost layer is the
    mirrored interior with the normal velocity negated, an outflow ghost
    repeats the last interior cell, the inflow ghosts hold (1, 1, 0, 0, 1)."""
    st = field(2, u0=1.0)
    s, _ = pair(BCS["inflow"], st)
    pr = s.primitive()
    ng = s.ng
    i = slice(ng, -ng)
    # z walls (axis 0 of the (z, y, x) arrays), low and high side
    for k in ("rho", "u", "v", "p"):
        assert np.array_equal(pr[k][ng - 1::-1, i, i], pr[k][ng:2 * ng, i, i]), k
        assert np.array_equal(pr[k][-ng:, i, i], pr[k][-ng - 1:-2 * ng - 1:-1, i, i]), k
    assert np.array_equal(pr["w"][ng - 1::-1, i, i], -pr["w"][ng:2 * ng, i, i])
    assert np.array_equal(pr["w"][-ng:, i, i], -pr["w"][-ng - 1:-2 * ng - 1:-1, i, i])
    # x outflow at xmax, inflow at xmin
    for k in ("rho", "u", "v", "w", "p"):
        last = pr[k][i, i, -ng - 1:-ng]
        assert np.array_equal(pr[k][i, i, -ng:], np.repeat(last, ng, axis=2)), k
    for k, want in zip(("rho", "u", "v", "w", "p"), (1.0, 1.0, 0.0, 0.0, 1.0)):
        assert np.all(pr[k][i, i, :ng] == want), k


@pytest.mark.parametrize("vis", ["C8th", "C6th", "C4th"])
@pytest.mark.parametrize("case", ["channel", "box"])
def test_gradients_with_closures(gpu, case, vis):
    """compute_gradients near non-periodic ends (one-sided and reduced-order
    central formulas): identical to the oracle."""
    s, o = pair(BCS[case], field(3), vis_scheme=vis)
    gg, og = s.compute_gradients(), o.compute_gradients()
    for m in range(12):
        assert np.array_equal(s.interior(gg[m]), o.interior(og[m])), m


RHS_CFGS = [dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th"),
            dict(recon="MDCD", fvs_type="LaxFriedrichs", vis_scheme="C6th"),
            dict(recon="UPWIND_7TH", fvs_type="StagerWarming", vis_scheme="C4th"),
            dict(recon="MDCD_HYBRID", fvs_type="LaxFriedrichs", vis_scheme="C8th"),
            dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", char_recon=1)]


@pytest.mark.parametrize("cfg", RHS_CFGS, ids=lambda c: "-".join(str(v) for v in c.values()))
@pytest.mark.parametrize("case", list(BCS))
def test_rhs(gpu, case, cfg):
    """compute_rhs with the boundary fills and closures; 1e-13 rel."""
    s, o = pair(BCS[case], field(4, u0=1.0 if case == "inflow" else 0.0), **cfg)
    for m, (a, b) in enumerate(zip(s.compute_rhs(), o.compute_rhs())):
        err = relmax(s.interior(a), o.interior(b))
        assert err <= 1e-13, (m, err)


@pytest.mark.parametrize("case", list(BCS))
def test_steps(gpu, case):
    """Six RK3 steps with the CFL dt (both sides use the GPU's): rel-L2(U)
    <= 1e-12 and the same dt."""
    s, o = pair(BCS[case], field(5, u0=1.0 if case == "inflow" else 0.0), cfl=0.4)
    for _ in range(6):
        dt = s.timestep()
        assert dt == pytest.approx(o.timestep(), rel=1e-14)
        s.rk3_step(dt)
        o.step(dt)
    errs = [rel_l2(s.interior(a), o.interior(b)) for a, b in zip(s.conserved(), o.conserved())]
    assert max(errs) <= 1e-12, errs


@pytest.mark.parametrize("case", list(BCS))
def test_mixed_against_fp64(gpu, case):
    """The mixed path with the same fills and closures: one RHS within 2e-5
    of the FP64 one (per-component max relative to scale) and 10 steps with
    the FP64 path's dt within 1e-5 rel-L2 on U."""
    st = field(6, u0=1.0 if case == "inflow" else 0.0)
    kw = dict(Re=100.0, Ma=MA, Pr=0.72, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", cfl=0.4)
    sv = {}
    for prec in ("fp64", "mixed"):
        p, g = ns.make_params(SHAPE, bc=BCS[case], precision=prec, **kw)
        sv[prec] = ns.Solver(p, g)
        sv[prec].upload_primitive(*st)
    d, m = sv["fp64"], sv["mixed"]
    for k, (a, b) in enumerate(zip(m.compute_rhs(), d.compute_rhs())):
        err = relmax(m.interior(a), d.interior(b))
        assert err < 2e-5, (k, err)
    for _ in range(10):
        dt = d.timestep()
        d.rk3_step(dt)
        m.rk3_step(dt)
    errs = [rel_l2(m.interior(a), d.interior(b)) for a, b in zip(m.conserved(), d.conserved())]
    assert max(errs) < 1e-5, errs


def test_mixed_mdcd_hybrid_nonperiodic_x(gpu):
    """The FP64-reconstruction x rows of the mixed path (MDCD_HYBRID) with a
    non-periodic x axis: one RHS against the oracle within 2e-5."""
    st = field(7)
    kw = dict(Re=100.0, Ma=MA, Pr=0.72, recon="MDCD_HYBRID", fvs_type="StagerWarming", vis_scheme="C8th")
    p, g = ns.make_params(SHAPE, bc=BCS["slab"], precision="mixed", **kw)
    m = ns.Solver(p, g)
    m.upload_primitive(*st)
    pr, gr = ns.make_params(SHAPE, bc=BCS["slab"], **kw)
    o = Oracle(pr, gr)
    o.set_primitive(*st)
    for k, (a, b) in enumerate(zip(m.compute_rhs(), o.compute_rhs())):
        err = relmax(m.interior(a), o.interior(b))
        assert err < 2e-5, (k, err)


def test_mismatched_periodic_pair_is_refused(gpu):
    p, g = ns.make_params(SHAPE, bc=["Periodic", "Wall", "Periodic", "Periodic", "Periodic", "Periodic"])
    with pytest.raises(ns.NS3DError) as e:
        ns.Solver(p, g)
    assert e.value.status == ns._lib.NS3D_ERR_CONFIG
