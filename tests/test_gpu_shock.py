"""Sod's shock tube along each axis (outflow ends, periodic cross axes):
strong discontinuities drive the WENO5 weights far from their linear values
and the split fluxes through large cell-to-cell jumps, which the smooth and
noisy states of the other tests do not.  The FP64 path follows the oracle
(bounded_fill + du_inviscous on non-periodic axes, boundary.cpp:58-207,
InviscousFlux_new.cpp:4-163); the mixed path stays positive and close to the
FP64 path; the x, y and z runs are the same solution, permuted."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from oracle import Oracle  # noqa: E402

pytestmark = pytest.mark.gpu

N, M = 64, 8          # cells along the tube, across it
STEPS = 20


def setup(axis, fvs, precision="fp64", p_left=1.0):
    shape = [M, M, M]
    shape[axis] = N
    L = [M / N, M / N, M / N]
    L[axis] = 1.0
    bc = ["Periodic"] * 6
    bc[2 * axis] = bc[2 * axis + 1] = "Outflow"
    p, g = ns.make_params(tuple(shape), Lx=L[0], Ly=L[1], Lz=L[2], Re=1e4, Ma=1.0, Pr=0.72, recon="WENO5",
                          fvs_type=fvs, vis_scheme="C8th", cfl=0.5, bc=bc, precision=precision)
    nx, ny, nz = shape
    idx = np.indices((nz, ny, nx))[2 - axis]          # cell index along the tube
    left = idx < N // 2
    rho = np.where(left, 1.0, 0.125)
    pr = np.where(left, p_left, 0.1) * p.Rgas         # p_left = 1: T = 1 on the left
    z = np.zeros_like(rho)
    return p, g, [rho, z.copy(), z.copy(), z.copy(), pr]


def run(p, g, st, oracle=False):
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    o = None
    if oracle:
        o = Oracle(p, g)
        o.set_primitive(*st)
    for _ in range(STEPS):
        dt = s.timestep()
        s.rk3_step(dt)
        if o is not None:
            o.step(dt)
    return s, o


def relmax(a, b):
    return np.abs(a - b).max() / (np.abs(b).max() + 1e-300)


# Sod's pressure ratio 10, and 1000 (the expansion then turns supersonic,
# Mach 1.47, so both splits' supersonic branches are taken)
STRENGTH = {"sod": 1.0, "p1000": 100.0}


@pytest.mark.parametrize("strength", list(STRENGTH))
@pytest.mark.parametrize("fvs", ["StagerWarming", "LaxFriedrichs"])
def test_sod_fp64_matches_oracle(gpu, fvs, strength):
    p, g, st = setup(0, fvs, p_left=STRENGTH[strength])
    s, o = run(p, g, st, oracle=True)
    for m, (a, b) in enumerate(zip(s.conserved(), o.conserved())):
        assert relmax(s.interior(a), o.interior(b)) <= 1e-12, m
    # a shock has formed: the density has a steep jump right of the middle
    rho = s.interior(s.primitive()["rho"])[M // 2, M // 2, :]
    assert np.max(np.abs(np.diff(rho))) > 0.05


@pytest.mark.parametrize("strength", list(STRENGTH))
@pytest.mark.parametrize("fvs", ["StagerWarming", "LaxFriedrichs"])
def test_sod_mixed_close_to_fp64(gpu, fvs, strength):
    p, g, st = setup(0, fvs, p_left=STRENGTH[strength])
    d, _ = run(p, g, st)
    pm, gm, _ = setup(0, fvs, "mixed", p_left=STRENGTH[strength])
    m, _ = run(pm, gm, st)
    prim = m.primitive()
    for c in ("rho", "p", "T"):
        v = m.interior(prim[c])
        assert np.all(np.isfinite(v)) and v.min() > 0.0, c
    errs = []
    for a, b in zip(m.conserved(), d.conserved()):
        a, b = d.interior(a), d.interior(b)
        errs.append(float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300))))
    print("sod mixed vs fp64 rel-L2", fvs, strength, errs)
    # rho, rho u, E; rho v and rho w stay zero to rounding in both
    for q in (0, 1, 4):
        assert errs[q] <= 1e-5, (q, errs)


@pytest.mark.parametrize("axis", [1, 2])
def test_sod_along_each_axis_is_the_same_solution(gpu, axis):
    """The tube along y or z gives the x run's fields, permuted."""
    px, gx, stx = setup(0, "StagerWarming")
    sx, _ = run(px, gx, stx)
    pa, ga, sta = setup(axis, "StagerWarming")
    sa, _ = run(pa, ga, sta)
    ux = [sx.interior(a) for a in sx.conserved()]
    ua = [sa.interior(a) for a in sa.conserved()]
    # arrays are (k, j, i): move the tube axis of the x run (last) to
    # `axis`, and its tube-normal momentum rho u to rho v or rho w
    perm = (0, 2, 1) if axis == 1 else (2, 1, 0)
    order = [0, 2, 1, 3, 4] if axis == 1 else [0, 3, 2, 1, 4]
    for m in range(5):
        want = np.transpose(ux[order[m]], perm)
        assert relmax(ua[m], want) <= 1e-13, m
