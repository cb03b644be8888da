"""Seeded random configurations across the option space (reconstruction,
split, viscous scheme, viscosity law, boundary layout, grid shape, state):
the FP64 path against the oracle on one RHS and two CFL steps, and the mixed
path against the oracle on one RHS.  The draws are fixed by the seed, so
the same 40 cases run every time."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from oracle import Oracle  # noqa: E402

pytestmark = pytest.mark.gpu

RECONS = ["WENO5", "MDCD", "MDCD_HYBRID", "UPWIND_7TH", "UD7L", "LINEAR"]
FVSS = ["StagerWarming", "LaxFriedrichs"]
VISS = ["C8th", "C6th", "C4th"]
BCK = ["Periodic", "Wall", "Symmetry", "Outflow"]


def draw(i):
    rng = np.random.default_rng(1000 + i)
    cfg = dict(recon=RECONS[rng.integers(6)], fvs_type=FVSS[rng.integers(2)], vis_scheme=VISS[rng.integers(3)],
               Re=float(rng.choice([50.0, 400.0, 2000.0])), Ma=float(rng.uniform(0.1, 0.6)), Pr=0.72,
               cfl=float(rng.uniform(0.2, 0.6)))
    law = rng.integers(3)
    if law == 1:
        cfg["mu_law"] = "constant"
    elif law == 2:
        cfg.update(mu_law="power", mu_power=float(rng.uniform(0.5, 0.9)))
    if rng.random() < 0.3:
        cfg["char_recon"] = 1
    if rng.random() < 0.2 and cfg["fvs_type"] == "StagerWarming":
        cfg["sw_eps"] = float(rng.uniform(0.01, 0.2))
    if cfg["recon"] == "UD7L":
        cfg["ud7l_alpha"] = float(rng.uniform(0, 1))
    bc = []
    for axis in range(3):
        if rng.random() < 0.6:
            bc += ["Periodic", "Periodic"]
        else:
            bc += [BCK[rng.integers(1, 4)], BCK[rng.integers(1, 4)]]
    cfg["bc"] = bc
    shape = tuple(int(v) for v in rng.integers(6, 22, 3))
    return cfg, shape, int(rng.integers(1 << 30))


def state(shape, seed, Ma):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    X, Y, Z = (2 * np.pi * np.arange(n) / n for n in (nx, ny, nz))
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    pinf = 1.0 / (1.4 * Ma * Ma)
    out = []
    for base, amp in ((1.0, 0.1), (0.0, 0.4), (0.0, 0.4), (0.0, 0.4), (pinf, 0.1 * pinf)):
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(-2, 3, 3)
            f += np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        f += 0.2 * rng.standard_normal(X.shape)
        out.append(base + amp * f / 3.0)
    return out


def relmax(a, b):
    return np.abs(a - b).max() / (np.abs(b).max() + 1e-300)


@pytest.mark.parametrize("i", range(40))
def test_random_config(gpu, i):
    cfg, shape, seed = draw(i)
    print(i, shape, cfg, flush=True)
    st = state(shape, seed, cfg["Ma"])
    p, g = ns.make_params(shape, **cfg)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    o = Oracle(p, g)
    o.set_primitive(*st)
    for m, (a, b) in enumerate(zip(s.compute_rhs(), o.compute_rhs())):
        assert relmax(s.interior(a), o.interior(b)) <= 1e-12, ("rhs", m)
    for _ in range(2):
        dt = s.timestep()
        assert dt == pytest.approx(o.timestep(), rel=1e-13)
        s.rk3_step(dt)
        o.step(dt)
    for m, (a, b) in enumerate(zip(s.conserved(), o.conserved())):
        assert relmax(s.interior(a), o.interior(b)) <= 1e-12, ("U", m)
    pm, gm = ns.make_params(shape, precision="mixed", **cfg)
    x = ns.Solver(pm, gm)
    x.upload_primitive(*st)
    o.set_primitive(*st)
    for m, (a, b) in enumerate(zip(x.compute_rhs(), o.compute_rhs())):
        assert relmax(x.interior(a), o.interior(b)) < 5e-5, ("mixed rhs", m)
