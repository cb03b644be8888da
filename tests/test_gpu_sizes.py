"""Grid sizes at the edges of what the kernels tile: the smallest legal box
(every axis = ng), odd extents that leave partial waves, blocks and column
segments in every kernel, and one long thin axis.  Each runs the FP64 path
against the oracle (RHS and RK3 steps), the mixed path against the oracle's
RHS, and, where the adjoint applies (every axis >= 8), its dot-product
identity."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ns3d_cpp_amd as ns  # noqa: E402
from oracle import Oracle  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4), (13, 7, 5), (5, 9, 31), (70, 6, 4), (67, 17, 9)]
WALLS = ["Wall", "Wall", "Outflow", "Outflow", "Symmetry", "Symmetry"]


def field(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    X, Y, Z = (2 * np.pi * np.arange(n) / n for n in (nx, ny, nz))
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    out = []
    for m, (base, amp) in enumerate(((1.0, 0.08), (0.0, 0.3), (0.0, 0.3), (0.0, 0.3), (20.0, 1.0))):
        f = np.sin(X + 2 * Y + rng.uniform(0, 6)) + np.cos(Y - Z + rng.uniform(0, 6)) + np.sin(Z + X)
        out.append(base + amp * (f + 0.2 * rng.standard_normal(X.shape)))
    return out


def params(shape, bc=None, **kw):
    cfg = dict(Re=100.0, Ma=0.17, Pr=0.72, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", cfl=0.4)
    cfg.update(kw)
    if bc:
        cfg["bc"] = bc
    return ns.make_params(shape, **cfg)


def relmax(a, b):
    return np.abs(a - b).max() / (np.abs(b).max() + 1e-300)


@pytest.mark.parametrize("bc", [None, WALLS], ids=["periodic", "walls"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fp64_rhs_and_steps(gpu, shape, bc):
    st = field(shape, 1)
    p, g = params(shape, bc)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    o = Oracle(p, g)
    o.set_primitive(*st)
    for m, (a, b) in enumerate(zip(s.compute_rhs(), o.compute_rhs())):
        assert relmax(s.interior(a), o.interior(b)) <= 1e-13, (shape, m)
    for _ in range(3):
        dt = s.timestep()
        s.rk3_step(dt)
        o.step(dt)
    for m, (a, b) in enumerate(zip(s.conserved(), o.conserved())):
        assert relmax(s.interior(a), o.interior(b)) <= 1e-13, (shape, m)


@pytest.mark.parametrize("bc", [None, WALLS], ids=["periodic", "walls"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_mixed_rhs(gpu, shape, bc):
    st = field(shape, 2)
    p, g = params(shape, bc, precision="mixed")
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    po, go = params(shape, bc)
    o = Oracle(po, go)
    o.set_primitive(*st)
    for m, (a, b) in enumerate(zip(s.compute_rhs(), o.compute_rhs())):
        assert relmax(s.interior(a), o.interior(b)) < 2e-5, (shape, m)


@pytest.mark.parametrize("shape", [(8, 8, 8), (13, 9, 11), (67, 17, 9)], ids=lambda s: "x".join(map(str, s)))
def test_adjoint_dot_product(gpu, shape):
    st = field(shape, 3)
    p, g = params(shape)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    rng = np.random.default_rng(4)
    nx, ny, nz = shape
    x = [s._full(rng.normal(size=(nz, ny, nx))) for _ in range(5)]
    y = [s._full(rng.normal(size=(nz, ny, nx))) for _ in range(5)]
    Rx = s.tlm_rhs(x)
    RTy, _ = s.adj_rhs(y, 0.0)

    def dot(a, b):
        return float(sum(np.sum(s.interior(u) * s.interior(v)) for u, v in zip(a, b)))
    lhs, rhs = dot(y, Rx), dot(RTy, x)
    assert abs(lhs - rhs) <= 1e-12 * np.sqrt(dot(y, y) * dot(Rx, Rx)), (lhs, rhs)
