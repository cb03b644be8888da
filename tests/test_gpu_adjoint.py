"""Tangent-linear and adjoint RHS, and the checkpointed adjoint of the RK3
time loop, on the GPU.

Checks that pin them independently of the code under test:
  * the TLM against central finite differences of the GPU RHS (O(h^2));
  * the adjoint against the TLM by the dot-product identity
    <y, R' x> = <R'^T y, x> (the TLM is forward-mode dual numbers, the
    adjoint hand-written reverse rules: they share no derivative code);
  * the gradient of J = sum_n dt * eps(U^n) against finite differences of J
    evaluated by the GPU forward path and by the CPU oracle's RK3 loop, the
    latter with eps computed in numpy from the oracle's gradients;
  * d/d(1/Re) against finite differences in Re;
  * checkpoint spacing changes nothing (the recomputation is bitwise);
  * the characteristic reconstruction (char_recon) and non-periodic sides
    (wall, symmetry, outflow, inflow): the same checks;
  * in mixed-precision contexts (mixed steps forward, the FP64 scheme's
    derivatives at each stage state's FP64 primitives): the RHS derivatives
    against the FP64 context's, the gradient against the FP64 gradient and
    FD of the FP64 J, and bitwise checkpoint-spacing invariance.
The split uses sw_eps > 0 where finite differences are taken, so the flux is
smooth; the dot-product test also covers the plain |lambda| splits.
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

BASE = dict(Re=100.0, Pr=0.72, sutherland_s=110.4 / 273.15, cfl=0.5)


def state(n, seed=3, Mt=0.3):
    """Smooth compressible state: HIT velocity with density/pressure ripples."""
    st, Ma = fields.hit(n, k0=2, Mt=Mt, seed=seed)
    rho, u, v, w, p = [np.array(a) for a in st]
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    X, Y, Z = X.transpose(2, 1, 0), Y.transpose(2, 1, 0), Z.transpose(2, 1, 0)
    rho = rho * (1.0 + 0.05 * np.sin(X + 2 * Y) * np.cos(Z))
    p = p * (1.0 + 0.05 * np.cos(2 * X - Z) * np.sin(Y))
    return (rho, u, v, w, p), Ma


def conserved_of(prims, gamma=1.4):
    rho, u, v, w, p = prims
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return [rho, rho * u, rho * v, rho * w, E]


def smooth_dir(n, seed, scale):
    """Random smooth direction (low Fourier modes), scaled per variable."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    out = []
    for m in range(5):
        a = np.zeros((n, n, n))
        for _ in range(4):
            k = rng.integers(-2, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            a += rng.normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + ph)
        out.append(scale[m] * a.transpose(2, 1, 0))
    return out


def interior_dot(s, a, b):
    return float(sum(np.sum(s.interior(x) * s.interior(y)) for x, y in zip(a, b)))


def solver(n, prims, Ma, **cfg):
    kw = dict(BASE)
    kw.update(cfg)
    p, g = ns.make_params(n, Ma=Ma, **kw)
    s = ns.Solver(p, g)
    s.upload_primitive(*prims)
    return s


DOT_CASES = [
    dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th"),
    dict(recon="WENO5", fvs_type="LaxFriedrichs", vis_scheme="C6th"),
    dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C4th", sw_eps=0.2, mu_law="power", mu_power=0.76),
    dict(recon="LINEAR", fvs_type="StagerWarming", vis_scheme="C8th", mu_law="constant"),
    dict(recon="MDCD", fvs_type="LaxFriedrichs", vis_scheme="C8th"),
    dict(recon="UPWIND_7TH", fvs_type="StagerWarming", vis_scheme="C8th"),
    dict(recon="UD7L", fvs_type="StagerWarming", vis_scheme="C6th", ud7l_alpha=0.3),
    dict(recon="MDCD_HYBRID", fvs_type="StagerWarming", vis_scheme="C8th"),
]


def front_state(n):
    """state() with density and pressure fronts about one cell wide, so that
    the MDCD-hybrid sensor selects its nonlinear branch near them."""
    prims, Ma = state(n)
    rho, u, v, w, p = prims
    x = np.arange(n) * 2 * np.pi / n
    f = 0.25 * np.tanh(4.0 * np.sin(x) * n / (2 * np.pi))[None, None, :]
    return (rho * (1.0 + f), u, v, w, p * (1.0 + 0.5 * f)), Ma


@pytest.mark.parametrize("cfg", DOT_CASES, ids=lambda c: f"{c['recon']}-{c['fvs_type']}-{c['vis_scheme']}")
def test_dot_product_identity(gpu, cfg):
    n = 16
    prims, Ma = state(n)
    s = solver(n, prims, Ma, **cfg)
    rng = np.random.default_rng(11)
    x = [s._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    y = [s._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    Rx = s.tlm_rhs(x)
    RTy, _ = s.adj_rhs(y, 0.0)
    lhs = interior_dot(s, y, Rx)
    rhs = interior_dot(s, RTy, x)
    scale = np.sqrt(interior_dot(s, y, y) * interior_dot(s, Rx, Rx))
    print(cfg, lhs, rhs, scale)
    assert abs(lhs - rhs) <= 1e-12 * scale
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("cfg", DOT_CASES[:2] + DOT_CASES[5:6] + DOT_CASES[7:8],
                         ids=["weno5-sw", "weno5-lf", "ud7", "mdcd-hybrid"])
def test_dot_product_marching_sweeps(gpu, cfg, monkeypatch):
    """The column-marching adjoint sweeps (used for lines too long for LDS)
    against the TLM, and against the whole-line sweeps."""
    n = 16
    prims, Ma = state(n)
    rng = np.random.default_rng(19)
    x = [rng.normal(size=(n, n, n)) for _ in range(5)]
    y = [rng.normal(size=(n, n, n)) for _ in range(5)]
    ref = solver(n, prims, Ma, **cfg)
    RTy_line, th_line = ref.adj_rhs(y, 0.5)
    monkeypatch.setenv("NS3D_ADJ_MARCH", "1")
    s = solver(n, prims, Ma, **cfg)
    RTy, th = s.adj_rhs(y, 0.5)
    Rx = s.tlm_rhs(x)
    RTy0, _ = s.adj_rhs(y, 0.0)
    lhs = interior_dot(s, [s._full(a) for a in y], Rx)
    rhs = interior_dot(s, RTy0, [s._full(a) for a in x])
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    for a, b in zip(RTy, RTy_line):
        assert np.allclose(s.interior(a), s.interior(b), rtol=1e-11, atol=1e-11 * np.abs(b).max())
    assert abs(th - th_line) <= 1e-12 * abs(th_line)


@pytest.mark.parametrize("vis", ["C8th", "C6th", "C4th"])
def test_z_marching_viscous_adjoint_is_bitwise(gpu, vis, monkeypatch):
    """The viscous adjoint kernels as z marches with LDS-staged planes (boxes
    the 64 x 4 blocks tile) against the cached-load forms: the adjoint RHS,
    the 1/Re term and a two-step reverse pass bit for bit, on a 64 x 16 x 40
    box (two z segments, the second shorter)."""
    nx, ny, nz = 64, 16, 40
    rng = np.random.default_rng(41)
    st, Ma = fields.hit(16, k0=2, Mt=0.3, seed=9)
    shape = (nz, ny, nx)
    prims = [np.resize(a, shape) * (1.0 + 0.01 * rng.standard_normal(shape)) for a in st]
    y = [rng.normal(size=shape) for _ in range(5)]
    out = []
    for zm in ("1", "0"):
        monkeypatch.setenv("NS3D_ADJ_ZMARCH", zm)
        kw = dict(BASE, recon="WENO5", fvs_type="StagerWarming", vis_scheme=vis)
        p, g = ns.make_params((nx, ny, nz), Ma=Ma, **kw)
        s = ns.Solver(p, g)
        s.upload_primitive(*prims)
        r, th = s.adj_rhs([s._full(a) for a in y], 0.4)
        J = s.adj_forward(2, 1e-3, 1)
        lam0, dth = s.adj_backward()
        out.append(([s.interior(a).copy() for a in r + lam0], (th, J, dth)))
        s.close()
    assert out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)


def _fd_rhs(n, prims, Ma, cfg, dU, h):
    s = solver(n, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    out = []
    for sg in (+1, -1):
        s.upload_conserved(*[a + sg * h * d for a, d in zip(U, dU)])
        out.append(s.compute_rhs())
    return [(a - b) / (2 * h) for a, b in zip(*out)]


@pytest.mark.parametrize("cfg", [
    dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3),
    dict(recon="UPWIND_7TH", fvs_type="StagerWarming", vis_scheme="C6th", sw_eps=0.3, mu_law="power",
         mu_power=0.7),
], ids=["weno5", "ud7"])
def test_tlm_against_finite_differences(gpu, cfg):
    n = 16
    prims, Ma = state(n)
    s = solver(n, prims, Ma, **cfg)
    dU = [s._full(a) for a in smooth_dir(n, 5, [0.05, 0.05, 0.05, 0.05, 1.0])]
    tl = s.tlm_rhs(dU)
    fd = _fd_rhs(n, prims, Ma, cfg, dU, 1e-5)
    num = np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(tl, fd)))
    den = np.sqrt(sum(np.sum(s.interior(b) ** 2) for b in fd))
    print("TLM vs FD rel", num / den)
    assert num / den < 1e-7


def test_tlm_mdcd_hybrid_fronts(gpu):
    """MDCD_HYBRID on a state with fronts (both sensor branches in use): the
    TLM against central differences, and the dot-product identity."""
    n = 16
    cfg = dict(recon="MDCD_HYBRID", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = front_state(n)
    s = solver(n, prims, Ma, **cfg)
    dU = [s._full(a) for a in smooth_dir(n, 7, [0.05, 0.05, 0.05, 0.05, 1.0])]
    tl = s.tlm_rhs(dU)
    fd = _fd_rhs(n, prims, Ma, cfg, dU, 1e-6)
    num = np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(tl, fd)))
    den = np.sqrt(sum(np.sum(s.interior(b) ** 2) for b in fd))
    print("hybrid TLM vs FD rel", num / den)
    assert num / den < 1e-6
    rng = np.random.default_rng(13)
    y = [s._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    RTy, _ = s.adj_rhs(y, 0.0)
    lhs, rhs = interior_dot(s, y, tl), interior_dot(s, RTy, dU)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def _sum_phi(s):
    g = s.grid
    return s.dissipation() * g.nx * g.ny * g.nz


def test_adjoint_objective_and_theta(gpu):
    """w_obj * d(sum Phi)/dU and d/d(1/Re) of lam.R + w_obj sum Phi."""
    n = 16
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = state(n)
    s = solver(n, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    rng = np.random.default_rng(2)
    lam = [s._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    w = 0.37
    g, th = s.adj_rhs(lam, w)
    # state direction
    dU = [s._full(a) for a in smooth_dir(n, 9, [0.05, 0.05, 0.05, 0.05, 1.0])]
    h = 1e-5
    val = []
    for sg in (+1, -1):
        s.upload_conserved(*[a + sg * h * d for a, d in zip(U, dU)])
        R = s.compute_rhs()
        val.append(interior_dot(s, lam, R) + w * _sum_phi(s))
    fd = (val[0] - val[1]) / (2 * h)
    ad = interior_dot(s, g, dU)
    print("state", fd, ad)
    assert abs(fd - ad) <= 1e-7 * abs(fd)
    # 1/Re direction
    theta = 1.0 / BASE["Re"]
    ht = 1e-6 * theta
    val = []
    for sg in (+1, -1):
        st_ = solver(n, prims, Ma, **dict(cfg, Re=1.0 / (theta + sg * ht)))
        R = st_.compute_rhs()
        val.append(interior_dot(st_, lam, R) + w * _sum_phi(st_))
    fdt = (val[0] - val[1]) / (2 * ht)
    print("theta", fdt, th)
    assert abs(fdt - th) <= 1e-7 * abs(fdt)


def _J_gpu(n, prims, Ma, cfg, U, dU, h, nsteps, dt):
    s = solver(n, prims, Ma, **cfg)
    s.upload_conserved(*[a + h * d for a, d in zip(U, dU)])
    return s.adj_forward(nsteps, dt, nsteps)


def test_gradient_of_time_integrated_dissipation(gpu):
    n, nsteps, dt = 16, 7, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = state(n)
    s = solver(n, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    s.upload_conserved(*U)
    J = s.adj_forward(nsteps, dt, 3)
    lam0, dth = s.adj_backward()
    dU = [s._full(a) for a in smooth_dir(n, 21, [0.05, 0.05, 0.05, 0.05, 1.0])]
    h = 1e-5
    Jp = _J_gpu(n, prims, Ma, cfg, U, dU, +h, nsteps, dt)
    Jm = _J_gpu(n, prims, Ma, cfg, U, dU, -h, nsteps, dt)
    J0 = _J_gpu(n, prims, Ma, cfg, U, dU, 0.0, nsteps, dt)
    assert J0 == J
    fd = (Jp - Jm) / (2 * h)
    ad = interior_dot(s, lam0, dU)
    print("J", J, "fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-6 * abs(fd)
    # 1/Re
    theta = 1.0 / BASE["Re"]
    ht = 1e-6 * theta
    Jt = []
    for sg in (+1, -1):
        st_ = solver(n, prims, Ma, **dict(cfg, Re=1.0 / (theta + sg * ht)))
        st_.upload_conserved(*U)
        Jt.append(st_.adj_forward(nsteps, dt, nsteps))
    fdt = (Jt[0] - Jt[1]) / (2 * ht)
    print("dJ/dtheta fd", fdt, "adjoint", dth)
    assert abs(fdt - dth) <= 1e-6 * abs(fdt)


def test_terminal_adjoint_and_checkpoint_spacing(gpu):
    """lam0 = dJ/dU0 + (dU^N/dU^0)^T lamN, the same for any checkpoint spacing."""
    n, nsteps, dt = 16, 5, 0.01
    cfg = dict(recon="WENO5", fvs_type="LaxFriedrichs", vis_scheme="C6th")
    prims, Ma = state(n, seed=8)
    s = solver(n, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    rng = np.random.default_rng(4)
    lamN = [s._full(rng.normal(size=(n, n, n)) * 1e-3) for _ in range(5)]
    res = []
    for ck in (1, 2, 5):
        s.upload_conserved(*U)
        s.adj_forward(nsteps, dt, ck)
        res.append(s.adj_backward(lamN))
    for lam0, th in res[1:]:
        assert th == res[0][1]
        for a, b in zip(lam0, res[0][0]):
            assert np.array_equal(s.interior(a), s.interior(b))
    # the context is left at U^N
    UN = s.conserved()
    s2 = solver(n, prims, Ma, **cfg)
    s2.upload_conserved(*U)
    s2.advance(nsteps, dt)
    for a, b in zip(UN, s2.conserved()):
        assert np.array_equal(s.interior(a), s.interior(b))
    # against FD of J + <lamN, U^N> (LF: |u_n| kinks are far from the
    # perturbation size used here unless u_n ~ 0 within 1e-9 somewhere)
    dU = [s._full(a) for a in smooth_dir(n, 17, [0.05, 0.05, 0.05, 0.05, 1.0])]
    h = 1e-6
    val = []
    for sg in (+1, -1):
        s3 = solver(n, prims, Ma, **cfg)
        s3.upload_conserved(*[a + sg * h * d for a, d in zip(U, dU)])
        J = s3.adj_forward(nsteps, dt, nsteps)
        val.append(J + interior_dot(s3, lamN, s3.conserved()))
    fd = (val[0] - val[1]) / (2 * h)
    ad = interior_dot(s, res[0][0], dU)
    print("fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-5 * abs(fd)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_overlapped_reverse_pass_is_bitwise(gpu, prec, monkeypatch):
    """The reverse pass that recomputes the next checkpoint segment on a
    second stream while it reverses the current one gives the one-stream
    result bit for bit, also when called again (which recomputes the last
    segment as well), and leaves the same final state."""
    n, nsteps, dt = 16, 7, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = state(n, seed=5)
    res = []
    for ov in ("1", "0"):
        monkeypatch.setenv("NS3D_ADJ_OVERLAP", ov)
        s = solver(n, prims, Ma, precision=prec, **cfg)
        U = [s._full(a) for a in conserved_of(prims)]
        s.upload_conserved(*U)
        J = s.adj_forward(nsteps, dt, 2)
        first = s.adj_backward()
        again = s.adj_backward()
        res.append((J, first, again, s.conserved(), s.adj_memory(), s))
    (J1, f1, a1, U1, m1, s1), (J0, f0, a0, U0, m0, _) = res
    assert J1 == J0
    for lam, th in (f1, a1, a0):
        assert th == f0[1]
        for x, y in zip(lam, f0[0]):
            assert np.array_equal(s1.interior(x), s1.interior(y))
    for x, y in zip(U1, U0):
        assert np.array_equal(s1.interior(x), s1.interior(y))
    assert m1 > m0   # the second snapshot buffer


def _phi_sum_oracle(o, gamma_unused=None):
    """sum over interior cells of Phi from the oracle's gradients and mu(T)."""
    G = [o.interior(a) for a in o.compute_gradients()]
    T = o.interior(o.primitive()["T"])
    p = o.params
    S = p.sutherland_s
    mu = p.mu * T ** 1.5 * (1.0 + S) / (T + S)
    ux, uy, uz, vx, vy, vz, wx, wy, wz = G[:9]
    div = ux + vy + wz
    phi = mu * (2 * (ux * ux + vy * vy + wz * wz) - (2.0 / 3.0) * div * div
                + (uy + vx) ** 2 + (uz + wx) ** 2 + (vz + wy) ** 2)
    return float(np.sum(phi))


def test_gradient_against_oracle_finite_differences(gpu):
    """The adjoint gradient of J against FD of J computed by the CPU oracle."""
    n, nsteps, dt = 12, 4, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = state(n, seed=12)
    s = solver(n, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    s.upload_conserved(*U)
    J = s.adj_forward(nsteps, dt, 2)
    lam0, _ = s.adj_backward()
    dU = [s._full(a) for a in smooth_dir(n, 33, [0.05, 0.05, 0.05, 0.05, 1.0])]
    kw = dict(BASE)
    kw.update(cfg)
    p, g = ns.make_params(n, Ma=Ma, **kw)

    def J_oracle(h):
        o = Oracle(p, g)
        o.set_conserved(*[a + h * d for a, d in zip(U, dU)])
        tot = 0.0
        for _ in range(nsteps):
            tot += dt * _phi_sum_oracle(o) / n ** 3
            st, _ = o.step(dt)
            assert st == 0
        return tot

    Jo = J_oracle(0.0)
    assert abs(Jo - J) <= 1e-12 * abs(J)
    h = 1e-5
    fd = (J_oracle(h) - J_oracle(-h)) / (2 * h)
    ad = interior_dot(s, lam0, dU)
    print("oracle fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-6 * abs(fd)


# ---- mixed-precision contexts: the mixed steps forward and recomputed, the
# FP64 scheme's derivatives at the FP64 primitives of each stage state

MIXED_CASES = [DOT_CASES[0], DOT_CASES[1], DOT_CASES[5]]


def _rel_close(s, a, b, tol):
    for x, y in zip(a, b):
        ref = np.max(np.abs(s.interior(x)))
        assert np.max(np.abs(s.interior(x) - s.interior(y))) <= tol * ref


@pytest.mark.parametrize("cfg", MIXED_CASES, ids=["weno5-sw", "weno5-lf", "ud7"])
def test_mixed_rhs_derivatives_match_fp64(gpu, cfg):
    """A mixed context holds U = P2C(W) in FP64; its TLM / adjoint RHS,
    d/d(1/Re) and eps are the FP64 scheme's at C2P(U), so they equal the
    FP64 context's to rounding, and the pair passes the dot-product test."""
    n = 16
    prims, Ma = state(n)
    s64 = solver(n, prims, Ma, **cfg)
    sm = solver(n, prims, Ma, precision="mixed", **cfg)
    rng = np.random.default_rng(11)
    x = [s64._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    y = [s64._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    Rx64, Rxm = s64.tlm_rhs(x), sm.tlm_rhs(x)
    _rel_close(s64, Rx64, Rxm, 1e-11)
    (g64, th64), (gm, thm) = s64.adj_rhs(y, 0.3), sm.adj_rhs(y, 0.3)
    _rel_close(s64, g64, gm, 1e-11)
    assert abs(th64 - thm) <= 1e-11 * abs(th64)
    e64, em = s64.dissipation(), sm.dissipation()
    assert abs(e64 - em) <= 1e-12 * abs(e64)
    RTy, _ = sm.adj_rhs(y, 0.0)
    lhs, rhs = interior_dot(sm, y, Rxm), interior_dot(sm, RTy, x)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_mixed_gradient_and_checkpoint_spacing(gpu):
    """The mixed forward/backward: J within the mixed step's own error of
    the FP64 J; the gradient against the FP64 adjoint and against central
    differences of the FP64 J; any checkpoint spacing gives the same bits;
    the context is left at the state a plain mixed advance reaches."""
    n, nsteps, dt = 16, 7, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = state(n)
    s64 = solver(n, prims, Ma, **cfg)
    U = [s64._full(a) for a in conserved_of(prims)]
    s64.upload_conserved(*U)
    J64 = s64.adj_forward(nsteps, dt, 3)
    lam64, th64 = s64.adj_backward()
    sm = solver(n, prims, Ma, precision="mixed", **cfg)
    res = []
    for ck in (3, 1, 7):
        sm.upload_conserved(*U)
        J = sm.adj_forward(nsteps, dt, ck)
        res.append((J,) + sm.adj_backward())
    Jm, lamm, thm = res[0]
    for J, lam, th in res[1:]:
        assert J == Jm and th == thm
        for a, b in zip(lam, lamm):
            assert np.array_equal(sm.interior(a), sm.interior(b))
    UN = sm.conserved()
    s2 = solver(n, prims, Ma, precision="mixed", **cfg)
    s2.upload_conserved(*U)
    s2.advance(nsteps, dt)
    for a, b in zip(UN, s2.conserved()):
        assert np.array_equal(sm.interior(a), sm.interior(b))
    print("J fp64", J64, "mixed", Jm, "theta", th64, thm)
    assert abs(Jm - J64) <= 1e-7 * abs(J64)
    assert abs(thm - th64) <= 1e-7 * abs(th64)
    num = np.sqrt(sum(np.sum((s64.interior(a) - s64.interior(b)) ** 2) for a, b in zip(lamm, lam64)))
    den = np.sqrt(sum(np.sum(s64.interior(a) ** 2) for a in lam64))
    print("gradient rel L2 mixed vs fp64", num / den)
    assert num <= 1e-6 * den
    dU = [s64._full(a) for a in smooth_dir(n, 21, [0.05, 0.05, 0.05, 0.05, 1.0])]
    h = 1e-5
    fd = (_J_gpu(n, prims, Ma, cfg, U, dU, +h, nsteps, dt) - _J_gpu(n, prims, Ma, cfg, U, dU, -h, nsteps, dt)) / (2 * h)
    ad = interior_dot(sm, lamm, dU)
    print("fd (fp64 J)", fd, "mixed adjoint", ad)
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def test_mixed_adjoint_memory_and_precision_switch(gpu):
    """Mixed checkpoints hold 8 doubles per cell (U + the FP32 shadow);
    switching precision drops the recorded pass."""
    n, nsteps, dt = 16, 4, 0.01
    prims, Ma = state(n)
    sm = solver(n, prims, Ma, precision="mixed", **DOT_CASES[0])
    base = sm.adj_memory()
    sm.adj_forward(nsteps, dt, 2)
    states = 2 + 3 * 2 + 1     # checkpoints, one segment's stage inputs, the final state
    assert sm.adj_memory() - base >= states * 8 * int(np.prod(sm.shape)) * 8
    sm.set_precision("fp64")
    with pytest.raises(ns.NS3DError):
        sm.adj_backward()
    sm.adj_forward(nsteps, dt, 2)
    sm.adj_backward()
    sm.set_precision("mixed")
    sm.adj_forward(nsteps, dt, 2)
    sm.adj_backward()


# ---- characteristic reconstruction (char_recon): the TLM carries dual
# numbers through the Roe state, its eigenvector matrices, the projections
# and the reconstructions; the adjoint reverses the projections and the
# reconstructions by hand and the Roe state's matrices by forward mode, so
# the dot-product test still compares independent derivative code

CHAR_CASES = [
    dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", char_recon=1),
    dict(recon="WENO5", fvs_type="LaxFriedrichs", vis_scheme="C6th", char_recon=1),
    dict(recon="UPWIND_7TH", fvs_type="StagerWarming", vis_scheme="C4th", sw_eps=0.2, char_recon=1),
    dict(recon="LINEAR", fvs_type="LaxFriedrichs", vis_scheme="C8th", mu_law="constant", char_recon=1),
    dict(recon="MDCD_HYBRID", fvs_type="StagerWarming", vis_scheme="C8th", char_recon=1),
]


@pytest.mark.parametrize("cfg", CHAR_CASES, ids=["weno5-sw", "weno5-lf", "ud7-swe", "linear-lf", "mdcd-hybrid"])
def test_char_dot_product_identity(gpu, cfg):
    n = 16
    prims, Ma = state(n)
    s = solver(n, prims, Ma, **cfg)
    rng = np.random.default_rng(23)
    x = [s._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    y = [s._full(rng.normal(size=(n, n, n))) for _ in range(5)]
    Rx = s.tlm_rhs(x)
    RTy, _ = s.adj_rhs(y, 0.0)
    lhs, rhs = interior_dot(s, y, Rx), interior_dot(s, RTy, x)
    scale = np.sqrt(interior_dot(s, y, y) * interior_dot(s, Rx, Rx))
    print(cfg, lhs, rhs, scale)
    assert abs(lhs - rhs) <= 1e-12 * scale
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_char_tlm_against_finite_differences(gpu, prec):
    """The characteristic TLM against central differences of the FP64
    characteristic RHS (whose faces the oracle pins); a mixed context gives
    the FP64 scheme's TLM at C2P(U)."""
    n = 16
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3, char_recon=1)
    prims, Ma = state(n)
    s = solver(n, prims, Ma, precision=prec, **cfg)
    dU = [s._full(a) for a in smooth_dir(n, 5, [0.05, 0.05, 0.05, 0.05, 1.0])]
    tl = s.tlm_rhs(dU)
    fd = _fd_rhs(n, prims, Ma, cfg, dU, 1e-5)
    num = np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(tl, fd)))
    den = np.sqrt(sum(np.sum(s.interior(b) ** 2) for b in fd))
    print(prec, "char TLM vs FD rel", num / den)
    assert num / den < 1e-7


def test_char_gradient_against_finite_differences(gpu):
    """The gradient of J through the characteristic scheme: against central
    differences of J (GPU forward) along a smooth direction and in 1/Re; the
    mixed context's gradient against the FP64 one."""
    n, nsteps, dt = 12, 4, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3, char_recon=1)
    prims, Ma = state(n, seed=12)
    s = solver(n, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    s.upload_conserved(*U)
    J = s.adj_forward(nsteps, dt, 2)
    lam0, dth = s.adj_backward()
    dU = [s._full(a) for a in smooth_dir(n, 33, [0.05, 0.05, 0.05, 0.05, 1.0])]
    h = 1e-5
    fd = (_J_gpu(n, prims, Ma, cfg, U, dU, +h, nsteps, dt) - _J_gpu(n, prims, Ma, cfg, U, dU, -h, nsteps, dt)) / (2 * h)
    ad = interior_dot(s, lam0, dU)
    print("char J", J, "fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-6 * abs(fd)
    theta = 1.0 / BASE["Re"]
    ht = 1e-6 * theta
    Jt = []
    for sg in (+1, -1):
        st_ = solver(n, prims, Ma, **dict(cfg, Re=1.0 / (theta + sg * ht)))
        st_.upload_conserved(*U)
        Jt.append(st_.adj_forward(nsteps, dt, nsteps))
    fdt = (Jt[0] - Jt[1]) / (2 * ht)
    print("char dJ/dtheta fd", fdt, "adjoint", dth)
    assert abs(fdt - dth) <= 1e-6 * abs(fdt)
    sm = solver(n, prims, Ma, precision="mixed", **cfg)
    sm.upload_conserved(*U)
    Jm = sm.adj_forward(nsteps, dt, 2)
    lamm, thm = sm.adj_backward()
    num = np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(lamm, lam0)))
    den = np.sqrt(sum(np.sum(s.interior(a) ** 2) for a in lam0))
    print("char mixed: J", Jm, "gradient rel L2 vs fp64", num / den, "theta", thm)
    assert abs(Jm - J) <= 1e-7 * abs(J)
    assert num <= 1e-6 * den
    assert abs(thm - dth) <= 1e-6 * abs(dth)


# ---- non-periodic sides: the linearised ghost fill (and its transpose in
# the adjoint), the first cell's left face, lambda = 0 outside the box, and
# the transposed closure rows of the viscous derivatives

BOX = (16, 12, 10)
LAYOUTS = {
    "channel": ["Periodic", "Periodic", "Wall", "Wall", "Outflow", "Outflow"],
    "slab": ["Outflow", "Outflow", "Periodic", "Periodic", "Symmetry", "Symmetry"],
    "box": ["Wall", "Wall", "Symmetry", "Symmetry", "Wall", "Outflow"],
    "inflow": ["Inflow", "Outflow", "Periodic", "Periodic", "Wall", "Wall"],
}


def box_state(seed, Ma=0.3):
    """Smooth state on BOX (z, y, x arrays): rho ~ 1, |u| ~ 0.3, p ~ 1/(gamma Ma^2)."""
    nx, ny, nz = BOX
    rng = np.random.default_rng(seed)
    X, Y, Z = (2 * np.pi * np.arange(n) / n for n in (nx, ny, nz))
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    p0 = 1.0 / (1.4 * Ma * Ma)
    out = []
    for base, amp in ((1.0, 0.05), (0.1, 0.3), (0.0, 0.3), (0.0, 0.3), (p0, 0.05 * p0)):
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(-2, 3, 3)
            f += np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        out.append(base + amp * f / 3.0)
    return out, Ma


def bsolver(bcs, prims, Ma, **cfg):
    kw = dict(BASE)
    kw.update(cfg)
    p, g = ns.make_params(BOX, Ma=Ma, bc=bcs, **kw)
    s = ns.Solver(p, g)
    s.upload_primitive(*prims)
    return s


def box_dir(s, seed, scale):
    rng = np.random.default_rng(seed)
    nx, ny, nz = BOX
    X, Y, Z = (2 * np.pi * np.arange(n) / n for n in (nx, ny, nz))
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    out = []
    for m in range(5):
        a = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(-2, 3, 3)
            a += rng.normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        out.append(s._full(scale[m] * a))
    return out


BC_SCHEMES = {
    "weno5-sw-c8": dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th"),
    "ud7-lf-c4": dict(recon="UPWIND_7TH", fvs_type="LaxFriedrichs", vis_scheme="C4th", mu_law="power",
                      mu_power=0.76),
    "char-weno5-c6": dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C6th", sw_eps=0.2, char_recon=1),
}


@pytest.mark.parametrize("scheme", list(BC_SCHEMES))
@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_bc_dot_product_identity(gpu, layout, scheme):
    prims, Ma = box_state(3)
    s = bsolver(LAYOUTS[layout], prims, Ma, **BC_SCHEMES[scheme])
    rng = np.random.default_rng(29)
    x = [s._full(rng.normal(size=BOX[::-1])) for _ in range(5)]
    y = [s._full(rng.normal(size=BOX[::-1])) for _ in range(5)]
    Rx = s.tlm_rhs(x)
    RTy, _ = s.adj_rhs(y, 0.0)
    lhs, rhs = interior_dot(s, y, Rx), interior_dot(s, RTy, x)
    scale = np.sqrt(interior_dot(s, y, y) * interior_dot(s, Rx, Rx))
    print(layout, scheme, lhs, rhs, scale)
    assert abs(lhs - rhs) <= 1e-12 * scale
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_bc_tlm_against_finite_differences(gpu, layout):
    """The TLM against 4th-order central differences of the GPU RHS, whose
    boundary handling the oracle pins.  A wall mirrors the normal velocity,
    so a state that does not vanish there has a jump across the wall, where
    WENO5's weights vary fast: 2nd-order differences converge as h^2 from
    1.9e-5 (h = 1e-4) to 1.8e-7 (1e-5), 4th-order ones as h^4 to the 9e-10
    rounding floor."""
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = box_state(5)
    s = bsolver(LAYOUTS[layout], prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    dU = box_dir(s, 7, [0.05, 0.05, 0.05, 0.05, 1.0])
    tl = s.tlm_rhs(dU)
    h = 1e-5
    R = {}
    for q in (-2, -1, 1, 2):
        s.upload_conserved(*[a + q * h * d for a, d in zip(U, dU)])
        R[q] = s.compute_rhs()
    fd = [(8.0 * (a - b) - (c - d)) / (12.0 * h) for a, b, c, d in zip(R[1], R[-1], R[2], R[-2])]
    num = np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(tl, fd)))
    den = np.sqrt(sum(np.sum(s.interior(b) ** 2) for b in fd))
    print(layout, "TLM vs FD rel", num / den)
    assert num / den < 1e-8


@pytest.mark.parametrize("layout,char", [("box", 0), ("channel", 0), ("inflow", 1)])
def test_bc_gradient_against_finite_differences(gpu, layout, char):
    """The gradient of J with non-periodic sides against central differences
    of J (GPU forward) along a smooth direction and in 1/Re, the same for
    any checkpoint spacing; a mixed context's gradient against the FP64 one."""
    nsteps, dt = 4, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3, char_recon=char)
    bcs = LAYOUTS[layout]
    prims, Ma = box_state(11)
    s = bsolver(bcs, prims, Ma, **cfg)
    U = [s._full(a) for a in conserved_of(prims)]
    res = []
    for ck in (2, 1):
        s.upload_conserved(*U)
        J = s.adj_forward(nsteps, dt, ck)
        res.append((J,) + s.adj_backward())
    J, lam0, dth = res[0]
    assert res[1][0] == J and res[1][2] == dth
    for a, b in zip(res[1][1], lam0):
        assert np.array_equal(s.interior(a), s.interior(b))
    dU = box_dir(s, 33, [0.05, 0.05, 0.05, 0.05, 1.0])

    def J_of(h, **over):
        t = bsolver(bcs, prims, Ma, **dict(cfg, **over))
        t.upload_conserved(*[a + h * d for a, d in zip(U, dU)])
        return t.adj_forward(nsteps, dt, nsteps)

    h = 1e-5
    fd = (8.0 * (J_of(h) - J_of(-h)) - (J_of(2 * h) - J_of(-2 * h))) / (12.0 * h)
    ad = interior_dot(s, lam0, dU)
    print(layout, "J", J, "fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-6 * abs(fd)
    theta = 1.0 / BASE["Re"]
    ht = 1e-6 * theta
    fdt = (J_of(0.0, Re=1.0 / (theta + ht)) - J_of(0.0, Re=1.0 / (theta - ht))) / (2 * ht)
    print(layout, "dJ/dtheta fd", fdt, "adjoint", dth)
    assert abs(fdt - dth) <= 1e-6 * abs(fdt)
    sm = bsolver(bcs, prims, Ma, precision="mixed", **cfg)
    sm.upload_conserved(*U)
    Jm = sm.adj_forward(nsteps, dt, 2)
    lamm, thm = sm.adj_backward()
    num = np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(lamm, lam0)))
    den = np.sqrt(sum(np.sum(s.interior(a) ** 2) for a in lam0))
    print(layout, "mixed: J", Jm, "gradient rel L2 vs fp64", num / den)
    assert abs(Jm - J) <= 1e-6 * abs(J)
    assert num <= 1e-5 * den


def test_unsupported_settings(gpu):
    """Every scheme a context accepts is differentiated (WCNS is refused at
    creation); an extent below 8 cells is not."""
    kw = dict(BASE)
    p, g = ns.make_params((16, 16, 6), Ma=0.3, **kw)
    s = ns.Solver(p, g)
    s.upload_primitive(*[np.full((6, 16, 16), v) for v in (1.0, 0.1, 0.0, 0.0, 8.0)])
    with pytest.raises(ns.NS3DError):
        s.tlm_rhs([np.zeros((6, 16, 16))] * 5)
    n = 16
    prims, Ma = state(n)
    s = solver(n, prims, Ma)
    with pytest.raises(ns.NS3DError):
        s.adj_backward()   # no forward pass recorded


@pytest.mark.parametrize("kind", ["smooth", "white"])
def test_gradient_fd_32_isotropic_turbulence(gpu, kind):
    """The adjoint workload's setting (HIT k0=4, Mt=0.3, Re=100, WENO5 +
    Steger-Warming, Sutherland) at 32^3 over 10 steps, along a seeded random
    direction: adjoint gradient against a fourth-order central difference of
    J.  The split carries sw_eps = 0.05 so that J is smooth at the step sizes
    used (with the plain |lambda| a few of the 10^7 eigenvalue evaluations
    cross zero inside the difference stencil)."""
    n, nsteps = 32, 10
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.05)
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=77)
    s = solver(n, st, Ma, **cfg)
    dt = s.timestep()
    U = [s._full(a) for a in conserved_of(st)]
    s.upload_conserved(*U)
    J = s.adj_forward(nsteps, dt, 4)
    lam0, _ = s.adj_backward()
    scale = [0.01, 0.01, 0.01, 0.01, 0.2]
    if kind == "smooth":
        dU = [s._full(a) for a in smooth_dir(n, 123, scale)]
    else:
        rng = np.random.default_rng(123)
        dU = [s._full(c * rng.normal(size=(n, n, n))) for c in scale]
    h = 1e-4
    val = {}
    for m in (-2, -1, 1, 2):
        s2 = solver(n, st, Ma, **cfg)
        s2.upload_conserved(*[a + m * h * d for a, d in zip(U, dU)])
        val[m] = s2.adj_forward(nsteps, dt, nsteps)
    fd = (8.0 * (val[1] - val[-1]) - (val[2] - val[-2])) / (12.0 * h)
    ad = interior_dot(s, lam0, dU)
    print(kind, "J", J, "fd", fd, "adjoint", ad, "rel", abs(fd - ad) / abs(fd))
    assert abs(fd - ad) <= 1e-6 * abs(fd)


class _DissipationJ:
    """torch.autograd.Function over the C-ABI forward/backward pair: the
    time-integrated dissipation J(U0, 1/Re).  Only the tests use torch here;
    the library itself takes and returns plain host arrays."""

    @staticmethod
    def build():
        import torch

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, U0, inv_re, solver_factory, nsteps, dt, ckpt):
                s = solver_factory(float(inv_re))
                arrs = [s._full(U0[m].detach().cpu().numpy()) for m in range(5)]
                s.upload_conserved(*arrs)
                J = s.adj_forward(nsteps, dt, ckpt)
                ctx.solver = s
                return torch.tensor(J, dtype=torch.float64)

            @staticmethod
            def backward(ctx, gJ):
                s = ctx.solver
                lam0, dth = s.adj_backward()
                g = torch.tensor(np.stack([s.interior(a) for a in lam0]), dtype=torch.float64)
                return gJ * g, gJ * torch.tensor(dth, dtype=torch.float64), None, None, None, None

        return Fn


def test_torch_autograd_wrapper(gpu):
    import torch
    n, nsteps, dt = 16, 4, 0.01
    cfg = dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.3)
    prims, Ma = state(n, seed=41)

    def factory(inv_re):
        return solver(n, prims, Ma, **dict(cfg, Re=1.0 / inv_re))

    Fn = _DissipationJ.build()
    U0 = torch.tensor(np.stack(conserved_of(prims)), dtype=torch.float64, requires_grad=True)
    th = torch.tensor(1.0 / BASE["Re"], dtype=torch.float64, requires_grad=True)
    J = Fn.apply(U0, th, factory, nsteps, dt, 2)
    (2.0 * J).backward()
    # the same gradient through the Solver API
    s = factory(1.0 / BASE["Re"])
    s.upload_conserved(*[s._full(a) for a in conserved_of(prims)])
    J2 = s.adj_forward(nsteps, dt, 2)
    lam0, dth = s.adj_backward()
    assert J.item() == J2
    assert np.array_equal(U0.grad.numpy(), 2.0 * np.stack([s.interior(a) for a in lam0]))
    assert th.grad.item() == 2.0 * dth
    # a directional derivative through the wrapper against central differences
    d = torch.tensor(np.stack(smooth_dir(n, 3, [0.05, 0.05, 0.05, 0.05, 1.0])), dtype=torch.float64)
    h = 1e-5
    with torch.no_grad():
        Jp = Fn.apply(U0 + h * d, th, factory, nsteps, dt, 2)
        Jm = Fn.apply(U0 - h * d, th, factory, nsteps, dt, 2)
    fd = (Jp - Jm).item() / (2 * h)
    ad = float(torch.sum(U0.grad * d)) / 2.0
    assert abs(fd - ad) <= 1e-6 * abs(fd)
