"""GPU parity: the HIP path through the C ABI against the CPU FP64 oracle.

Small grids run both sides on the same seeded inputs; the FP64 path keeps
the reference's operation order with no FMA contraction, so results are
expected to agree to the last bits (pow() in Sutherland's law may differ by
an ulp).  Tolerances are stated per test.  Full-size (256^3) cases check
size-independent properties: free-stream preservation, conservation and
mirror symmetry.
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

RECONS = ["WENO5", "MDCD", "MDCD_HYBRID", "UPWIND_7TH", "UD7L", "LINEAR"]
FVSS = ["StagerWarming", "LaxFriedrichs"]


def rough_field(n, seed=0, Ma=0.3):
    """Smooth low modes plus cell noise: WENO weights and the MDCD-hybrid
    sensor take both branches, velocities change sign."""
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    out = []
    for m, base in enumerate((1.0, 0.0, 0.0, 0.0, 1.0 / (1.4 * Ma * Ma))):
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(-2, 3, 3)
            f += np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        f += 0.3 * rng.standard_normal(X.shape)
        amp = 0.1 * base if m in (0, 4) else 0.5
        out.append(base + amp * f)
    return out


def pair(n, state, **cfg):
    kw = dict(Re=100.0, Ma=0.3, Pr=0.72)
    kw.update(cfg)
    p, g = ns.make_params(n, **kw)
    s = ns.Solver(p, g)
    s.upload_primitive(*state)
    o = Oracle(p, g)
    o.set_primitive(*state)
    return s, o


def relmax(a, b):
    return np.abs(a - b).max() / (np.abs(b).max() + 1e-300)


def rel_l2(a, b):
    return np.sqrt(np.sum((a - b) ** 2)) / (np.sqrt(np.sum(b * b)) + 1e-300)


def test_upload_matches_oracle_state(gpu):
    n = 12
    st = rough_field(n, 1)
    s, o = pair(n, st, recon="WENO5", vis_scheme="C8th")
    gp, op = s.primitive(), o.primitive()
    ng = s.ng
    for k in ("rho", "u", "v", "w", "p", "T"):
        a, b = gp[k], op[k]
        # interior and face ghosts (edges/corners are never read)
        assert np.array_equal(s.interior(a), o.interior(b)), k
        assert np.array_equal(a[ng:-ng, ng:-ng, :ng], b[ng:-ng, ng:-ng, :ng]), k
        assert np.array_equal(a[:ng, ng:-ng, ng:-ng], b[:ng, ng:-ng, ng:-ng]), k
    for a, b in zip(s.conserved(), o.conserved()):
        assert np.array_equal(s.interior(a), o.interior(b))


def test_upload_conserved_round_trip(gpu):
    n = 12
    st = rough_field(n, 2)
    p, g = ns.make_params(n, Re=100.0, Ma=0.3)
    o = Oracle(p, g)
    o.set_primitive(*st)
    U = [o.interior(a) for a in o.conserved()]
    s = ns.Solver(p, g)
    s.upload_conserved(*U)
    o2 = Oracle(p, g)
    o2.set_conserved(*U)
    gp, op = s.primitive(), o2.primitive()
    for k in ("rho", "u", "v", "w", "p", "T"):
        assert np.array_equal(s.interior(gp[k]), o2.interior(op[k])), k


@pytest.mark.parametrize("fvs", FVSS)
@pytest.mark.parametrize("recon", RECONS)
def test_face_fluxes(gpu, recon, fvs):
    """du_inviscous face arrays (InviscousFlux_new.cpp:4-163); tol 1e-14 rel."""
    n = 12
    s, o = pair(n, rough_field(n, 3), recon=recon, fvs_type=fvs, vis_scheme="C8th")
    s.du_inviscous()
    o.du_inviscous()
    for d in range(3):
        for m, (a, b) in enumerate(zip(s.face_flux(d), o.face_flux(d))):
            assert relmax(a, b) <= 1e-14, (d, m, relmax(a, b))


@pytest.mark.parametrize("fvs", FVSS)
@pytest.mark.parametrize("recon", ["WENO5", "UPWIND_7TH"])
def test_face_fluxes_characteristic(gpu, recon, fvs):
    """du_inviscous_character (InviscousFlux_new.cpp:165-437); tol 1e-13 rel."""
    n = 12
    st = rough_field(n, 4)
    s, o = pair(n, st, recon=recon, fvs_type=fvs, vis_scheme="C8th", char_recon=1)
    s.du_inviscous()
    o.du_inviscous()
    for d in range(3):
        for m, (a, b) in enumerate(zip(s.face_flux(d), o.face_flux(d))):
            assert relmax(a, b) <= 1e-13, (d, m, relmax(a, b))


@pytest.mark.parametrize("vis", ["C8th", "C6th", "C4th"])
@pytest.mark.parametrize("law", [dict(mu_law="sutherland"), dict(mu_law="constant"),
                                 dict(mu_law="power", mu_power=0.76),
                                 dict(mu_law="sutherland", sutherland_s=110.4 / 273.15)])
def test_gradients_and_viscous_flux(gpu, vis, law):
    """compute_gradients / compute_viscous_flux (ViscousFlux.cpp:240-339)."""
    n = 12
    s, o = pair(n, rough_field(n, 5), recon="WENO5", vis_scheme=vis, **law)
    gg = s.compute_gradients()
    og = o.compute_gradients()
    for m in range(12):
        assert np.array_equal(s.interior(gg[m]), o.interior(og[m])), m
    gv = s.compute_viscous_flux()
    ov = o.compute_viscous_flux()
    for m in range(12):
        # pow() of the device library may differ from glibc's by an ulp
        assert relmax(s.interior(gv[m]), o.interior(ov[m])) <= 1e-14, m


CFGS = [dict(recon=r, fvs_type=f, vis_scheme=v) for r, f, v in [
    ("WENO5", "StagerWarming", "C8th"), ("WENO5", "LaxFriedrichs", "C6th"), ("MDCD", "LaxFriedrichs", "C8th"),
    ("MDCD_HYBRID", "StagerWarming", "C4th"), ("UPWIND_7TH", "StagerWarming", "C8th"),
    ("UD7L", "LaxFriedrichs", "C8th"), ("LINEAR", "StagerWarming", "C4th")]]
CFGS += [dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", sw_eps=0.05),
         dict(recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th", char_recon=1),
         dict(recon="MDCD", fvs_type="LaxFriedrichs", vis_scheme="C8th", char_recon=1)]


@pytest.mark.parametrize("cfg", CFGS, ids=lambda c: "-".join(str(v) for v in c.values()))
def test_rhs(gpu, cfg):
    """compute_rhs (range_kutta.cpp:11-99) through the fused sweeps; 1e-13 rel."""
    n = 16
    s, o = pair(n, rough_field(n, 6), **cfg)
    gr = s.compute_rhs()
    orr = o.compute_rhs()
    for m in range(5):
        assert relmax(s.interior(gr[m]), o.interior(orr[m])) <= 1e-13, (m, relmax(s.interior(gr[m]), o.interior(orr[m])))


def test_rhs_nonuniform_sizes(gpu):
    """nx != ny != nz with several sweep segments per column."""
    p, g = ns.make_params((40, 24, 20), Re=100.0, Ma=0.3, recon="WENO5", vis_scheme="C8th")
    rng = np.random.default_rng(7)
    st = [1.0 + 0.1 * rng.standard_normal((20, 24, 40)) for _ in range(5)]
    st[4] = st[4] + 8.0
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    o = Oracle(p, g)
    o.set_primitive(*st)
    for a, b in zip(s.compute_rhs(), o.compute_rhs()):
        assert relmax(s.interior(a), o.interior(b)) <= 1e-13


def _u_rel_l2(s, o):
    return [rel_l2(s.interior(a), o.interior(b)) for a, b in zip(s.conserved(), o.conserved())]


def test_tgv_steps_fixed_dt(gpu):
    """TGV (64^3 config at 32^3): 10 RK3 steps, dt = 1e-3; rel-L2(U) <= 1e-12."""
    n = 32
    s, o = pair(n, fields.tgv(n), Re=1600.0, Ma=0.1, Pr=0.72, recon="WENO5",
                fvs_type="StagerWarming", vis_scheme="C8th")
    for _ in range(10):
        s.rk3_step(1e-3)
        o.step(1e-3)
        assert s.total_energy() == pytest.approx(o.total_energy(), rel=1e-13)
    assert max(_u_rel_l2(s, o)) <= 1e-12, _u_rel_l2(s, o)


def test_hit_steps_cfl(gpu):
    """HIT (MDCD + LF, CFL 0.5, Sutherland S/T_ref=110.4/273.15) at 32^3:
    8 steps with the CFL dt of each side; rel-L2(U) <= 1e-12."""
    n = 32
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=11)
    s, o = pair(n, st, Re=100.0, Ma=Ma, Pr=0.72, recon="MDCD", fvs_type="LaxFriedrichs",
                vis_scheme="C8th", cfl=0.5, sutherland_s=110.4 / 273.15)
    for _ in range(8):
        dt = s.timestep()
        assert dt == o.timestep()
        s.rk3_step(dt)
        o.step(dt)
    assert max(_u_rel_l2(s, o)) <= 1e-12, _u_rel_l2(s, o)


def test_advance_matches_stepwise(gpu):
    n = 16
    st, Ma = fields.hit(n, k0=2, Mt=0.3, seed=3)
    p, g = ns.make_params(n, Re=100.0, Ma=Ma, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th")
    a = ns.Solver(p, g)
    b = ns.Solver(p, g)
    a.upload_primitive(*st)
    b.upload_primitive(*st)
    a.advance(5, 2e-3)
    for _ in range(5):
        b.rk3_step(2e-3)
    for x, y in zip(a.conserved(), b.conserved()):
        assert np.array_equal(x, y)


def test_reference_kinetic_energy_history_gpu(gpu):
    """The reference's printed Ek_avg for the 64^3 TGV, on the GPU path."""
    import json
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "tgv64_ek_reference.json")))
    c = ref["case"]
    p, g = ns.make_params(c["n"], Re=c["Re"], Ma=c["Ma"], Pr=c["Pr"], recon=c["recon"],
                          fvs_type=c["fvs_type"], vis_scheme=c["vis_scheme"])
    s = ns.Solver(p, g)
    s.upload_primitive(*fields.tgv(c["n"], Ma=c["Ma"]))
    for want in ref["ek"]:
        s.rk3_step(c["dt"])
        assert round(s.total_energy(), ref["printed_digits"]) == pytest.approx(want, abs=1e-9)


def test_tgv64_many_steps(gpu):
    """The 64^3 TGV workload (WENO5 + SW + C8, dt 1e-3) for 40 steps against
    the oracle: rel-L2(U) <= 1e-12 and the KE history to 1e-13."""
    n = 64
    s, o = pair(n, fields.tgv(n), Re=1600.0, Ma=0.1, Pr=0.72, recon="WENO5",
                fvs_type="StagerWarming", vis_scheme="C8th")
    for it in range(40):
        s.rk3_step(1e-3)
        o.step(1e-3)
        if it % 10 == 9:
            assert s.total_energy() == pytest.approx(o.total_energy(), rel=1e-13)
    assert max(_u_rel_l2(s, o)) <= 1e-12, _u_rel_l2(s, o)


def test_nonphysical_detection(gpu):
    """A cell with E below its kinetic energy (p < 0) makes the step report
    NS3D_ERR_NONPHYSICAL (range_kutta.cpp:150-161) with the first bad cell
    in (k, j, i) order."""
    n = 12
    st = [np.ones((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)),
          np.full((n, n, n), 10.0)]
    p, g = ns.make_params(n, Re=100.0, Ma=0.3, recon="WENO5", vis_scheme="C8th")
    o = Oracle(p, g)
    o.set_primitive(*st)
    U = [o.interior(a).copy() for a in o.conserved()]
    U[1][5, 6, 7] = 3.0
    U[4][5, 6, 7] = 1.0          # E = 1 < 0.5 * 3^2: p < 0 there
    s = ns.Solver(p, g)
    s.upload_conserved(*U)
    with pytest.raises(ns.NS3DError) as e:
        s.rk3_step(1e-3)
    assert e.value.status == ns._lib.NS3D_ERR_NONPHYSICAL
    i, j, k = s.nonphysical_cell()
    prim = s.primitive()
    ng = s.ng
    bad = ~(prim["rho"] > 0) | ~(prim["p"] > 0) | ~(prim["T"] > 0)
    inter = np.zeros_like(bad)
    inter[ng:-ng, ng:-ng, ng:-ng] = True
    kk, jj, ii = np.argwhere(bad & inter)[0]
    assert (i, j, k) == (ii, jj, kk)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("comp", [0, 4])
def test_nan_is_reported_nonphysical(gpu, prec, comp):
    """A NaN in one cell's density or energy: the step reports
    NS3D_ERR_NONPHYSICAL at the first non-physical cell in (k, j, i) order
    (NaN included).  The reference's scan (range_kutta.cpp:150-161) tests
    rho <= 0 and p <= 0, which are false for NaN, so it would go on; the
    oracle restates that and reports nothing."""
    n = 12
    st = rough_field(n, 9)
    p, g = ns.make_params(n, Re=100.0, Ma=0.3, recon="WENO5", vis_scheme="C8th", precision=prec)
    po, go = ns.make_params(n, Re=100.0, Ma=0.3, recon="WENO5", vis_scheme="C8th")
    o = Oracle(po, go)
    o.set_primitive(*st)
    U = [o.interior(a).copy() for a in o.conserved()]
    U[comp][6, 3, 9] = np.nan
    o.set_conserved(*U)
    status, _ = o.step(1e-4)
    assert status == 0            # the reference's check misses NaN
    s = ns.Solver(p, g)
    s.upload_conserved(*U)
    with pytest.raises(ns.NS3DError) as e:
        s.rk3_step(1e-4)
    assert e.value.status == ns._lib.NS3D_ERR_NONPHYSICAL
    i, j, k = s.nonphysical_cell()
    prim = s.primitive()
    ng = s.ng
    bad = ~(prim["rho"] > 0) | ~(prim["p"] > 0) | ~(prim["T"] > 0)
    inter = np.zeros_like(bad)
    inter[ng:-ng, ng:-ng, ng:-ng] = True
    kk, jj, ii = np.argwhere(bad & inter)[0]
    assert (i, j, k) == (ii, jj, kk)
    s.close()


def test_healthy_step_reports_ok(gpu):
    n = 12
    s, _ = pair(n, rough_field(n, 8), recon="WENO5", vis_scheme="C8th")
    s.rk3_step(1e-4)
    assert s.nonphysical_cell() == (-1, -1, -1)


# ------------------------------------------------- full-size properties
N_FULL = 256


def test_full_size_free_stream(gpu):
    """256^3 uniform flow: the mass RHS is exactly 0 and a step leaves the
    state unchanged to rounding (size-independent)."""
    n = N_FULL
    st = [np.full((n, n, n), v) for v in (1.1, 0.3, -0.2, 0.25, 20.0)]
    p, g = ns.make_params(n, Re=200.0, Ma=0.17, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th")
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    r = s.compute_rhs()
    assert np.all(s.interior(r[0]) == 0.0)
    for m in range(1, 5):
        assert np.abs(s.interior(r[m])).max() < 1e-30


def test_full_size_conservation_and_mirror(gpu):
    """256^3 isotropic turbulence: sum of each RHS component vanishes
    (telescoping fluxes) and the RHS of the x-mirrored field is the mirrored
    RHS with u -> -u."""
    n = N_FULL
    (rho, u, v, w, p), Ma = fields.hit(n, k0=4, Mt=0.3, seed=42)
    pr, g = ns.make_params(n, Re=200.0, Ma=Ma, Pr=0.72, recon="WENO5", fvs_type="StagerWarming",
                           vis_scheme="C8th", sutherland_s=110.4 / 273.15)
    s = ns.Solver(pr, g)
    s.upload_primitive(rho, u, v, w, p)
    r = [s.interior(a).copy() for a in s.compute_rhs()]
    for m in range(5):
        assert abs(r[m].sum()) <= 1e-12 * np.abs(r[m]).sum(), m

    def mx(a):
        return np.roll(a[:, :, ::-1], 1, axis=2)
    s.upload_primitive(mx(rho), -mx(u), mx(v), mx(w), mx(p))
    rm = [s.interior(a) for a in s.compute_rhs()]
    sign = [1, -1, 1, 1, 1]
    for m in range(5):
        assert np.abs(rm[m] - sign[m] * mx(r[m])).max() <= 1e-11 * np.abs(r[m]).max(), m


@pytest.mark.parametrize("vis", ["C8th", "C6th", "C4th"])
def test_compute_vis_flux_matches_oracle(gpu, vis):
    """compute_vis_flux (ViscousFlux.cpp:341-386) on rhs = 0 after
    compute_gradients + compute_viscous_flux + the periodic flux exchange."""
    import oracle as orc
    n = 12
    st = rough_field(n, 9)
    s, o = pair(n, st, recon="WENO5", vis_scheme=vis)
    s.compute_gradients()
    s.compute_viscous_flux()
    r = s.compute_vis_flux()
    for m in range(5):
        o.arr(orc.RHS0 + m)[...] = 0.0
    o.compute_gradients()
    o.compute_viscous_flux()
    o.L.orc_exchange_vflux(o.h)
    o.L.orc_compute_vis_flux(o.h)
    for m in range(5):
        a, b = s.interior(r[m]), o.interior(o.arr(orc.RHS0 + m))
        assert np.array_equal(a, b) or relmax(a, b) <= 1e-15, (m, relmax(a, b))


def test_set_precision_round_trip(gpu):
    """Switching an FP64 context to mixed gives the state a mixed context
    builds from the same upload (bitwise), the steps agree bitwise, and
    switching back gives that context's primitives."""
    n = 16
    st, Ma = fields.hit(n, k0=2, Mt=0.3, seed=6)
    cfg = dict(Re=100.0, Ma=Ma, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th")
    p, g = ns.make_params(n, **cfg)
    a = ns.Solver(p, g)
    a.upload_primitive(*st)
    a.set_precision("mixed")
    pm, gm = ns.make_params(n, precision="mixed", **cfg)
    b = ns.Solver(pm, gm)
    b.upload_primitive(*st)
    for x, y in zip(a.conserved(), b.conserved()):
        assert np.array_equal(a.interior(x), b.interior(y))
    for _ in range(3):
        a.rk3_step(2e-3)
        b.rk3_step(2e-3)
    for x, y in zip(a.conserved(), b.conserved()):
        assert np.array_equal(a.interior(x), b.interior(y))
    a.set_precision("fp64")
    pa, pb = a.primitive(), b.primitive()
    for k in ("rho", "u", "v", "w", "p", "T"):
        assert np.array_equal(a.interior(pa[k]), b.interior(pb[k])), k
    a.rk3_step(2e-3)        # the FP64 path runs on the converted state


def rough_box(n, seed):
    """rough_field on an (nx, ny, nz) box (an int: the cube)."""
    if isinstance(n, int):
        return rough_field(n, seed)
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    Z, Y, X = np.meshgrid(*(2 * np.pi * np.arange(k) / k for k in (nz, ny, nx)), indexing="ij")
    out = []
    for m, base in enumerate((1.0, 0.0, 0.0, 0.0, 1.0 / (1.4 * 0.09))):
        f = np.zeros_like(X)
        for _ in range(3):
            k = rng.integers(-2, 3, 3)
            f += np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.3))
        f += 0.3 * rng.standard_normal(X.shape)
        out.append(base + (0.1 * base if m in (0, 4) else 0.5) * f)
    return out


def blob_field(n, seed, scale):
    """rough_box with rho and p multiplied by `scale` inside a ball (same
    temperature and sound speed): split fluxes there leave the range in
    which the sweeps take WENO5's quotients by refined reciprocals, so the
    columns and rows through the ball switch to the literal quotients."""
    st = rough_box(n, seed)
    nz, ny, nx = (n, n, n) if isinstance(n, int) else n[::-1]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ball = (x - 0.3 * nx) ** 2 + (y - 0.6 * ny) ** 2 + (z - 0.5 * nz) ** 2 < (0.2 * min(nx, ny, nz)) ** 2
    for m in (0, 4):
        st[m] = np.where(ball, st[m] * scale, st[m])
    return st


RCP_CASES = [
    ("in-range", 32, None, dict(fvs_type="StagerWarming")),
    ("in-range", 32, None, dict(fvs_type="StagerWarming", sw_eps=0.1)),
    ("in-range", 32, None, dict(fvs_type="LaxFriedrichs")),
    ("in-range", (40, 24, 20), None, dict(fvs_type="StagerWarming")),
    ("in-range-walls", 32, None, dict(fvs_type="StagerWarming",
                                      bc=["Wall", "Wall", "Periodic", "Periodic", "Periodic", "Periodic"])),
    ("large-ball", 32, 1e62, dict(fvs_type="StagerWarming")),
    ("large-ball", (40, 24, 20), 1e62, dict(fvs_type="LaxFriedrichs")),
    ("tiny-ball", 32, 1e-285, dict(fvs_type="StagerWarming")),
    ("large-ball-walls", 32, 1e62, dict(fvs_type="StagerWarming",
                                        bc=["Wall", "Wall", "Periodic", "Periodic", "Periodic", "Periodic"])),
]


@pytest.mark.parametrize("case,n,scale,cfg", RCP_CASES,
                         ids=[f"{c[0]}-{i}" for i, c in enumerate(RCP_CASES)])
def test_weno5_sweeps_rhs_bitwise(gpu, case, n, scale, cfg):
    """The FP64 sweeps take WENO5's quotients (reconstruction.cpp:106-116)
    with one refined reciprocal per denominator where every split flux of a
    wave's stencils is zero or within [2^-900, 2^200], and by the literal
    quotients elsewhere.  Both must give the oracle's right-hand side bit
    for bit (constant mu, so no pow() in the comparison): in range, and
    with a ball of 1e62 or 1e-285 times the density and pressure that sends
    the columns and rows through it to the literal form part way along."""
    st = rough_box(n, 11) if scale is None else blob_field(n, 11, scale)
    s, o = pair(n, st, recon="WENO5", vis_scheme="C8th", mu_law="constant", **cfg)
    gr = s.compute_rhs()
    orr = o.compute_rhs()
    for m in range(5):
        a, b = s.interior(gr[m]), o.interior(orr[m])
        assert np.all(np.isfinite(b)), m
        assert np.array_equal(a, b), (m, relmax(a, b), int(np.sum(a != b)))


@pytest.mark.parametrize("vis", ["C8th", "C6th", "C4th"])
@pytest.mark.parametrize("shape", [(64, 32, 24), (64, 20, 16)], ids=["ny32", "ny20"])
def test_fp64_viscous_pass_forms_bitwise(gpu, vis, shape, monkeypatch):
    """The FP64 viscous passes on periodic boxes with nx % 64 == 0 and
    ny % 4 == 0 stage each plane's x / y neighbourhood in LDS: the gradient/
    flux pass (NS3D_GF64_LDS=1, the default; 0 loads the stencils through
    the caches) and the divergence/stage pass (NS3D_VD64_LDS, likewise), in
    tiles of 64 x 4 or 64 x 8 cells (NS3D_FP64_TILE_Y).  All the
    combinations evaluate ViscousFlux.cpp:240-386 and
    range_kutta.cpp:11-161 in the same order on the same values, so the RHS,
    the divergence of stored viscous fluxes, two fixed-dt RK3 steps and
    three CFL steps must agree bit for bit.  A 64 x 32 x 24 box (a different
    spacing per axis) in 2 z segments of unequal length, and a 64 x 20 x 16
    box, whose 20 rows the 8-row tiles do not divide (those forms then run
    in 4-row tiles)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(5)
    st, Ma = fields.hit(nz, k0=2, Mt=0.3, seed=13)
    shape = (nz, ny, nx)
    x = [np.resize(a, shape) * (1.0 + 0.01 * rng.standard_normal(shape)) for a in st]
    out = {}
    for form in (("0", "0", "4"), ("1", "0", "4"), ("0", "1", "4"), ("1", "1", "4"), ("1", "1", "8")):
        monkeypatch.setenv("NS3D_GF64_LDS", form[0])
        monkeypatch.setenv("NS3D_VD64_LDS", form[1])
        monkeypatch.setenv("NS3D_FP64_TILE_Y", form[2])
        monkeypatch.setenv("NS3D_SEG_Z", "16")
        p, g = ns.make_params((nx, ny, nz), L=(2 * np.pi), Ma=Ma, Re=100.0, recon="WENO5",
                              vis_scheme=vis, Ly=np.pi, Lz=0.75 * np.pi)
        s = ns.Solver(p, g)
        s.upload_primitive(*x)
        r = [s.interior(a).copy() for a in s.compute_rhs()]
        s.compute_viscous_flux()
        dv = [s.interior(a).copy() for a in s.compute_vis_flux()]
        s.rk3_step(1e-3)
        s.rk3_step(1e-3)
        u2 = [s.interior(a).copy() for a in s.conserved()]
        s.advance_cfl(3)
        u5 = [s.interior(a).copy() for a in s.conserved()]
        out[form] = (r, dv, u2, u5)
        s.close()
    ref = out[("0", "0", "4")]
    for form, got in out.items():
        for part in range(4):
            for m, (a, b) in enumerate(zip(got[part], ref[part])):
                assert np.array_equal(a, b), (form, part, m, relmax(a, b))
