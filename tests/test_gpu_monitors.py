"""Device monitors against their host restatements: the energy spectrum and
the turbulence statistics (numpy FFT, oracle/monitors.py, pinned by
closed forms in test_monitors.py) and the residual diagnostics."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import monitors  # noqa: E402

import ns3d_cpp_amd as ns  # noqa: E402
from ns3d_cpp_amd import fields  # noqa: E402

pytestmark = pytest.mark.gpu

CFG = dict(Re=100.0, Pr=0.72, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th",
           sutherland_s=110.4 / 273.15, cfl=0.5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spectrum_and_statistics(gpu, precision):
    n = 32
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=12)
    p, g = ns.make_params(n, Ma=Ma, precision=precision, **CFG)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    s.rk3_step(s.timestep())
    pr = s.primitive()
    it = s.interior
    k, ek = s.energy_spectrum()
    kh, ekh = monitors.energy_spectrum(it(pr["u"]), it(pr["v"]), it(pr["w"]))
    assert len(ek) == 2 * n + 1
    mask = ekh > 1e-30
    assert np.all(np.abs(ek[mask] - ekh[mask]) <= 1e-12 * ekh[mask] + 1e-18)
    assert np.all(np.abs(ek[~mask]) <= 1e-25)
    stats = s.turbulence_statistics()
    p64 = s.params
    if precision == "fp64":
        du_dx = it(s.compute_gradients()[0])
    else:
        du_dx = monitors._diff_x(it(pr["u"]), s)
    ref = monitors.turbulence_statistics(p64, it(pr["rho"]), it(pr["u"]), it(pr["v"]), it(pr["w"]),
                                         it(pr["T"]), du_dx)
    for key, val in ref.items():
        assert stats[key] == pytest.approx(val, rel=1e-11), key


def test_single_mode_spectrum_on_device(gpu):
    n = 16
    x = np.arange(n) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = 0.1 * np.sin(X) * np.cos(2 * Y)        # |k|^2 = 5 -> shell 2
    st = [np.ones_like(X), u, 0 * X, 0 * X, np.full_like(X, 1.0 / (1.4 * 0.1 ** 2))]
    p, g = ns.make_params(n, Ma=0.1, **CFG)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    k, ek = s.energy_spectrum()
    assert ek[2] == pytest.approx(0.5 * np.mean(u * u), rel=1e-13)
    assert np.sum(np.abs(np.delete(ek, 2))) < 1e-28


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_residuals(gpu, precision):
    n = 16
    st, Ma = fields.hit(n, k0=2, Mt=0.3, seed=4)
    p, g = ns.make_params(n, Ma=Ma, precision=precision, **CFG)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    U0 = [s.interior(a).copy() for a in s.conserved()]
    s.rk3_step(2e-3)
    U1 = [s.interior(a) for a in s.conserved()]
    res = s.residuals()
    dx3 = s.grid.dx ** 3
    for m in range(5):
        want = np.sqrt(np.sum((U1[m] - U0[m]) ** 2) * dx3) / np.abs(U1[m]).max()
        assert res[m] == pytest.approx(want, rel=1e-9), m
