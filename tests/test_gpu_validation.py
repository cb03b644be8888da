"""Validation runs of the workloads at their stated sizes and lengths, in the
default GPU suite:

* TGV 64^3, WENO5 + epsilon-modified Steger-Warming (README.md:179, eps =
  0.1, 1 % of the initial sound speed 1/Ma = 10) + C8 viscous terms,
  constant mu, fixed dt = 1e-3 for 1000 steps: the FP64 GPU path against the
  CPU oracle on U, on the kinetic-energy and dissipation histories and on the
  final energy spectrum (rel-L2 over the shells), each within 1e-12.
* Isotropic turbulence 128^3 (k0 = 4, Mt = 0.3, Re = 100), MDCD + local
  Lax-Friedrichs, Sutherland mu (S/T_ref = 110.4/273.15), CFL 0.5 for 500
  steps: the same comparison, plus the CFL time step of both sides.
* Isotropic turbulence 256^3 (Re = 200) with WENO5 and with UD7, 500 CFL
  steps: the mixed-precision path against the FP64 path (rel-L2 of U within
  1e-5; kinetic energy, dissipation and spectrum to k = 85 recorded).
* Isotropic turbulence 512^3 (k0 = 8, Mt = 0.5, Re = 400), WENO5 +
  Steger-Warming: conservation and mirror symmetry of the FP64 and mixed
  right-hand sides, and 20 CFL steps of mixed against FP64 within 1e-5.

The long trajectories run in segments of at most a few hundred steps, one
test per segment, each continuing the previous one's state (a segment run
alone replays the earlier ones first).  Each case writes a JSON record to
gpurun_out/validation/<case>.json.
"""
import json
import os
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import monitors  # noqa: E402
from oracle import Oracle, load  # noqa: E402

import ns3d_cpp_amd as ns  # noqa: E402
from ns3d_cpp_amd import fields  # noqa: E402

pytestmark = pytest.mark.gpu

OUT = os.path.join(ROOT, "gpurun_out", "validation")
C1_SW_EPS = 0.1


def record(name, data):
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, name + ".json"), "w") as f:
        json.dump(data, f, indent=1)


def rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2)) / (np.sqrt(np.sum(b * b)) + 1e-300))


def oracle_mean_phi(o, params):
    """Volume-mean viscous dissipation of the oracle's state from its own
    gradients (the quantity ns3d_dissipation reduces on the device)."""
    G = [o.interior(a) for a in o.compute_gradients()]
    T = o.interior(o.primitive()["T"])
    mu = monitors.mu_of_T(params, T)
    ux, uy, uz, vx, vy, vz, wx, wy, wz = G[:9]
    div = ux + vy + wz
    phi = mu * (2 * (ux * ux + vy * vy + wz * wz) - (2.0 / 3.0) * div * div
                + (uy + vx) ** 2 + (uz + wx) ** 2 + (vz + wy) ** 2)
    return float(np.sum(phi)) / T.size


class OracleRun:
    """A GPU FP64 trajectory and the oracle's, stepped side by side."""

    def __init__(self, name, n, state, dt, every, **cfg):
        self.name, self.n, self.dt, self.every = name, n, dt, every
        self.p, self.g = ns.make_params(n, **cfg)
        self.s = ns.Solver(self.p, self.g)
        self.s.upload_primitive(*state)
        self.o = Oracle(self.p, self.g)
        self.o.set_primitive(*state)
        self.step = 0
        self.t0 = time.time()
        self.hist = {"step": [], "ek_gpu": [], "ek_oracle": [], "eps_gpu": [], "eps_oracle": [],
                     "dt": [], "dt_rel_diff": []}
        self.monitor()

    def monitor(self):
        h = self.hist
        h["step"].append(self.step)
        h["ek_gpu"].append(self.s.total_energy())
        h["ek_oracle"].append(self.o.total_energy())
        h["eps_gpu"].append(self.s.dissipation())
        h["eps_oracle"].append(oracle_mean_phi(self.o, self.p))

    def advance_to(self, last):
        while self.step < last:
            d = self.dt if self.dt > 0 else self.s.timestep()
            if self.dt <= 0:
                # both sides step with the GPU's CFL dt; the oracle's own
                # may differ in the last bits once the states do
                self.hist["dt_rel_diff"].append(abs(self.o.timestep() / d - 1.0))
            self.hist["dt"].append(d)
            self.s.rk3_step(d)
            st, used = self.o.step(d)
            assert st == 0 and used == d
            self.step += 1
            if self.step % self.every == 0:
                self.monitor()

    def errors(self):
        h = self.hist
        U = [rel_l2(self.s.interior(a), self.o.interior(b)) for a, b in zip(self.s.conserved(), self.o.conserved())]
        ek = max(abs(a - b) / abs(b) for a, b in zip(h["ek_gpu"], h["ek_oracle"]))
        eps = max(abs(a - b) / abs(b) for a, b in zip(h["eps_gpu"], h["eps_oracle"]))
        return {"step": self.step, "rel_l2_U": U, "max_rel_l2_U": max(U), "ek_hist_max_rel": ek,
                "eps_hist_max_rel": eps,
                "cfl_dt_max_rel_diff": max(h["dt_rel_diff"]) if h["dt_rel_diff"] else 0.0}

    def spectrum_error(self):
        """rel-L2 of the energy spectrum E(k) over all shells, and the largest
        relative difference of one shell (the weakest shells carry it)."""
        pg, po = self.s.primitive(), self.o.primitive()
        _, eg = monitors.energy_spectrum(*(self.s.interior(pg[c]) for c in "uvw"))
        _, eo = monitors.energy_spectrum(*(self.o.interior(po[c]) for c in "uvw"))
        mask = eo > 1e-30
        return rel_l2(eg, eo), float(np.max(np.abs(eg[mask] - eo[mask]) / eo[mask]))

    def finish(self, extra=None):
        res = {"case": self.name, "n": self.n, "steps": self.step, **self.errors(),
               "oracle_threads": int(load().orc_num_threads()), "wall_s": time.time() - self.t0,
               "history": self.hist}
        res.update(extra or {})
        record(self.name, res)
        self.s.close()
        self.o.close()
        return res


_RUNS = {}


def c1_run():
    if "c1" not in _RUNS:
        _RUNS["c1"] = OracleRun("c1_tgv64", 64, fields.tgv(64, Ma=0.1), 1e-3, 50, Re=1600.0, Ma=0.1, Pr=0.72,
                                recon="WENO5", fvs_type="StagerWarming", sw_eps=C1_SW_EPS, vis_scheme="C8th",
                                mu_law="constant")
    return _RUNS["c1"]


@pytest.mark.parametrize("last", [250, 500, 750, 1000])
def test_c1_tgv64_eps_sw_constant_mu_against_oracle(gpu, last):
    r = c1_run()
    r.advance_to(last)
    e = r.errors()
    assert e["max_rel_l2_U"] <= 1e-12, e
    assert e["ek_hist_max_rel"] <= 1e-12 and e["eps_hist_max_rel"] <= 1e-12, e
    if last == 1000:
        spec, shell = r.spectrum_error()
        r.finish({"spectrum_rel_l2": spec, "spectrum_shell_max_rel": shell, "sw_eps": C1_SW_EPS,
                  "mu_law": "constant"})
        del _RUNS["c1"]
        assert spec <= 1e-12 and shell <= 1e-10, (spec, shell)


def c2_run():
    if "c2" not in _RUNS:
        st, Ma = fields.hit(128, k0=4, Mt=0.3, seed=2024)
        _RUNS["c2"] = OracleRun("c2_hit128", 128, st, 0.0, 50, Re=100.0, Ma=Ma, Pr=0.72, recon="MDCD",
                                fvs_type="LaxFriedrichs", vis_scheme="C8th", cfl=0.5, sutherland_s=110.4 / 273.15)
    return _RUNS["c2"]


@pytest.mark.parametrize("last", [100, 200, 300, 400, 500])
def test_c2_hit128_mdcd_lf_against_oracle(gpu, last):
    r = c2_run()
    r.advance_to(last)
    e = r.errors()
    assert e["max_rel_l2_U"] <= 1e-12, e
    assert e["ek_hist_max_rel"] <= 1e-12 and e["eps_hist_max_rel"] <= 1e-12, e
    assert e["cfl_dt_max_rel_diff"] <= 1e-14, e
    if last == 500:
        spec, shell = r.spectrum_error()
        r.finish({"spectrum_rel_l2": spec, "spectrum_shell_max_rel": shell})
        del _RUNS["c2"]
        assert spec <= 1e-12 and shell <= 1e-10, (spec, shell)


def _fp64_dissipation_of(n, cfg, solver):
    """Physical-space dissipation of a context's state through an FP64
    context (the mixed context has no FP64 dissipation kernel)."""
    pm = solver.primitive()
    p64, g64 = ns.make_params(n, **{**cfg, "precision": "fp64"})
    tmp = ns.Solver(p64, g64)
    tmp.upload_primitive(*(solver.interior(pm[c]) for c in ("rho", "u", "v", "w", "p")))
    val = tmp.dissipation()
    tmp.close()
    return val


@pytest.mark.parametrize("recon", ["WENO5", "UPWIND_7TH"])
def test_c3_hit256_mixed_against_fp64(gpu, recon):
    n, nsteps, every = 256, 500, 50
    st, Ma = fields.hit(n, k0=4, Mt=0.3, seed=2024)
    cfg = dict(Re=200.0, Ma=Ma, Pr=0.72, recon=recon, fvs_type="StagerWarming", vis_scheme="C8th", cfl=0.5,
               sutherland_s=110.4 / 273.15)
    sv = {}
    for prec in ("fp64", "mixed"):
        p, g = ns.make_params(n, precision=prec, **cfg)
        sv[prec] = ns.Solver(p, g)
        sv[prec].upload_primitive(*st)
    del st
    d, m = sv["fp64"], sv["mixed"]
    hist = {"step": [], "ek_fp64": [], "ek_mixed": [], "eps_fp64": [], "eps_mixed": [],
            "eps_spec_fp64": [], "eps_spec_mixed": []}
    t0 = time.time()
    for it in range(nsteps + 1):
        if it % every == 0:
            hist["step"].append(it)
            for prec, s in sv.items():
                hist["ek_" + prec].append(s.total_energy())
                hist["eps_spec_" + prec].append(monitors.solver_statistics(s)["dissipation"])
            hist["eps_fp64"].append(d.dissipation())
            hist["eps_mixed"].append(_fp64_dissipation_of(n, cfg, m))
        if it == nsteps:
            break
        dt = d.timestep()
        d.rk3_step(dt)
        m.rk3_step(dt)
    U = [rel_l2(d.interior(b), d.interior(a)) for a, b in zip(d.conserved(), m.conserved())]
    pd, pm = d.primitive(), m.primitive()
    k, ed = monitors.energy_spectrum(*(d.interior(pd[c]) for c in "uvw"))
    _, em = monitors.energy_spectrum(*(m.interior(pm[c]) for c in "uvw"))
    d.close()
    m.close()
    kk = k <= 85
    spec = float(np.max(np.abs(em[kk][1:] - ed[kk][1:]) / ed[kk][1:]))
    ek_err = max(abs(a - b) / abs(b) for a, b in zip(hist["ek_mixed"], hist["ek_fp64"]))
    eps_err = max(abs(a - b) / abs(b) for a, b in zip(hist["eps_mixed"], hist["eps_fp64"]))
    res = {"case": f"c3_hit256_{recon}", "steps": nsteps, "rel_l2_U": U, "max_rel_l2_U": max(U),
           "ek_hist_max_rel": ek_err, "eps_hist_max_rel": eps_err, "spectrum_k_le_85_max_rel": spec,
           "wall_s": time.time() - t0, "history": hist,
           "spectrum_fp64": ed[kk].tolist(), "spectrum_mixed": em[kk].tolist()}
    record(f"c3_hit256_{recon}", res)
    assert max(U) <= 1e-5, U
    assert ek_err <= 1e-5 and eps_err <= 1e-4, (ek_err, eps_err)


# ------------------------------------------------------------------ 512^3
C4 = dict(n=512, k0=8.0, Mt=0.5, Re=400.0, seed=2024)
_C4_FIELD = {}


def c4_field():
    if "st" not in _C4_FIELD:
        _C4_FIELD["st"] = fields.hit(C4["n"], k0=C4["k0"], Mt=C4["Mt"], seed=C4["seed"])
    return _C4_FIELD["st"]


def c4_params(prec):
    (_, Ma) = c4_field()
    return ns.make_params(C4["n"], Re=C4["Re"], Ma=Ma, Pr=0.72, recon="WENO5", fvs_type="StagerWarming",
                          vis_scheme="C8th", cfl=0.5, sutherland_s=110.4 / 273.15, precision=prec)


def _mirror_x(a):
    # x_i -> x_{-i} on the periodic grid x_i = i dx
    return np.roll(a[:, :, ::-1], 1, axis=2)


@pytest.mark.parametrize("prec,tol_sum,tol_mirror", [("fp64", 1e-12, 1e-11), ("mixed", 1e-6, 1e-5)])
def test_c4_hit512_rhs_conservation_and_mirror(gpu, prec, tol_sum, tol_mirror):
    """512^3 (k0 = 8, Mt = 0.5: shocklets): each RHS component sums to zero
    over the periodic box (telescoping fluxes), and the RHS of the x-mirrored
    field is the mirrored RHS with u -> -u."""
    (rho, u, v, w, p), _ = c4_field()
    pr, g = c4_params(prec)
    s = ns.Solver(pr, g)
    s.upload_primitive(rho, u, v, w, p)
    r = [s.interior(a).copy() for a in s.compute_rhs()]
    for m in range(5):
        assert abs(r[m].sum()) <= tol_sum * np.abs(r[m]).sum(), (prec, m)
    s.upload_primitive(_mirror_x(rho), -_mirror_x(u), _mirror_x(v), _mirror_x(w), _mirror_x(p))
    rm = s.compute_rhs()
    s.close()
    sign = [1, -1, 1, 1, 1]
    for m in range(5):
        dev = np.abs(s.interior(rm[m]) - sign[m] * _mirror_x(r[m])).max()
        assert dev <= tol_mirror * np.abs(r[m]).max(), (prec, m, dev)


def test_c4_hit512_mixed_against_fp64_steps(gpu):
    """20 CFL steps at 512^3, mixed against FP64 from the same state with the
    FP64 path's time step: rel-L2 of U within 1e-5."""
    st, _ = c4_field()
    sv = {}
    for prec in ("fp64", "mixed"):
        pr, g = c4_params(prec)
        sv[prec] = ns.Solver(pr, g)
        sv[prec].upload_primitive(*st)
    _C4_FIELD.clear()
    d, m = sv["fp64"], sv["mixed"]
    t = 0.0
    for _ in range(20):
        dt = d.timestep()
        d.rk3_step(dt)
        m.rk3_step(dt)
        t += dt
    U = [rel_l2(d.interior(b), d.interior(a)) for a, b in zip(d.conserved(), m.conserved())]
    d.close()
    m.close()
    record("c4_hit512_mixed_vs_fp64", {"case": "c4_hit512", **C4, "steps": 20, "time": t, "rel_l2_U": U})
    assert max(U) <= 1e-5, U
