"""Long validation runs of the workloads (minutes each): the FP64 GPU path
against the CPU oracle over full trajectories (TGV 64^3 x 1000 steps, HIT
128^3 MDCD + LF x 500 steps), and the mixed-precision path against the FP64
path at 256^3 (WENO5 and UD7, 500 steps) with the kinetic-energy and
dissipation histories and the energy spectrum.

Skipped unless NS3D_VALIDATE=1 (the default GPU suite stays short).  Each
case prints its progress and writes a JSON record to
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

import ns3d_cpp_amd as ns  # noqa: E402
from ns3d_cpp_amd import fields, monitors  # noqa: E402
from oracle import Oracle  # noqa: E402

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("NS3D_VALIDATE") != "1", reason="long run: set NS3D_VALIDATE=1")]

OUT = os.path.join(ROOT, "gpurun_out", "validation")


def record(name, data):
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, name + ".json"), "w") as f:
        json.dump(data, f, indent=1)


def say(*a):
    print(*a, flush=True)


def rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2)) / (np.sqrt(np.sum(b * b)) + 1e-300))


def oracle_sum_phi(o, params):
    G = [o.interior(a) for a in o.compute_gradients()]
    T = o.interior(o.primitive()["T"])
    mu = monitors.mu_of_T(params, T)
    ux, uy, uz, vx, vy, vz, wx, wy, wz = G[:9]
    div = ux + vy + wz
    phi = mu * (2 * (ux * ux + vy * vy + wz * wz) - (2.0 / 3.0) * div * div
                + (uy + vx) ** 2 + (uz + wx) ** 2 + (vz + wy) ** 2)
    return float(np.sum(phi))


def run_against_oracle(name, n, state, nsteps, dt, every, **cfg):
    p, g = ns.make_params(n, **cfg)
    s = ns.Solver(p, g)
    s.upload_primitive(*state)
    o = Oracle(p, g)
    o.set_primitive(*state)
    n3 = float(n ** 3)
    hist = {"step": [], "ek_gpu": [], "ek_oracle": [], "eps_gpu": [], "eps_oracle": [], "dt": [],
            "dt_rel_diff": []}
    t0 = time.time()
    for it in range(nsteps + 1):
        if it % every == 0 or it == nsteps:
            hist["step"].append(it)
            hist["ek_gpu"].append(s.total_energy())
            hist["ek_oracle"].append(o.total_energy())
            hist["eps_gpu"].append(s.dissipation())
            hist["eps_oracle"].append(oracle_sum_phi(o, p) / n3)
            say(f"{name} step {it} ek {hist['ek_gpu'][-1]:.12e} / {hist['ek_oracle'][-1]:.12e} "
                f"eps {hist['eps_gpu'][-1]:.12e} / {hist['eps_oracle'][-1]:.12e} ({time.time() - t0:.0f} s)")
        if it == nsteps:
            break
        d = dt if dt > 0 else s.timestep()
        if dt <= 0:
            # both sides step with the GPU's dt; the oracle's own CFL dt may
            # differ in the last bit once the states do
            hist["dt_rel_diff"].append(abs(o.timestep() / d - 1.0))
        hist["dt"].append(d)
        s.rk3_step(d)
        st, used = o.step(d)
        assert st == 0 and used == d
    U = [rel_l2(s.interior(a), o.interior(b)) for a, b in zip(s.conserved(), o.conserved())]
    pg, po = s.primitive(), o.primitive()
    kg, eg = monitors.energy_spectrum(*(s.interior(pg[c]) for c in "uvw"))
    ko, eo = monitors.energy_spectrum(*(o.interior(po[c]) for c in "uvw"))
    mask = eo > 1e-30
    spec = float(np.max(np.abs(eg[mask] - eo[mask]) / eo[mask]))
    ek_err = max(abs(a - b) / abs(b) for a, b in zip(hist["ek_gpu"], hist["ek_oracle"]))
    eps_err = max(abs(a - b) / abs(b) for a, b in zip(hist["eps_gpu"], hist["eps_oracle"]))
    res = {"case": name, "n": n, "steps": nsteps, "rel_l2_U": U, "max_rel_l2_U": max(U),
           "cfl_dt_max_rel_diff": max(hist["dt_rel_diff"]) if hist["dt_rel_diff"] else 0.0,
           "ek_hist_max_rel": ek_err, "eps_hist_max_rel": eps_err, "spectrum_max_rel": spec,
           "oracle_threads": int(__import__("oracle").load().orc_num_threads()),
           "wall_s": time.time() - t0, "history": hist}
    record(name, res)
    say(name, json.dumps({k: v for k, v in res.items() if k != "history"}))
    return res


def test_c1_tgv64_1000_steps_against_oracle(gpu):
    res = run_against_oracle("c1_tgv64", 64, fields.tgv(64, Ma=0.1), 1000, 1e-3, 50, Re=1600.0, Ma=0.1,
                             Pr=0.72, recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th",
                             mu_law="constant")
    assert res["max_rel_l2_U"] <= 1e-12
    assert res["ek_hist_max_rel"] <= 1e-12 and res["eps_hist_max_rel"] <= 1e-12
    assert res["spectrum_max_rel"] <= 1e-10


def test_c2_hit128_mdcd_lf_500_steps_against_oracle(gpu):
    st, Ma = fields.hit(128, k0=4, Mt=0.3, seed=2024)
    res = run_against_oracle("c2_hit128", 128, st, 500, 0.0, 50, Re=100.0, Ma=Ma, Pr=0.72, recon="MDCD",
                             fvs_type="LaxFriedrichs", vis_scheme="C8th", cfl=0.5,
                             sutherland_s=110.4 / 273.15)
    assert res["max_rel_l2_U"] <= 1e-12
    assert res["ek_hist_max_rel"] <= 1e-12 and res["eps_hist_max_rel"] <= 1e-12
    assert res["cfl_dt_max_rel_diff"] <= 1e-14


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
                stats = monitors.solver_statistics(s)
                hist["eps_spec_" + prec].append(stats["dissipation"])
            hist["eps_fp64"].append(d.dissipation())
            # the mixed context has no FP64 dissipation kernel: physical-space
            # dissipation from its downloaded primitives through an FP64 context
            pm = m.primitive()
            p64, g64 = ns.make_params(n, precision="fp64", **cfg)
            tmp = ns.Solver(p64, g64)
            tmp.upload_primitive(*(m.interior(pm[c]) for c in ("rho", "u", "v", "w", "p")))
            hist["eps_mixed"].append(tmp.dissipation())
            tmp.close()
            say(f"c3 {recon} step {it} ek {hist['ek_fp64'][-1]:.10e} / {hist['ek_mixed'][-1]:.10e} "
                f"({time.time() - t0:.0f} s)")
        if it == nsteps:
            break
        dt = d.timestep()
        d.rk3_step(dt)
        m.rk3_step(dt)
    U = [rel_l2(d.interior(b), d.interior(a)) for a, b in zip(d.conserved(), m.conserved())]
    pd, pm = d.primitive(), m.primitive()
    k, ed = monitors.energy_spectrum(*(d.interior(pd[c]) for c in "uvw"))
    _, em = monitors.energy_spectrum(*(m.interior(pm[c]) for c in "uvw"))
    kk = k <= 85
    spec = float(np.max(np.abs(em[kk][1:] - ed[kk][1:]) / ed[kk][1:]))
    ek_err = max(abs(a - b) / abs(b) for a, b in zip(hist["ek_mixed"], hist["ek_fp64"]))
    eps_err = max(abs(a - b) / abs(b) for a, b in zip(hist["eps_mixed"], hist["eps_fp64"]))
    res = {"case": f"c3_hit256_{recon}", "steps": nsteps, "rel_l2_U": U, "max_rel_l2_U": max(U),
           "ek_hist_max_rel": ek_err, "eps_hist_max_rel": eps_err, "spectrum_k_le_85_max_rel": spec,
           "wall_s": time.time() - t0, "history": hist,
           "spectrum_fp64": ed[kk].tolist(), "spectrum_mixed": em[kk].tolist()}
    record(f"c3_hit256_{recon}", res)
    say(json.dumps({k: v for k, v in res.items() if k not in ("history", "spectrum_fp64", "spectrum_mixed")}))
    assert max(U) <= 1e-5
