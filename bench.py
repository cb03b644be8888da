"""Throughput of the RK3 RHS path on one MI355X: FP64 vs mixed precision.

Workload (default): decaying compressible isotropic turbulence on a 256^3
periodic box [0, 2pi)^3 -- seeded solenoidal random-phase velocity with
E(k) = 16 sqrt(2/pi) (u0^2/k0) (k/k0)^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 1,
rho = 1, T = 1, Mt = 0.3, Re = 200, Pr = 0.72, Sutherland mu(T) with
S/T_ref = 110.4/273.15; WENO5 + Steger-Warming + C8 central viscous terms;
SSP-RK3 with the CFL (0.5) time step recomputed every step.

A step = one RK3 step (3 RHS + 3 stage updates + the CFL reduction), run as
ns3d_advance_cfl runs them: dt = cfl * min stays on the device, so the timed
loop has no host round trip per step.  In mixed mode the third stage reduces
the next step's CFL minimum from the state it writes; only the first step of
a call runs the CFL pass.
Metric: cell-updates/s = N^3 / (seconds per step), for the mixed-precision
path (`value`) and the FP64 path (`fp64`).  Inputs are resident in HBM
before the timed region.  Per-kernel durations come from HIP events recorded
around every launch on the solver's stream inside the timed region.  The
mixed path's error is measured against the FP64 path over --err-steps steps
from the same initial state, with the FP64 path's CFL time step.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK3 cell-updates/s @256^3 FP64 vs mixed-FP32; % HBM peak; adjoint/fwd cost"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per interior cell per launch of each kernel (every
# array touched once per launch; halos reused on chip).  The stress tensor is
# symmetric, so 9 distinct viscous-flux arrays are written and read.
ALG_BYTES = {
    "fp64": {
        "sweep_x": 80,        # read rho,u,v,w,p; write A(5)
        "sweep_y": 120,       # + read A(5)
        "sweep_z": 120,
        "gradflux": 104,      # read u,v,w,T; write 9 viscous fluxes
        "vdiv_stage": 240,    # read 9 fluxes, A(5), W(5), U0(5); write W(6)
        "cfl": 40,            # read rho,u,v,w,p
    },
    "mixed": {
        "sweep_x": 40,        # read FP32 shadow (5); write FP32 A(5)
        "sweep_y": 60,        # + read A(5)
        "sweep_z": 60,
        "gradflux": 52,       # read u,v,w,T-1 (FP32); write 9 FP32 fluxes
        # read 9 FP32 fluxes, A32, U64 (+ dU32 in stages 2, 3); write dU32,
        # shadow(6) (+ U64 in stage 3): 140, 160, 200 B, mean over a step
        "vdiv_stage": 500 / 3,
        # the stage kernel with the z sweep folded in: the stage's bytes and
        # the sweep's shadow read (5); the sweep's A read and write are gone:
        # 160, 180, 220 B
        "vdiv_zs_stage": 500 / 3 + 20,
        "cfl": 20,            # read the FP32 shadow
    },
}
# Per cell-update byte model of the full step used for "% of HBM peak"
# (3 RHS + 3 stages + the CFL read; FP64 2448 B, mixed 1564 B), as the
# metric defines it; the mixed step no longer reads its 20 B of CFL input
# (its third stage takes the minimum from values it already holds).
STEP_MODEL_BYTES = {"fp64": 2448, "mixed": 1564}


def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    lib.hipEventSynchronize.argtypes = [C.c_void_p]
    lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def make(ns, n, Ma, Re, recon, precision):
    return ns.make_params(n, Re=Re, Ma=Ma, Pr=0.72, cfl=0.5, recon=recon, fvs_type="StagerWarming",
                          vis_scheme="C8th", sutherland_s=110.4 / 273.15, precision=precision)


def traffic_from_profiles(prec, kernel, n, recon="WENO5"):
    """HBM bytes per launch of `kernel` at n^3 measured by rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration) and
    committed under profiles/traffic_<prec>_hit<n>.json; None where no
    profile of that size and scheme exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{prec}_hit{n}.json")
    if recon != "WENO5" or not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(kernel)
    except (OSError, ValueError):
        return None


def time_path(ns, H, n, st, Ma, Re, recon, prec, steps, warmup):
    p, g = make(ns, n, Ma, Re, recon, prec)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    s.advance_cfl(warmup)
    s.sync()
    H.hipDeviceSynchronize()
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    H.hipEventCreate(C.byref(ev0))
    H.hipEventCreate(C.byref(ev1))
    s.profile(True)
    t0 = time.perf_counter()
    H.hipEventRecord(ev0, s.stream())
    # each step: the CFL reduction, dt = cfl * min on the device, 3 stages
    # (mixed: the reduction rides on the previous step's third stage)
    s.advance_cfl(steps)
    H.hipEventRecord(ev1, s.stream())
    H.hipEventSynchronize(ev1)
    s.sync()
    H.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    ms = C.c_float()
    H.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    rep = s.profile_report()
    s.profile(False)
    s.close()
    ms_step = ms.value / steps
    value = n ** 3 / (ms_step * 1e-3)
    alg = ALG_BYTES[prec]
    kernels = {}
    for k, (tot, cnt) in sorted(rep.items(), key=lambda kv: -kv[1][0]):
        e = {"avg_ms": round(tot / cnt, 4), "launches": cnt, "share": round(tot / ms.value, 4)}
        if k in alg:
            e["GBps"] = round(alg[k] * n ** 3 / (tot / cnt * 1e-3) / 1e9, 1)
        kernels[k] = e
    dom = max(rep, key=lambda k: rep[k][0])
    dom_avg = rep[dom][0] / rep[dom][1]
    achieved = alg.get(dom, 0) * n ** 3 / (dom_avg * 1e-3) / 1e9
    traffic = traffic_from_profiles(prec, dom, n, recon)
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}
    if prec == "fp64" and dom.startswith("sweep"):
        # the fraction is of HBM, but what bounds these kernels is VALU issue
        roof["note"] = ("VALU-bound: the reference's correctly rounded FP64 divisions "
                        "(94 per cell and direction) dominate the sweep")
    # the kernels the timed steps ran, each with its launches per step
    step_bytes = round(sum(alg[k] * rep[k][1] / steps for k in rep if k in alg), 1)
    return {"value": value, "ms_per_step": ms_step, "wall_s_timed": round(wall, 4), "kernels": kernels,
            "roofline": roof,
            "kernel_bytes_per_cell_update": step_bytes,
            "kernel_bytes_hbm_frac": round(step_bytes * value / (HBM_PEAK_GBS * 1e9), 4),
            "step_model": {"bytes_per_cell_update": STEP_MODEL_BYTES[prec],
                           "hbm_frac": round(STEP_MODEL_BYTES[prec] * value / (HBM_PEAK_GBS * 1e9), 4)}}


def mixed_error(ns, n, st, Ma, Re, recon, nsteps):
    """rel-L2 of each conserved component, mixed vs FP64, after nsteps."""
    pd, gd = make(ns, n, Ma, Re, recon, "fp64")
    pm, gm = make(ns, n, Ma, Re, recon, "mixed")
    d, m = ns.Solver(pd, gd), ns.Solver(pm, gm)
    d.upload_primitive(*st)
    m.upload_primitive(*st)
    t = 0.0
    for _ in range(nsteps):
        dt = d.timestep()
        d.rk3_step(dt)
        m.rk3_step(dt)
        t += dt
    ud, um = d.conserved(), m.conserved()
    errs = []
    for a, b in zip(um, ud):
        a, b = d.interior(a), d.interior(b)
        errs.append(float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b * b))))
    ek = (d.total_energy(), m.total_energy())
    d.close()
    m.close()
    return {"steps": nsteps, "time": t, "rel_l2_U": errs, "max_rel_l2": max(errs),
            "kinetic_energy_fp64": ek[0], "kinetic_energy_mixed": ek[1]}


def adjoint_cost(ns, H, n, nsteps, ckpt, seed):
    """Checkpointed adjoint of J = sum_n dt * eps(U^n) over nsteps FP64 RK3
    steps (isotropic turbulence n^3, k0=4, Mt=0.3, Re=100, WENO5+SW+C8,
    fixed dt = the CFL dt of the initial field), against the plain forward
    pass over the same steps.  Times are HIP events on the solver's stream."""
    from ns3d_cpp_amd import fields
    st, Ma = fields.hit(n, k0=4.0, Mt=0.3, seed=seed)
    p, g = make(ns, n, Ma, 100.0, "WENO5", "fp64")
    s = ns.Solver(p, g)
    U0 = None
    s.upload_primitive(*st)
    U0 = s.conserved()
    dt = s.timestep()
    ev = [C.c_void_p() for _ in range(4)]
    for e in ev:
        H.hipEventCreate(C.byref(e))
    # warm-up of every kernel involved
    s.advance(2, dt)
    s.upload_conserved(*U0)
    s.adj_forward(2, dt, 1)
    s.adj_backward()
    s.upload_conserved(*U0)
    s.sync()
    H.hipEventRecord(ev[0], s.stream())
    s.advance(nsteps, dt)
    H.hipEventRecord(ev[1], s.stream())
    H.hipEventSynchronize(ev[1])
    fwd = C.c_float()
    H.hipEventElapsedTime(C.byref(fwd), ev[0], ev[1])
    # timed without per-kernel profiling: the reverse pass then recomputes
    # the next checkpoint segment on a second stream while it reverses one
    s.upload_conserved(*U0)
    s.sync()
    H.hipEventRecord(ev[0], s.stream())
    J = s.adj_forward(nsteps, dt, ckpt)
    H.hipEventRecord(ev[1], s.stream())
    lam0, dJ_dinvRe = s.adj_backward()
    H.hipEventRecord(ev[2], s.stream())
    H.hipEventSynchronize(ev[2])
    t_f, t_b = C.c_float(), C.c_float()
    H.hipEventElapsedTime(C.byref(t_f), ev[0], ev[1])
    H.hipEventElapsedTime(C.byref(t_b), ev[1], ev[2])
    mem = s.adj_memory()
    # the per-kernel breakdown, from a profiled pass (one stream, so each
    # kernel's time is its own); its gradient is the same bit for bit
    s.upload_conserved(*U0)
    s.profile(True)
    s.sync()
    H.hipEventRecord(ev[0], s.stream())
    Jp = s.adj_forward(nsteps, dt, ckpt)
    lamp, thp = s.adj_backward()
    H.hipEventRecord(ev[2], s.stream())
    H.hipEventSynchronize(ev[2])
    t_seq = C.c_float()
    H.hipEventElapsedTime(C.byref(t_seq), ev[0], ev[2])
    rep = s.profile_report()
    s.profile(False)
    seq_same = Jp == J and thp == dJ_dinvRe and all(
        np.array_equal(s.interior(a), s.interior(b)) for a, b in zip(lamp, lam0))
    # the same pass on one stream without profiling (a context made with
    # NS3D_ADJ_OVERLAP=0): what the overlap itself gains
    prev = os.environ.get("NS3D_ADJ_OVERLAP")
    os.environ["NS3D_ADJ_OVERLAP"] = "0"
    try:
        s1 = ns.Solver(p, g)
    finally:
        if prev is None:
            del os.environ["NS3D_ADJ_OVERLAP"]
        else:
            os.environ["NS3D_ADJ_OVERLAP"] = prev
    s1.upload_conserved(*U0)
    s1.adj_forward(2, dt, 1)
    s1.adj_backward()
    s1.upload_conserved(*U0)
    s1.sync()
    H.hipEventRecord(ev[0], s1.stream())
    J1s = s1.adj_forward(nsteps, dt, ckpt)
    lam1s, th1s = s1.adj_backward()
    H.hipEventRecord(ev[2], s1.stream())
    H.hipEventSynchronize(ev[2])
    t_one = C.c_float()
    H.hipEventElapsedTime(C.byref(t_one), ev[0], ev[2])
    one_same = J1s == J and th1s == dJ_dinvRe and all(
        np.array_equal(s1.interior(a), s1.interior(b)) for a, b in zip(lam1s, lam0))
    one_mem = s1.adj_memory()
    s1.close()
    base = ns.lib().ns3d_cells(s.ctx) * 40 * 8
    gnorm = float(np.sqrt(sum(np.sum(s.interior(a) ** 2) for a in lam0)))
    # the same gradient with every stage state of the run kept in HBM (one
    # checkpoint segment: no recomputation), where the device memory allows
    s.upload_conserved(*U0)
    s.sync()
    H.hipEventRecord(ev[0], s.stream())
    J1 = s.adj_forward(nsteps, dt, nsteps)
    H.hipEventRecord(ev[1], s.stream())
    lam1, _ = s.adj_backward()
    H.hipEventRecord(ev[2], s.stream())
    H.hipEventSynchronize(ev[2])
    a_f, a_b = C.c_float(), C.c_float()
    H.hipEventElapsedTime(C.byref(a_f), ev[0], ev[1])
    H.hipEventElapsedTime(C.byref(a_b), ev[1], ev[2])
    same = J1 == J and all(np.array_equal(s.interior(a), s.interior(b)) for a, b in zip(lam0, lam1))
    store_all = {"ckpt_every": nsteps, "adj_forward_ms": round(a_f.value, 3), "adj_backward_ms": round(a_b.value, 3),
                 "cost_vs_forward": round((a_f.value + a_b.value) / fwd.value, 3),
                 "device_GB": round((s.adj_memory() + base) / 1e9, 3), "gradient_bitwise_equal": bool(same)}
    # the same gradient from the context switched to mixed precision: mixed
    # steps forward and recomputed, the FP64 scheme's derivatives at each
    # stage state (checkpoint every ckpt steps as above)
    s.set_precision("mixed")
    s.upload_conserved(*U0)
    s.advance(2, dt)
    s.upload_conserved(*U0)
    s.adj_forward(2, dt, 1)
    s.adj_backward()
    s.upload_conserved(*U0)
    s.sync()
    H.hipEventRecord(ev[0], s.stream())
    s.advance(nsteps, dt)
    H.hipEventRecord(ev[1], s.stream())
    H.hipEventSynchronize(ev[1])
    m_fwd = C.c_float()
    H.hipEventElapsedTime(C.byref(m_fwd), ev[0], ev[1])
    s.upload_conserved(*U0)
    s.sync()
    H.hipEventRecord(ev[0], s.stream())
    Jm = s.adj_forward(nsteps, dt, ckpt)
    H.hipEventRecord(ev[1], s.stream())
    lamm, thm = s.adj_backward()
    H.hipEventRecord(ev[2], s.stream())
    H.hipEventSynchronize(ev[2])
    m_f, m_b = C.c_float(), C.c_float()
    H.hipEventElapsedTime(C.byref(m_f), ev[0], ev[1])
    H.hipEventElapsedTime(C.byref(m_b), ev[1], ev[2])
    gdiff = float(np.sqrt(sum(np.sum((s.interior(a) - s.interior(b)) ** 2) for a, b in zip(lamm, lam0))))
    m_tot = m_f.value + m_b.value
    mixed = {"forward_ms": round(m_fwd.value, 3), "adj_forward_ms": round(m_f.value, 3),
             "adj_backward_ms": round(m_b.value, 3), "adjoint_total_ms": round(m_tot, 3),
             "cost_vs_mixed_forward": round(m_tot / m_fwd.value, 3),
             "cost_vs_fp64_forward": round(m_tot / fwd.value, 3),
             "J_rel_diff_vs_fp64": abs(Jm - J) / abs(J), "dJ_dinvRe_rel_diff_vs_fp64": abs(thm - dJ_dinvRe) / abs(dJ_dinvRe),
             "grad_rel_l2_vs_fp64": gdiff / gnorm,
             "device_GB": round((s.adj_memory() + base + ns.lib().ns3d_cells(s.ctx) * 20 * 4) / 1e9, 3)}
    s.close()
    total = t_f.value + t_b.value
    return {"config": f"isotropic turbulence {n}^3 FP64, k0=4, Mt=0.3, Re=100, WENO5+SW+C8, {nsteps} RK3 steps, "
                      f"fixed dt={dt:.4g}, checkpoint every {ckpt}, J = sum dt * mean dissipation",
            "J": J, "dJ_dinvRe": dJ_dinvRe, "grad_U0_l2": gnorm,
            "forward_ms": round(fwd.value, 3), "adj_forward_ms": round(t_f.value, 3),
            "adj_backward_ms": round(t_b.value, 3), "adjoint_total_ms": round(total, 3),
            "cost_vs_forward": round(total / fwd.value, 3),
            "one_stream": {"adjoint_total_ms": round(t_one.value, 3),
                           "cost_vs_forward": round(t_one.value / fwd.value, 3),
                           "device_GB": round((one_mem + base) / 1e9, 3),
                           "gradient_bitwise_equal": bool(one_same)},
            "one_stream_profiled": {"adjoint_total_ms": round(t_seq.value, 3),
                                    "cost_vs_forward": round(t_seq.value / fwd.value, 3),
                                    "gradient_bitwise_equal": bool(seq_same)},
            "device_bytes": int(mem + base), "device_GB": round((mem + base) / 1e9, 3),
            "store_all_stages": store_all, "mixed": mixed,
            "kernels_ms": {k: round(v[0], 2) for k, v in sorted(rep.items(), key=lambda kv: -kv[1][0])}}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_oracle(ns, Oracle, n, state, nsteps, dt, **cfg):
    p, g = ns.make_params(n, **cfg)
    o = Oracle(p, g)
    o.set_primitive(*state)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step(dt)       # dt <= 0: the CFL dt, as the GPU run
    t = time.perf_counter() - t0
    o.close()
    return n ** 3 * nsteps / t, t


def cpu_baseline(n, nsteps, Re, Mt, k0, seed):
    """The CPU FP64 oracle on the host's cores, on bounded samples: the
    benchmark's own workload at n^3 (the reported value), and the 64^3
    Taylor-Green vortex and the 128^3 MDCD + Lax-Friedrichs turbulence
    workloads beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ns3d_cpp_amd as ns
    from ns3d_cpp_amd import fields
    from oracle import Oracle, load
    cores = int(load().orc_num_threads())
    st, Ma = fields.hit(n, k0=k0, Mt=Mt, seed=seed)
    v, t = _time_oracle(ns, Oracle, n, st, nsteps, 0.0, Re=Re, Ma=Ma, Pr=0.72, cfl=0.5, recon="WENO5",
                        fvs_type="StagerWarming", vis_scheme="C8th", sutherland_s=110.4 / 273.15)
    del st
    legs = []
    v64, t64 = _time_oracle(ns, Oracle, 64, fields.tgv(64, Ma=0.1), 20, 1e-3, Re=1600.0, Ma=0.1, Pr=0.72,
                            recon="WENO5", fvs_type="StagerWarming", vis_scheme="C8th")
    legs.append({"workload": "Taylor-Green vortex 64^3, WENO5 + Steger-Warming + C8th, dt 1e-3",
                 "value": v64, "steps": 20, "seconds": round(t64, 2)})
    st2, Ma2 = fields.hit(128, k0=4, Mt=0.3, seed=seed)
    v128, t128 = _time_oracle(ns, Oracle, 128, st2, 5, 0.0, Re=100.0, Ma=Ma2, Pr=0.72, cfl=0.5, recon="MDCD",
                              fvs_type="LaxFriedrichs", vis_scheme="C8th", sutherland_s=110.4 / 273.15)
    legs.append({"workload": "isotropic turbulence 128^3, MDCD + Lax-Friedrichs + C8th, CFL 0.5",
                 "value": v128, "steps": 5, "seconds": round(t128, 2)})
    return {"value": v, "unit": "cell-updates/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{nsteps} RK3 step(s) of the same isotropic-turbulence workload at {n}^3 "
                      f"(WENO5+SW+C8, CFL dt) with the C oracle, OpenMP threads = cores; {t:.1f} s",
            "legs": legs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--recon", default="WENO5")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--err-steps", type=int, default=500)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--k0", type=float, default=4.0, help="spectrum peak wavenumber of the field")
    ap.add_argument("--mt", type=float, default=0.3, help="turbulent Mach number of the field")
    ap.add_argument("--re", type=float, default=200.0, help="Reynolds number")
    ap.add_argument("--adj-n", type=int, default=128)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-ckpt", type=int, default=20)
    ap.add_argument("--adj-only", action="store_true", help="only the adjoint leg (for profiling it)")
    ap.add_argument("--no-other-scheme", dest="other_scheme", action="store_false",
                    help="skip the timing of the workload's other scheme (UD7 beside WENO5)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built).", file=sys.stderr)
        sys.exit(2)

    import ns3d_cpp_amd as ns
    from ns3d_cpp_amd import fields

    n, Re, Mt, k0 = args.n, args.re, args.mt, args.k0
    H = hip()
    if args.adj_only:
        adj = adjoint_cost(ns, H, args.adj_n, args.adj_steps, args.adj_ckpt, args.seed)
        print(json.dumps({"adjoint": adj}), flush=True)
        return
    st, Ma = fields.hit(n, k0=k0, Mt=Mt, seed=args.seed)
    # achievable streaming bandwidth on this device: a grid-stride 16-byte
    # copy and a flat nontemporal one over 12 cell arrays
    p, g = make(ns, n, Ma, Re, args.recon, "fp64")
    probe = ns.Solver(p, g)
    probe.profile(True)
    copy_gbps = probe.copy_probe(10)
    prep = probe.profile_report()
    cbytes = 2 * 8 * 12 * ns.lib().ns3d_cells(probe.ctx)
    stream = {"best_GBps": round(copy_gbps, 1)}
    for kname in ("copy", "copy_nt"):
        if kname in prep:
            stream[kname + "_GBps"] = round(cbytes / (prep[kname][0] / prep[kname][1] * 1e-3) / 1e9, 1)
    probe.profile(False)
    probe.close()

    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {prec: time_path(ns, H, n, st, Ma, Re, args.recon, prec, args.steps, args.warmup) for prec in precs}
    # the other scheme of the metric's workload: UD7 (or WENO5 when UD7 ran)
    other = "UPWIND_7TH" if args.recon == "WENO5" else "WENO5"
    legs = {}
    if args.other_scheme:
        leg = {}
        for prec in precs:
            r = time_path(ns, H, n, st, Ma, Re, other, prec, args.steps, args.warmup)
            leg[prec] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "step_model": r["step_model"],
                         "roofline": r["roofline"]}
        if "mixed" in leg and "fp64" in leg:
            leg["speedup_mixed_vs_fp64"] = round(leg["mixed"]["value"] / leg["fp64"]["value"], 3)
        legs[other] = leg
    err = None
    if args.precision == "both" and args.err_steps > 0:
        err = mixed_error(ns, n, st, Ma, Re, args.recon, args.err_steps)
    del st
    adj = None
    if args.adj_steps > 0:
        adj = adjoint_cost(ns, H, args.adj_n, args.adj_steps, args.adj_ckpt, args.seed)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps, Re, Mt, k0, args.seed)

    head = "mixed" if "mixed" in res else "fp64"
    h = res[head]
    line = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"decaying isotropic turbulence {n}^3, k0={k0:g}, Mt={Mt:g}, Re={Re:g}, "
                               f"{args.recon} + Steger-Warming + C8th, Sutherland 110.4/273.15, "
                               f"RK3 with CFL 0.5 dt every step, {head} path",
                   "n": n, "k0": k0, "Mt": Mt, "Re": Re, "recon": args.recon, "fvs": "StagerWarming",
                   "vis": "C8th", "precision": head},
        "roofline": h["roofline"],
        "step_model": h["step_model"],
        "kernel_bytes_per_cell_update": h["kernel_bytes_per_cell_update"],
        "kernel_bytes_hbm_frac": h["kernel_bytes_hbm_frac"],
        "stream_probe": stream,
        "kernels": h["kernels"],
    }
    line["roofline"]["stream_GBps"] = stream["best_GBps"]
    line["roofline"]["frac_of_stream"] = round(line["roofline"]["achieved"] / stream["best_GBps"], 4)
    if legs:
        line["other_scheme"] = legs
    if head == "mixed" and "fp64" in res:
        line["fp64"] = res["fp64"]
        line["speedup_mixed_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    if err is not None:
        line["mixed_error_vs_fp64"] = err
    if adj is not None:
        line["adjoint"] = adj
    line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
