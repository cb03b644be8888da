"""Throughput of the RK3 RHS path on one MI355X.

Workload (default): decaying compressible isotropic turbulence on a 256^3
periodic box [0, 2pi)^3 -- seeded solenoidal random-phase velocity with
E(k) = 16 sqrt(2/pi) (u0^2/k0) (k/k0)^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 1,
rho = 1, T = 1, Mt = 0.3, Re = 200, Pr = 0.72, Sutherland mu(T) with
S/T_ref = 110.4/273.15; WENO5 + Steger-Warming + C8 central viscous terms;
SSP-RK3 with the CFL (0.5) time step recomputed every step.

A step = one RK3 step (3 RHS + 3 stage updates + the CFL reduction).
Metric: cell-updates/s = N^3 / (seconds per step).  Inputs are resident in
HBM before the timed region.  Per-kernel durations come from HIP events
recorded around every launch on the solver's stream during the timed region.
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
# Algorithmic HBM bytes per interior cell per launch of each FP64 kernel
# (every array touched once per launch; halos reused on chip).
ALG_BYTES_PER_CELL = {
    "sweep_x": 80,        # read rho,u,v,w,p; write A(5)
    "sweep_y": 120,       # + read A(5)
    "sweep_z": 120,
    "grad": 128,          # read u,v,w,T; write 12 gradients
    "vflux": 224,         # read 12 gradients + T,u,v,w; write 12 fluxes
    "vdiv_stage": 264,    # read 12 fluxes, A(5), W(5), U0(5); write W(6)
    "cfl": 40,            # read rho,u,v,w,p
}
# Per cell-update model of the full step (3 RHS + 3 stages + CFL read).
STEP_MODEL_BYTES_FP64 = 2448


def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    lib.hipEventSynchronize.argtypes = [C.c_void_p]
    lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def cpu_baseline(n, nsteps, Re, Mt, k0, seed):
    """The CPU FP64 oracle on the host's cores, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ns3d_cpp_amd as ns
    from ns3d_cpp_amd import fields
    from oracle import Oracle, load
    st, Ma = fields.hit(n, k0=k0, Mt=Mt, seed=seed)
    p, g = ns.make_params(n, Re=Re, Ma=Ma, Pr=0.72, cfl=0.5, recon="WENO5", fvs_type="StagerWarming",
                          vis_scheme="C8th", sutherland_s=110.4 / 273.15)
    o = Oracle(p, g)
    o.set_primitive(*st)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step(0.0)      # CFL dt, as the GPU run
    t = time.perf_counter() - t0
    o.close()
    return {"value": n ** 3 * nsteps / t, "unit": "cell-updates/s", "cores": int(load().orc_num_threads()),
            "kind": "port",
            "sample": f"{nsteps} RK3 step(s) of the same isotropic-turbulence workload at {n}^3 "
                      f"(WENO5+SW+C8, CFL dt) with the C oracle, OpenMP threads = cores; {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--recon", default="WENO5")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built).", file=sys.stderr)
        sys.exit(2)

    import ns3d_cpp_amd as ns
    from ns3d_cpp_amd import fields

    n, Re, Mt, k0 = args.n, 200.0, 0.3, 4.0
    st, Ma = fields.hit(n, k0=k0, Mt=Mt, seed=args.seed)
    p, g = ns.make_params(n, Re=Re, Ma=Ma, Pr=0.72, cfl=0.5, recon=args.recon, fvs_type="StagerWarming",
                          vis_scheme="C8th", sutherland_s=110.4 / 273.15)
    s = ns.Solver(p, g)
    s.upload_primitive(*st)
    H = hip()
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    H.hipEventCreate(C.byref(ev0))
    H.hipEventCreate(C.byref(ev1))

    del st
    # achievable streaming bandwidth on this device (writes gradient and
    # viscous-flux scratch only, which every RHS recomputes)
    copy_gbps = s.copy_probe(10)
    for _ in range(args.warmup):
        s.rk3_step(s.timestep())
    s.sync()
    H.hipDeviceSynchronize()

    s.profile(True)
    t0 = time.perf_counter()
    H.hipEventRecord(ev0, s.stream())
    for _ in range(args.steps):
        s.rk3_step(s.timestep())
    H.hipEventRecord(ev1, s.stream())
    H.hipEventSynchronize(ev1)
    s.sync()
    H.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    ms = C.c_float()
    H.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    rep = s.profile_report()
    s.profile(False)
    ek = s.total_energy()

    ms_per_step = ms.value / args.steps
    value = n ** 3 / (ms_per_step * 1e-3)
    # dominant kernel by total time
    dom = max(rep, key=lambda k: rep[k][0])
    dom_ms, dom_n = rep[dom]
    dom_avg = dom_ms / dom_n
    bytes_per_launch = ALG_BYTES_PER_CELL.get(dom, 0) * n ** 3
    achieved = bytes_per_launch / (dom_avg * 1e-3) / 1e9
    kernels = {k: {"avg_ms": round(v[0] / v[1], 4), "launches": v[1],
                   "share": round(v[0] / ms.value, 4)} for k, v in sorted(rep.items(), key=lambda kv: -kv[1][0])}
    for k in kernels:
        if k in ALG_BYTES_PER_CELL:
            kernels[k]["GBps"] = round(ALG_BYTES_PER_CELL[k] * n ** 3 / (kernels[k]["avg_ms"] * 1e-3) / 1e9, 1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps, Re, Mt, k0, args.seed)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"decaying isotropic turbulence {n}^3, k0=4, Mt=0.3, Re=200, "
                               f"{args.recon} + Steger-Warming + C8th, Sutherland 110.4/273.15, "
                               "RK3 with CFL 0.5 dt every step, FP64",
                   "n": n, "recon": args.recon, "fvs": "StagerWarming", "vis": "C8th", "precision": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "step_model": {"bytes_per_cell_update": STEP_MODEL_BYTES_FP64,
                       "hbm_frac": round(STEP_MODEL_BYTES_FP64 * value / (HBM_PEAK_GBS * 1e9), 4)},
        "copy_probe_GBps": round(copy_gbps, 1),
        "kernels": kernels,
        "wall_s_timed": round(wall, 4),
        "kinetic_energy_end": ek,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
