"""bench.py's contract parts that need no GPU: more than one GPU is refused
with a message, and the per-kernel byte model adds up to what its comments
state (the mixed stage kernel averages its three stages' 140 / 160 / 200 B,
160 / 180 / 220 B with the z sweep folded in; the mixed step has no CFL pass
of its own)."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_byte_model():
    b = load_bench()
    fp, mx = b.ALG_BYTES["fp64"], b.ALG_BYTES["mixed"]
    # mixed stage kernel: reads 9 fluxes (36), A (20), U (40), dU (20 in stages 2, 3);
    # writes dU (20), the shadow (24), U (40 in stage 3)
    stages = [36 + 20 + 40 + 20 + 24, 36 + 20 + 40 + 20 + 20 + 24, 36 + 20 + 40 + 20 + 20 + 24 + 40]
    assert stages == [140, 160, 200]
    assert abs(mx["vdiv_stage"] - sum(stages) / 3) < 1e-12
    # one RHS + stage per RK3 stage; the CFL pass once per step in FP64, and
    # not at all in mixed mode (its third stage reduces the next minimum)
    step = lambda a, ncfl: 3 * sum(v for k, v in a.items() if k != "cfl") + a["cfl"] * ncfl
    assert step(fp, 1) == 3 * (80 + 120 + 120 + 104 + 240) + 40
    sep = {k: v for k, v in mx.items() if k != "vdiv_zs_stage"}
    assert abs(step(sep, 0) - (3 * (40 + 60 + 60 + 52) + 500)) < 1e-9
    # the z sweep folded into the stage kernel: that kernel also reads the
    # shadow (20 B), and the sweep's accumulator read and write (40 B) are gone
    assert abs(mx["vdiv_zs_stage"] - (sum(stages) / 3 + 20)) < 1e-12
    fused = {k: v for k, v in mx.items() if k not in ("vdiv_stage", "sweep_z")}
    assert abs(step(fused, 0) - (3 * (40 + 60 + 52) + 500 + 60)) < 1e-9
    # the metric's model keeps the CFL read in both (2408 + 40, 1544 + 20)
    assert b.STEP_MODEL_BYTES == {"fp64": 2448, "mixed": 1564}


def test_traffic_record_is_keyed_by_size_and_scheme():
    """The roofline's measured traffic comes from the profile of the same
    size and scheme, and is null where none was recorded."""
    b = load_bench()
    assert isinstance(b.traffic_from_profiles("mixed", "vdiv_zs_stage", 256), int)
    assert isinstance(b.traffic_from_profiles("fp64", "vdiv_stage", 256), int)
    assert b.traffic_from_profiles("mixed", "vdiv_zs_stage", 256, "UPWIND_7TH") is None
    assert b.traffic_from_profiles("mixed", "vdiv_zs_stage", 96) is None


def test_prof_kinds_and_stages():
    sys.path.insert(0, ROOT)
    from ns3d_cpp_amd import prof
    assert prof.kind_of("void ns3d::km_vdiv_z<0, 2, 0>(ns3d::DevParams, ns3d::MFields, double)") == "vdiv_stage"
    assert prof.stage_of("void ns3d::km_vdiv_z<0, 2, 0>(ns3d::DevParams)") == 2
    assert prof.stage_of("void ns3d::km_vdiv_z<0, 0, 0>(ns3d::DevParams)") is None
    assert prof.kind_of("ns3d::k_copy_nt(double2 const*, double2*, long)") == "copy_nt"
    assert prof.kind_of("void ns3d::km_vdiv_zs<0, 3, 0, 0>(ns3d::DevParams, ns3d::MFields)") == "vdiv_zs_stage"
    assert prof.stage_of("void ns3d::km_vdiv_zs<0, 3, 0, 0>(ns3d::DevParams, ns3d::MFields)") == 3
    assert prof.kind_of("void ns3d::km_vdiv_lds<0, 1, true>(ns3d::DevParams)") == "vdiv_stage"
    assert prof.stage_of("void ns3d::km_vdiv_lds<0, 1, true>(ns3d::DevParams)") == 1
    assert prof.kind_of("ns3d::k_copy(double2 const*, double2*, long)") == "copy"
