"""CPU check of the quotient steps the FP64 sweeps use for WENO5
(rcp_refined / quot_rcp / recon_weno5_rcp in ns3d_device.hpp): restated in
C with fma() (tests/division_steps_check.c), they must give C's correctly
rounded division bit for bit on random quotients inside the range the device
comment states, WENO5 faces equal to the literal reconstruction.cpp:93-124 on
stencils of zeros and magnitudes in [2^-900, 2^200] (a zero's sign aside),
and the 1/6 quotients of integer combinations.  The same program built with
the last correction step left out must report mismatches.  The device's own
bits are compared on the GPU (tests/test_gpu_parity.py)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "division_steps_check.c")


def _build_run(tmp_path, extra):
    exe = str(tmp_path / ("divchk" + "".join(extra).replace("-", "_")))
    subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", *extra, SRC, "-o", exe, "-lm"], check=True)
    return subprocess.run([exe], capture_output=True, text=True, timeout=300)


def test_division_steps_equal_ieee_division(tmp_path):
    r = _build_run(tmp_path, [])
    assert r.returncode == 0, r.stdout + r.stderr
    q = re.search(r"quotients (\d+): (\d+) differ", r.stdout)
    w = re.search(r"weno5 stencils (\d+): (\d+) differ, (\d+) differ only in a zero's sign", r.stdout)
    p = re.search(r"sixths (\d+): (\d+) differ", r.stdout)
    assert q and w and p, r.stdout
    assert int(q.group(1)) > 5_000_000 and int(q.group(2)) == 0
    assert int(w.group(1)) == 3_000_000 and int(w.group(2)) == 0
    assert int(w.group(3)) < 100          # -0 numerators only
    assert int(p.group(1)) > 2_000_000 and int(p.group(2)) == 0


def test_planted_missing_correction_is_caught(tmp_path):
    r = _build_run(tmp_path, ["-DPLANT_NO_CORRECTION"])
    assert r.returncode != 0, r.stdout
    q = re.search(r"quotients (\d+): (\d+) differ", r.stdout)
    assert q and int(q.group(2)) > 0
