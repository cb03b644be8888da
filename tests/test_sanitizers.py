"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer (no
GPU): the oracle (oracle/ns3d_oracle.c, with the C ABI's parameter code
ns3d_params.cpp) driven over its whole option surface by
tests/oracle_sanitize_check.cpp.  The host file formats
(host/ns3d_io.cpp) run their whole test module under the same sanitizers:
test_host_io.py builds its harness both plainly and sanitized.  Any report
makes the program exit non-zero (-fno-sanitize-recover=all)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
           "OMP_NUM_THREADS": "2"}


def test_oracle_under_asan_ubsan(tmp_path):
    orc = str(tmp_path / "orc.o")
    par = str(tmp_path / "par.o")
    chk = str(tmp_path / "chk.o")
    exe = str(tmp_path / "oracle_sanitize_check")
    subprocess.run(["gcc", *SAN, "-std=c99", "-fopenmp", "-ffp-contract=off", "-c",
                    os.path.join(ROOT, "oracle", "ns3d_oracle.c"), "-o", orc], check=True)
    subprocess.run(["g++", *SAN, "-std=c++17", "-c", os.path.join(ROOT, "ns3d_cpp_amd", "csrc", "ns3d_params.cpp"),
                    "-o", par], check=True)
    subprocess.run(["g++", *SAN, "-std=c++17", "-c", os.path.join(ROOT, "tests", "oracle_sanitize_check.cpp"),
                    "-o", chk], check=True)
    subprocess.run(["g++", *SAN, "-fopenmp", orc, par, chk, "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, env={**os.environ, **SAN_ENV}, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle ok: 36 configurations, 0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
