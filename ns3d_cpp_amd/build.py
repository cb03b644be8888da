"""Build the GPU library (libns3d_gpu.so, gfx950) and the host driver.

Everything is compiled in-tree so the shared objects travel with the source
tree.  Run as ``python -m ns3d_cpp_amd.build``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libns3d_gpu.so")
DRIVER = os.path.join(HERE, "ns3d_gpu_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: the FP64 path forms no FMA (the reference's x86-64
# baseline build has none).  Packed FP32 ops are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
SOURCES = ["ns3d_kernels.hip", "ns3d_mixed.hip", "ns3d_adjoint.hip", "ns3d_monitors.hip", "ns3d_noalloc.hip",
           "ns3d_api.cpp", "ns3d_params.cpp"]
# the mixed-precision path and the derivative kernels do not reproduce the
# reference's rounding, so their translation units contract a*b+c into FMAs
# within an expression (-ffp-contract=on).  Not "fast": that lets the backend
# fuse across statements wherever the schedule allows, so one inline function
# can round differently at two call sites, and the sweeps difference split
# values computed at different call sites.
FMA_SOURCES = {"ns3d_mixed.hip", "ns3d_adjoint.hip"}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    return r


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        flags = list(HIP_FLAGS)
        if src in FMA_SOURCES:
            flags[flags.index("-ffp-contract=off")] = "-ffp-contract=on"
        jobs.append([HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB + ".tmp"
    # --no-undefined: a declaration without a definition fails here, not at load time
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs, "-lhipfft"],
         verbose)
    os.replace(tmp, LIB)
    drv_src = os.path.join(HOST, "ns3d_gpu_run.cpp")
    if os.path.exists(drv_src):
        _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HOST,
              drv_src, os.path.join(HOST, "ns3d_field.cpp"), os.path.join(HOST, "ns3d_io.cpp"), "-o", DRIVER,
              f"-L{HERE}", "-lns3d_gpu", f"-Wl,-rpath,{HERE}"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
