"""Builds libadseismic_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: ``python adseismic.jl_amd/build.py`` (or ``__graft_entry__.build()``).  Every HIP
translation unit is compiled with packed-FP32 instruction selection disabled; ads_ops.hip
(the bit-exact drop-in operators) and ads_acoustic.hip (whose single- and two-level stencil
passes must agree bit for bit) are additionally compiled without FMA contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIBNAME = "libadseismic_hip.so"
ARCH = os.environ.get("ADS_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]
SOURCES = {
    "ads_host.cpp": ["-x", "hip"],
    # no FMA contraction: the rectangle, two-level rectangle and frame kernels evaluate the
    # same stencil expression in different kernels and must round identically
    "ads_acoustic.hip": ["-ffp-contract=off"],
    "ads_elastic.hip": [],
    "ads_ops.hip": ["-ffp-contract=off"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build the HIP library)")


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    cc = hipcc()
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = []
    for src, extra in SOURCES.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJDIR, src + ".o")
        jobs.append((path, obj, [cc, *COMMON, *extra, "-c", path, "-o", obj]))

    def run(job):
        path, obj, cmd = job
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {os.path.basename(path)}:\n{r.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        objs = list(ex.map(run, jobs))
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
