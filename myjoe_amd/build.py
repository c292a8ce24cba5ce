"""Builds libjoemi.so in-tree: host C++ with g++, device code with hipcc for
gfx950, both without FMA contraction (-ffp-contract=off) so the FP64
residual path keeps the reference's rounding.

    python -m myjoe_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libjoemi.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SRC = ["jm_mesh.cpp", "jm_capi.cpp", "jm_server.cpp"]
DEV_SRC = ["jm_solver.hip", "jm_chemtable.hip"]

HOST_FLAGS = ["-O2", "-ffp-contract=off", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]
DEV_FLAGS = ["-O3", "-ffp-contract=off", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HOST_SRC:
        o = os.path.join(BUILD, s + ".o")
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o], verbose)
        objs.append(o)
    for s in DEV_SRC:
        o = os.path.join(BUILD, s + ".o")
        _run([hipcc, *DEV_FLAGS, "-c", os.path.join(CSRC, s), "-o", o], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> None:
    """The C++ JoeWithModels mirror (libjoemi_host.a, host code over the C
    ABI only) and the example case program bin/joe_cases linked against it."""
    flags = ["-O2", "-std=c++17", "-Wall", "-Wno-pmf-conversions", "-fPIC"]
    objs = []
    for src in ("UgpWithCvCompFlow", "JoeWithModels"):
        o = os.path.join(BUILD, src + ".o")
        _run(["g++", *flags, "-c", os.path.join(HERE, "host", src + ".cpp"), "-o", o], verbose)
        objs.append(o)
    ar = os.path.join(HERE, "libjoemi_host.a")
    if os.path.exists(ar):
        os.remove(ar)
    _run(["ar", "rcs", ar, *objs], verbose)
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    exe = os.path.join(HERE, "bin", "joe_cases")
    _run(["g++", *flags, os.path.join(HERE, "host", "cases", "joe_cases.cpp"), "-o", exe + ".tmp", ar,
          f"-L{HERE}", "-ljoemi", "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose=True))
