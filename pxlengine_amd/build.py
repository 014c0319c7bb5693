"""Builds the HIP library ``pxlengine_amd/libpxlflip.so`` for gfx950.

Host-only sources (JSON / scene block / seeding) are compiled with g++, HIP sources with
hipcc.  Every source is compiled with -ffp-contract=off: the FLIP kernels reproduce the
reference's fp32 operation order exactly.  HIP code also gets
``-Xclang -target-feature -Xclang -packed-fp32-ops`` (no packed fp32 instructions).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpxlflip.so")
ARCH = os.environ.get("PXL_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["json.cpp", "spec.cpp"]
HIP_SOURCES = ["particles.hip", "seed.hip", "grid.hip", "solver.hip", "mg.hip", "surface.hip", "agents.hip", "flip.hip",
               "server.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIP_FLAGS = [f"-I{INCLUDE}", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fno-gpu-rdc", "-munsafe-fp-atomics"]


def _hipcc() -> str:
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(OUT_DIR, src + ".o")
        jobs.append(["g++", *COMMON, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(OUT_DIR, src + ".o")
        jobs.append([_hipcc(), *COMMON, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([_hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
