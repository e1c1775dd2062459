"""Compiles the HIP lattice library for gfx950, in-tree.

    python -m banjo_amd.build

Floating-point contract (the reference's, cmake/FloatingPointModel.cmake:14-18,
and its no-FMA GPU build, tests/fast_lattice_tests.cpp:474-479): no fast math,
no FMA contraction, IEEE fp32 division and square root, denormals kept. Packed
fp32 instructions are switched off.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBRARY = os.path.join(LIBDIR, "libbanjo_lattice.so")

SOURCES = ["lattice_backend.hip", "lattice_host.cpp", "fragment_geometry.cpp", "fragment_lattice.cpp"]
HEADERS = ["lattice_device.hpp", "lattice_kernels.hpp", "lattice_errors.hpp"]

FP_FLAGS = ["-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-fno-gpu-flush-denormals-to-zero"]
GPU_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIBRARY + ".tmp"
    cmd = [hipcc(), "-std=c++20", "-O3", "-fPIC", "-shared", *GPU_FLAGS, *FP_FLAGS,
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", tmp]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building the lattice library")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
