"""Build libanacor_mi355x.so in-tree (gfx950).

    python -m anacor_public_amd.build

Host geometry (csrc/geometry.cpp) is compiled with g++ against glibc libm and
without FMA contraction, so its FP64 results carry the reference's bits; the
HIP kernels and the C-ABI (csrc/trace.hip, csrc/mlp.hip) are compiled with hipcc for gfx950,
also with -ffp-contract=off (bit-exact stepping) and packed-FP32 code
generation disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libanacor_mi355x.so")
# The same library under the names the reference's drivers load
# (utils/utils_mp.py:51-55 and :119-125 with --gpu-card mi355x).
ALIASES = ("ray_tracing_cpu.so", "ray_tracing_gpu_mi355x.so")

ARCH = os.environ.get("ANACOR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
              "-Wextra"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]


HIP_SOURCES = ("trace.hip", "mlp.hip", "probe.hip")   # tracer + C-ABI; surrogate; gather probe


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    geo_o = os.path.join(BUILD, "geometry.o")
    _run([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, "geometry.cpp"), "-o", geo_o])
    objs = [geo_o]
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.insert(0, obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    for alias in ALIASES:
        dst = os.path.join(LIBDIR, alias)
        shutil.copyfile(LIB, dst + ".tmp")
        os.replace(dst + ".tmp", dst)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
