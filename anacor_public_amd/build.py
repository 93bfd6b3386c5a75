"""Build libanacor_mi355x.so in-tree (gfx950).

    python -m anacor_public_amd.build

Host code (csrc/geometry.cpp, csrc/risk_fit.cpp) is compiled with g++ against
glibc libm and without FMA contraction, so the geometry's FP64 results carry
the reference's bits; the
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
HOST_SOURCES = ("geometry.cpp", "risk_fit.cpp")       # FP64 geometry; risk-tree trainer


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build for the CPU
# tests: the host code of every source (geometry.cpp with the ROCm clang, the
# host half of the .hip files through -Xarch_host) is instrumented, the GPU code
# is not.  Run with the clang runtime preloaded, e.g.
#   LD_PRELOAD=$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)
#   ASAN_OPTIONS=detect_leaks=0 ANACOR_LIB=anacor_public_amd/_sanitize/libanacor_mi355x.so
#   python -m pytest tests -m "not gpu"
SAN_BUILD = os.path.join(PKG, "_sanitize")
SAN_CXX = os.environ.get("SAN_CXX", "/opt/rocm/llvm/bin/clang++")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-fno-omit-frame-pointer", "-g"]


def build(verbose: bool = True, sanitize: bool = False) -> str:
    out_dir = SAN_BUILD if sanitize else BUILD
    lib_dir = SAN_BUILD if sanitize else LIBDIR
    os.makedirs(out_dir, exist_ok=True)
    os.makedirs(lib_dir, exist_ok=True)
    lib = os.path.join(lib_dir, "libanacor_mi355x.so")
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        cxx = [SAN_CXX, *HOST_FLAGS, *SAN] if sanitize else [CXX, *HOST_FLAGS]
        _run([*cxx, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    host_san = [a for f in SAN for a in ("-Xarch_host", f)] if sanitize else []
    for src in HIP_SOURCES:
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, *host_san, "-I", os.path.join(ROOT, "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.insert(0, obj)
    tmp = lib + ".tmp"
    # -shared-libsan: the sanitizer runtime becomes a dependency of the
    # library, so a caller that resolves free through the library (as
    # utils/utils_mp.py:212 does) gets the runtime's free, matching its malloc
    link_san = [*host_san, "-shared-libsan"] if sanitize else []
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *link_san, "-o", tmp, *objs,
          "-lpthread"])
    os.replace(tmp, lib)
    if not sanitize:
        for alias in ALIASES:
            dst = os.path.join(LIBDIR, alias)
            shutil.copyfile(LIB, dst + ".tmp")
            os.replace(dst + ".tmp", dst)
    return lib


if __name__ == "__main__":
    build(sanitize="--sanitize" in sys.argv[1:])
    sys.exit(0)
