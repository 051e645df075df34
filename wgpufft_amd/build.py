"""Builds libmifft.so (HIP kernels + C++ host runtime) for gfx950 in-tree.

    python -m wgpufft_amd.build          # or __graft_entry__.build()

Each source compiles to an object under wgpufft_amd/_build/ in parallel, then
hipcc links wgpufft_amd/libmifft.so.  hipcc cross-compiles for gfx950 without a
GPU present.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmifft.so")
SOURCES = ["kern_line.hip", "kern_conv.hip", "kern_regconv.hip", "kern_util.hip", "kern_batch.hip", "kern_small.hip", "scratch.hip", "plan.cpp",
           "api.cpp", "diagnostics.cpp", "batch.cpp", "serve.cpp", "cpu_plan.cpp", "cpu_fft_avx2.cpp", "cpu_fft_base.cpp"]
# Host-only translation units of the CPU FFT engine: plain C++ (no device
# code), one built for AVX2 + FMA and one for the x86-64 baseline; the library
# picks one at run time from what the CPU reports.
HOST_ONLY = {"cpu_fft_avx2.cpp": ["-mavx2", "-mfma"], "cpu_fft_base.cpp": []}
ARCH = os.environ.get("MIFFT_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
# Packed-FP32 instructions are disabled for every HIP translation unit.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src in HOST_ONLY:
        cmd = [CXX, "-O3", "-std=c++17", "-fPIC", "-Wall", *HOST_ONLY[src], "-c", os.path.join(CSRC, src),
               "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return obj
    cmd = [HIPCC, *COMMON]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
