"""Build the C-ABI library euc_amd/libeuchip.so for gfx950 with hipcc.

Flags that exactness depends on:
  -ffp-contract=off                          no FMA contraction (each f32 op rounds once)
  -fno-fast-math                             no reassociation / approximate math
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE f32 division and sqrt
  (no -fgpu-flush-denormals-to-zero)         subnormals preserved, as on the CPU
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libeuchip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("kernels.hip", "runtime.cpp", "server.cpp", "noalloc.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


CPP_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "pipeline_check.cpp")
CPP_CHECK = os.path.join(ROOT, "tests", "cpp", "pipeline_check")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *SOURCES, "-I", os.path.join(ROOT, "include")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_cpp_check(verbose)
    return LIB


def build_cpp_check(verbose: bool = False) -> str:
    """The C++ host-mirror program the GPU tests run (g++ against libeuchip.so)."""
    tmp = CPP_CHECK + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), CPP_CHECK_SRC, "-o", tmp,
           "-L", HERE, "-leuchip", "-Wl,-rpath,$ORIGIN/../../euc_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, CPP_CHECK)
    return CPP_CHECK


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
