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
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("kernels.hip", "runtime.cpp", "server.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *SOURCES, "-I", os.path.join(ROOT, "include")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
