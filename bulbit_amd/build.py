"""Builds libbbgpu.so in-tree with hipcc for gfx950.

Flags that matter for parity with the reference:
  -ffp-contract=off                          no fused multiply-add unless written as fmaf()
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and sqrt in kernels
  no -ffast-math, no denormal flushing
and, for gfx950, packed FP32 instructions are disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbbgpu.so")
ARCH = os.environ.get("BBGPU_ARCH", "gfx950")

SOURCES = ["host.cpp", "kernels.hip"]
HEADERS = ["bb_math.h", "bb_scene.h", "bb_shade.h", os.path.join("..", "..", "include", "bbgpu.h")]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compiles every source into libbbgpu.so (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, "-I", CSRC, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libbbgpu.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
