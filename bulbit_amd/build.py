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

SOURCES = ["host.cpp", "tables.cpp", "kernels.hip", "nrc.hip"]
HEADERS = ["bb_math.h", "bb_scene.h", "bb_shade.h", "bb_texcoord.h", os.path.join("..", "..", "include", "bbgpu.h")]

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


CPP = os.path.join(HERE, "cpp")
CPP_LIB = os.path.join(HERE, "libbulbit.so")
CPP_EXAMPLE = os.path.join(CPP, "_build", "readme_example")
CPP_WRITE_IMAGES = os.path.join(CPP, "_build", "write_images")
CPP_TEXTURED = os.path.join(CPP, "_build", "textured_example")
CPP_LOAD_OBJ = os.path.join(CPP, "_build", "load_obj")
CPP_CONVERT_IMAGE = os.path.join(CPP, "_build", "convert_image")
CPP_DESCRIBE_SCENE = os.path.join(CPP, "_build", "describe_scene")
CPP_RENDER_SCENE = os.path.join(CPP, "_build", "render_scene")
# host C++ API (bulbit/bulbit.h) over the C ABI: plain g++, same FP rules
CXX_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-pthread"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{what} failed")


def build(verbose: bool = False) -> str:
    """Compiles libbbgpu.so, then the C++ API libbulbit.so and its README
    example (always recompiles)."""
    tmp = LIB + ".tmp"
    _run([hipcc(), *FLAGS, "-I", CSRC, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES],
         "hipcc building libbbgpu.so", verbose)
    os.replace(tmp, LIB)
    build_cpp(verbose)
    return LIB


def build_cpp(verbose: bool = False) -> str:
    cxx = os.environ.get("CXX", "g++")
    inc = os.path.join(CPP, "include")
    tmp = CPP_LIB + ".tmp"
    _run([cxx, *CXX_FLAGS, "-fPIC", "-shared", "-I", inc, os.path.join(CPP, "bulbit.cpp"),
          os.path.join(CPP, "image_io.cpp"), os.path.join(CPP, "jpeg_decode.cpp"), os.path.join(CPP, "model_loader.cpp"), os.path.join(CPP, "scene_loader.cpp"), "-o", tmp,
          "-L", HERE, "-lbbgpu", "-lz", "-Wl,-rpath,$ORIGIN"], "g++ building libbulbit.so", verbose)
    os.replace(tmp, CPP_LIB)
    os.makedirs(os.path.dirname(CPP_EXAMPLE), exist_ok=True)
    _run([cxx, *CXX_FLAGS, "-I", inc, os.path.join(CPP, "examples", "readme.cpp"), "-o", CPP_EXAMPLE,
          "-L", HERE, "-lbulbit", "-lbbgpu", "-Wl,-rpath,$ORIGIN/../.."], "g++ building the README example", verbose)
    _run([cxx, *CXX_FLAGS, "-I", inc, os.path.join(CPP, "examples", "write_images.cpp"), "-o", CPP_WRITE_IMAGES,
          "-L", HERE, "-lbulbit", "-lbbgpu", "-Wl,-rpath,$ORIGIN/../.."], "g++ building write_images", verbose)
    _run([cxx, *CXX_FLAGS, "-I", inc, os.path.join(CPP, "examples", "textured.cpp"), "-o", CPP_TEXTURED,
          "-L", HERE, "-lbulbit", "-lbbgpu", "-Wl,-rpath,$ORIGIN/../.."], "g++ building the textured example", verbose)
    _run([cxx, *CXX_FLAGS, "-I", inc, os.path.join(CPP, "examples", "load_obj.cpp"), "-o", CPP_LOAD_OBJ,
          "-L", HERE, "-lbulbit", "-lbbgpu", "-Wl,-rpath,$ORIGIN/../.."], "g++ building load_obj", verbose)
    _run([cxx, *CXX_FLAGS, "-I", inc, os.path.join(CPP, "examples", "convert_image.cpp"), "-o", CPP_CONVERT_IMAGE,
          "-L", HERE, "-lbulbit", "-lbbgpu", "-Wl,-rpath,$ORIGIN/../.."], "g++ building convert_image", verbose)
    for src, exe in (("describe_scene.cpp", CPP_DESCRIBE_SCENE), ("render_scene.cpp", CPP_RENDER_SCENE)):
        _run([cxx, *CXX_FLAGS, "-I", inc, os.path.join(CPP, "examples", src), "-o", exe,
              "-L", HERE, "-lbulbit", "-lbbgpu", "-Wl,-rpath,$ORIGIN/../.."], "g++ building " + src, verbose)
    return CPP_LIB


if __name__ == "__main__":
    print(build(verbose=True))
