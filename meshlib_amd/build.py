"""Build recipe of the native library ``meshlib_amd/libmeshlib_amd.so`` (gfx950).

``python -m meshlib_amd.build`` compiles the host tree builder, the HIP kernels and the
C-ABI into one shared library in-tree.  Flags:

* ``-ffp-contract=off``: no fused multiply-add contraction, so every float/double
  expression rounds exactly like the reference CPU build (bit-identical distances);
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division and square root;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmeshlib_amd.so")
ARCH = os.environ.get("MESHLIB_AMD_ARCH", "gfx950")

SOURCES = ["mi_tree.cpp", "mi_tree_gpu.hip", "mi_kernels.hip", "mi_capi.cpp", "mi_mlp.hip", "mi_mlp_capi.cpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


ADAPTER_SRC = os.path.join(ROOT, "tests", "cpp", "adapter_main.cpp")
ADAPTER_BIN = os.path.join(ROOT, "tests", "cpp", "adapter_main")


def build_adapter_driver(verbose: bool = False) -> str:
    """Host-only C++ program driving the drop-in classes of MRAmdDistanceField.h (linked against
    the library built above; runs on the GPU box from the GPU tests)."""
    cxx = shutil.which("g++") or shutil.which("c++")
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-DMESHLIB_AMD_USE_MIRROR", "-I", os.path.join(ROOT, "include"),
           ADAPTER_SRC, "-o", ADAPTER_BIN, "-L", HERE, "-lmeshlib_amd",
           "-Wl,-rpath,$ORIGIN/../../meshlib_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return ADAPTER_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_adapter_driver(verbose=True))
