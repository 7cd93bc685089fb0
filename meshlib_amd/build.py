"""Build recipe of the native library ``meshlib_amd/libmeshlib_amd.so`` (gfx950).

``python -m meshlib_amd.build`` compiles the host tree builder, the HIP kernels and the
C-ABI into one shared library in-tree: every source to its own object (in parallel, under
``meshlib_amd/_build/``), then one link.  Flags:

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
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmeshlib_amd.so")
ARCH = os.environ.get("MESHLIB_AMD_ARCH", "gfx950")

SOURCES = ["mi_tree.cpp", "mi_tree_gpu.hip", "mi_kernels.hip", "mi_kernels_keyed.hip", "mi_capi.cpp",
           "mi_mlp.hip", "mi_mlp_capi.cpp", "mi_noalloc.hip"]

FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ, src + ".o")
    cmd = [hipcc(), *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
           os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = max(1, min(len(SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
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
