"""Builds the native library libdfttest2_mi355x.so in-tree for gfx950.

    python vs-dfttest2_amd/build_native.py

Compiles csrc/dft2_kernels.hip (device + launch code), csrc/dft2_capi.cpp and
csrc/dft2_batch.cpp (C ABI) with hipcc and links them into
vs-dfttest2_amd/dfttest2_amd/_native/.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "dfttest2_amd", "_native")
LIB_NAME = "libdfttest2_mi355x.so"
ARCH = "gfx950"
# Kernels and launch code; the C ABI (plans, batches).
SOURCES = ("dft2_kernels.hip", "dft2_capi.cpp", "dft2_batch.cpp")

# Packed FP32 instructions are kept out of every gfx950 code object.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Only the C ABI (DFT2_API in include/dfttest2_mi355x.h) is exported from the library.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wno-unused-result"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    hc = hipcc()
    objs = []
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        run([hc, "-c", "-x", "hip", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS,
             os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + ".tmp"
    run([hc, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
