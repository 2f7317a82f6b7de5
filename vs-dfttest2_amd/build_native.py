"""Builds the native library libdfttest2_mi355x.so in-tree for gfx950.

    python vs-dfttest2_amd/build_native.py

Compiles csrc/dft2_kernels.hip (device + launch code) and csrc/dft2_capi.cpp (C ABI)
with hipcc and links them into vs-dfttest2_amd/dfttest2_amd/_native/.
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

# Packed FP32 instructions are kept out of every gfx950 code object.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


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
    k_obj = os.path.join(obj_dir, "dft2_kernels.o")
    c_obj = os.path.join(obj_dir, "dft2_capi.o")
    run([hc, "-c", "-x", "hip", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS,
         os.path.join(CSRC, "dft2_kernels.hip"), "-o", k_obj])
    run([hc, "-c", "-x", "hip", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS,
         os.path.join(CSRC, "dft2_capi.cpp"), "-o", c_obj])
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + ".tmp"
    run([hc, "-shared", f"--offload-arch={ARCH}", k_obj, c_obj, "-o", tmp])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
