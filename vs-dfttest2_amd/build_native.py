"""Builds the native library libdfttest2_mi355x.so in-tree for gfx950.

    python vs-dfttest2_amd/build_native.py

Compiles the gfx950 kernels (csrc/dft2_fused.hip once per radius, dft2_backward.hip,
dft2_kernels.hip) and the C ABI (dft2_capi.cpp, dft2_batch.cpp, dft2_noalloc.cpp) with
hipcc, in parallel, and links them into vs-dfttest2_amd/dfttest2_amd/_native/.  Where the
VapourSynth headers are found, it also builds the VapourSynth plugin
(vsplugin/dfttest2_mi355x_vs.cpp) beside it; otherwise it says why not.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "dfttest2_amd", "_native")
LIB_NAME = "libdfttest2_mi355x.so"
ARCH = "gfx950"
# (source, extra flags, object): the fused kernel is compiled once per radius so that its
# instantiations build in parallel; then the backward, finalize/launch code and C ABI.
SOURCES = tuple(("dft2_fused.hip", [f"-DDFT2_FUSED_RADIUS={r}"], f"dft2_fused_r{r}.o") for r in range(4)) + (
    ("dft2_backward.hip", [], "dft2_backward.o"),
    ("dft2_kernels.hip", [], "dft2_kernels.o"),
    ("dft2_generic.hip", [], "dft2_generic.o"),
    ("dft2_capi.cpp", [], "dft2_capi.o"),
    ("dft2_batch.cpp", [], "dft2_batch.o"),
    ("dft2_noalloc.cpp", [], "dft2_noalloc.o"),
)

# Packed FP32 instructions are kept out of every gfx950 code object.  The sign of a zero
# never reaches an output (integers are rounded, float results are compared by value), so
# x + 0 may be folded: the middle stage's terms that vanish for the temporal DC row then
# cost no instructions.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-signed-zeros"]
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


VS_PLUGIN_SRC = os.path.join(HERE, "vsplugin", "dfttest2_mi355x_vs.cpp")
VS_PLUGIN_NAME = "libdfttest2_mi355x_vs.so"


def find_vapoursynth_include():
    """Directory holding VapourSynth4.h: $VS_INCLUDE, then the usual include paths."""
    cands = [os.environ.get("VS_INCLUDE", "")] + [os.path.join(r, s) for r in ("/usr/include", "/usr/local/include")
                                                  for s in ("vapoursynth", "")]
    for d in cands:
        if d and os.path.exists(os.path.join(d, "VapourSynth4.h")) and os.path.exists(os.path.join(d, "VSHelper4.h")):
            return d
    return None


def build_vs_plugin(lib_path: str):
    """The VapourSynth plugin over the C ABI, where the VapourSynth headers are present."""
    inc = find_vapoursynth_include()
    if inc is None:
        print("VapourSynth4.h / VSHelper4.h not found (set VS_INCLUDE): the VapourSynth plugin "
              f"({os.path.relpath(VS_PLUGIN_SRC, HERE)}) is not built", flush=True)
        return None
    cxx = shutil.which("g++") or "c++"
    out = os.path.join(OUT_DIR, VS_PLUGIN_NAME)
    run([cxx, "-shared", "-fPIC", "-O2", "-std=c++17", f"-I{inc}", VS_PLUGIN_SRC, "-L" + os.path.dirname(lib_path),
         "-ldfttest2_mi355x", "-Wl,-rpath,$ORIGIN", "-o", out])
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    hc = hipcc()
    cmds, objs = [], []
    for src, extra, obj_name in SOURCES:
        obj = os.path.join(obj_dir, obj_name)
        cmds.append([hc, "-c", "-x", "hip", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, *extra,
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    workers = max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        for fut in [pool.submit(run, c) for c in cmds]:
            fut.result()
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + ".tmp"
    run([hc, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    build_vs_plugin(out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
