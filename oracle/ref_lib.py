"""ctypes access to oracle/_ref/libdft2ref.so -- TEST INFRASTRUCTURE ONLY.

The library wraps the reference GCC backend's fused() (see oracle/ref_harness.cpp);
it exists only where oracle/Makefile could build it from the reference sources.
Only tests, tests/golden/make_golden.py and bench.py's cpu_baseline leg use it.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libdft2ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"reference harness not built: {LIB_PATH}")
        _lib = ctypes.CDLL(LIB_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        _lib.ref_filter_plane.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, d, d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_int, ctypes.c_int, d, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
        ]
        _lib.ref_filter_plane.restype = ctypes.c_int
        _lib.ref_fused_block.argtypes = [
            ctypes.POINTER(ctypes.c_float), d, d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_int, ctypes.c_int, d, ctypes.c_int,
        ]
        _lib.ref_fused_block.restype = ctypes.c_int
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def filter_plane(frames: np.ndarray, p, bits: int, is_float: bool, syn_slice: str = "centre", threads: int = 1):
    """Reference fused() over one plane [N, H, W]; returns the filtered plane."""
    frames = np.ascontiguousarray(frames)
    n, h, w = frames.shape
    out = np.empty_like(frames)
    window = np.ascontiguousarray(p.window, dtype=np.float64)
    sigma = np.ascontiguousarray(p.sigma_array().reshape(-1), dtype=np.float64)
    wf = p.window_freq if p.window_freq is not None else np.zeros(p.t_len * 16 * 9 * 2)
    wf = np.ascontiguousarray(wf, dtype=np.float64)
    rc = lib().ref_filter_plane(
        frames.ctypes.data, n, w, h, w, frames.dtype.itemsize, 32 if is_float else bits, int(is_float),
        _dptr(window), _dptr(sigma), p.sigma2, p.pmin, p.pmax, p.filter_type, int(p.zero_mean), _dptr(wf),
        p.radius, p.block_step, 1 if syn_slice == "cpu" else 0, out.ctypes.data, threads,
    )
    if rc != 0:
        raise RuntimeError(f"ref_filter_plane failed: {rc}")
    return out


def fused_block(block: np.ndarray, p) -> np.ndarray:
    """Reference fused() on one windowed block [T, 16, 16] float32 -> centre slice [16, 16]."""
    b = np.ascontiguousarray(block, dtype=np.float32).copy()
    window = np.ascontiguousarray(p.window, dtype=np.float64)
    sigma = np.ascontiguousarray(p.sigma_array().reshape(-1), dtype=np.float64)
    wf = p.window_freq if p.window_freq is not None else np.zeros(p.t_len * 16 * 9 * 2)
    wf = np.ascontiguousarray(wf, dtype=np.float64)
    rc = lib().ref_fused_block(
        b.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _dptr(window), _dptr(sigma), p.sigma2, p.pmin, p.pmax,
        p.filter_type, int(p.zero_mean), _dptr(wf), p.radius,
    )
    if rc != 0:
        raise RuntimeError(f"ref_fused_block failed: {rc}")
    return b.reshape(-1)[:256].reshape(16, 16).copy()
