"""ctypes binding of the dfttest2_mi355x C ABI (include/dfttest2_mi355x.h).

The native library is built in-tree by vs-dfttest2_amd/build_native.py.  There is no
fallback: if the library is missing, loading raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_native", "libdfttest2_mi355x.so")

DFT2_SAMPLE_INT = 0
DFT2_SAMPLE_FLOAT = 1
DFT2_PRECISION_FP32 = 0
DFT2_PRECISION_SPLIT_BF16 = 1

EXPORTED_SYMBOLS = (
    "dft2_plan_create",
    "dft2_plan_destroy",
    "dft2_process_frame",
    "dft2_process_frame_keyed",
    "dft2_plan_forget_frames",
    "dft2_process_batch",
    "dft2_process_batch_sigma",
    "dft2_process_batch_decisions",
    "dft2_process_batch_noalloc",
    "dft2_backward_batch",
    "dft2_plan_set_timing",
    "dft2_plan_get_timing",
    "dft2_batch_create",
    "dft2_batch_launch",
    "dft2_batch_get_info",
    "dft2_batch_get_candidate",
    "dft2_batch_destroy",
    "dft2_rdft",
    "dft2_version",
    "dft2_last_error",
)

_dbl_p = ctypes.POINTER(ctypes.c_double)


class Dft2Params(ctypes.Structure):
    _fields_ = [
        ("radius", ctypes.c_int),
        ("block_size", ctypes.c_int),
        ("block_step", ctypes.c_int),
        ("filter_type", ctypes.c_int),
        ("zero_mean", ctypes.c_int),
        ("window", _dbl_p),
        ("window_len", ctypes.c_int),
        ("sigma", _dbl_p),
        ("sigma_len", ctypes.c_int),
        ("sigma2", ctypes.c_double),
        ("pmin", ctypes.c_double),
        ("pmax", ctypes.c_double),
        ("window_freq", _dbl_p),
        ("window_freq_len", ctypes.c_int),
    ]


class Dft2Job(ctypes.Structure):
    _fields_ = [
        ("src", ctypes.c_void_p),
        ("dst", ctypes.c_void_p),
        ("src_frame_pitch", ctypes.c_longlong),
        ("src_stride", ctypes.c_longlong),
        ("dst_frame_pitch", ctypes.c_longlong),
        ("dst_stride", ctypes.c_longlong),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("n_frames", ctypes.c_int),
        ("frame_begin", ctypes.c_int),
        ("frame_end", ctypes.c_int),
    ]


class Dft2GradJob(ctypes.Structure):
    _fields_ = [
        ("src", ctypes.c_void_p),
        ("grad_out", ctypes.c_void_p),
        ("grad_in", ctypes.c_void_p),
        ("src_frame_pitch", ctypes.c_longlong),
        ("src_stride", ctypes.c_longlong),
        ("go_frame_pitch", ctypes.c_longlong),
        ("go_stride", ctypes.c_longlong),
        ("gi_frame_pitch", ctypes.c_longlong),
        ("gi_stride", ctypes.c_longlong),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("n_frames", ctypes.c_int),
        ("frame_begin", ctypes.c_int),
        ("frame_end", ctypes.c_int),
    ]


class Dft2KernelTiming(ctypes.Structure):
    _fields_ = [
        ("fused_launches", ctypes.c_longlong),
        ("fused_ms", ctypes.c_double),
        ("finalize_launches", ctypes.c_longlong),
        ("finalize_ms", ctypes.c_double),
        ("blocks", ctypes.c_longlong),
        ("slice_spectra", ctypes.c_longlong),
    ]


class Dft2BatchOptions(ctypes.Structure):
    _fields_ = [
        ("autotune", ctypes.c_int),
        ("tile_w", ctypes.c_int),
        ("tile_h", ctypes.c_int),
        ("persistent", ctypes.c_int),
        ("sample_frames", ctypes.c_int),
    ]


class Dft2BatchInfo(ctypes.Structure):
    _fields_ = [
        ("tile_w", ctypes.c_int),
        ("tile_h", ctypes.c_int),
        ("persistent", ctypes.c_int),
        ("workgroups_per_cu", ctypes.c_int),
        ("compute_units", ctypes.c_int),
        ("lds_bytes", ctypes.c_int),
        ("n_chunks", ctypes.c_int),
        ("work_items", ctypes.c_longlong),
        ("scratch_bytes", ctypes.c_longlong),
        ("n_candidates", ctypes.c_int),
        ("tuned_ms", ctypes.c_float),
    ]


class Dft2BatchCandidate(ctypes.Structure):
    _fields_ = [
        ("tile_w", ctypes.c_int),
        ("tile_h", ctypes.c_int),
        ("persistent", ctypes.c_int),
        ("workgroups_per_cu", ctypes.c_int),
        ("lds_bytes", ctypes.c_int),
        ("ms", ctypes.c_float),
    ]


class Dft2Error(RuntimeError):
    pass


_lib = None


def load():
    """Loads the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Dft2Error(f"native library not built: {LIB_PATH} (run vs-dfttest2_amd/build_native.py)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.dft2_plan_create.argtypes = [
        ctypes.POINTER(Dft2Params), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(ctypes.c_void_p),
    ]
    lib.dft2_plan_create.restype = ctypes.c_int
    lib.dft2_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.dft2_plan_destroy.restype = ctypes.c_int
    lib.dft2_process_frame.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_longlong, ctypes.c_void_p, ctypes.c_longlong,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
    ]
    lib.dft2_process_frame.restype = ctypes.c_int
    lib.dft2_process_frame_keyed.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_longlong,
        ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
    ]
    lib.dft2_process_frame_keyed.restype = ctypes.c_int
    lib.dft2_plan_forget_frames.argtypes = [ctypes.c_void_p]
    lib.dft2_plan_forget_frames.restype = ctypes.c_int
    lib.dft2_process_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Dft2Job), ctypes.c_int, ctypes.c_void_p]
    lib.dft2_process_batch.restype = ctypes.c_int
    lib.dft2_process_batch_sigma.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(Dft2Job), ctypes.c_int, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
    ]
    lib.dft2_process_batch_sigma.restype = ctypes.c_int
    lib.dft2_process_batch_decisions.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(Dft2Job), ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
    ]
    lib.dft2_process_batch_decisions.restype = ctypes.c_int
    lib.dft2_process_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Dft2Job), ctypes.c_int,
                                               ctypes.c_void_p]
    lib.dft2_process_batch_noalloc.restype = ctypes.c_int
    lib.dft2_backward_batch.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(Dft2GradJob), ctypes.c_int, ctypes.c_void_p, ctypes.c_float,
        ctypes.c_void_p, ctypes.c_void_p,
    ]
    lib.dft2_backward_batch.restype = ctypes.c_int
    lib.dft2_plan_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.dft2_plan_set_timing.restype = ctypes.c_int
    lib.dft2_plan_get_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(Dft2KernelTiming)]
    lib.dft2_plan_get_timing.restype = ctypes.c_int
    lib.dft2_batch_create.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(Dft2Job), ctypes.c_int, ctypes.POINTER(Dft2BatchOptions),
        ctypes.POINTER(ctypes.c_void_p),
    ]
    lib.dft2_batch_create.restype = ctypes.c_int
    lib.dft2_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.dft2_batch_launch.restype = ctypes.c_int
    lib.dft2_batch_get_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(Dft2BatchInfo)]
    lib.dft2_batch_get_info.restype = ctypes.c_int
    lib.dft2_batch_get_candidate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Dft2BatchCandidate)]
    lib.dft2_batch_get_candidate.restype = ctypes.c_int
    lib.dft2_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.dft2_batch_destroy.restype = ctypes.c_int
    lib.dft2_rdft.argtypes = [_dbl_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _dbl_p, ctypes.c_int]
    lib.dft2_rdft.restype = ctypes.c_int
    lib.dft2_version.argtypes = []
    lib.dft2_version.restype = ctypes.c_char_p
    lib.dft2_last_error.argtypes = []
    lib.dft2_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = load().dft2_last_error().decode(errors="replace")
        raise Dft2Error(f"dfttest2_mi355x error {rc}: {msg}")


def version() -> str:
    return load().dft2_version().decode()


def rdft(data: Sequence[float], shape: Sequence[int]) -> np.ndarray:
    """RDFT through the C ABI (same contract as the plugin's RDFT)."""
    arr = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
    shp = (ctypes.c_int * len(shape))(*[int(s) for s in shape])
    n_out = int(np.prod(shape[:-1])) * (int(shape[-1]) // 2 + 1) * 2 if len(shape) else 0
    out = np.zeros(max(n_out, 1), dtype=np.float64)
    _check(load().dft2_rdft(arr.ctypes.data_as(_dbl_p), arr.size, shp, len(shape), out.ctypes.data_as(_dbl_p),
                            out.size))
    return out[:n_out]


# DFT2_DEFAULT_STREAM: the null (default) stream as a stream argument.  torch's default
# stream has handle 0, which the C ABI would take as "no stream" (the plan's own stream,
# and a blocking call).
DFT2_DEFAULT_STREAM = 1


def _stream_arg(stream: Optional[int]) -> ctypes.c_void_p:
    """None: the plan's (or batch's) own stream, and the call waits for completion.
    Otherwise a hipStream_t handle such as a torch stream's cuda_stream, where 0 (torch's
    default stream) is passed as DFT2_DEFAULT_STREAM: work is queued, not waited for."""
    if stream is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(int(stream) or DFT2_DEFAULT_STREAM)


class Plan:
    """One set of DFTTest plugin parameters on one GPU (dft2_plan_create)."""

    def __init__(self, params, sample_type: int, bits_per_sample: int, device_id: int = 0, num_streams: int = 1,
                 precision: int = DFT2_PRECISION_FP32):
        lib = load()
        self._keep = []
        window = np.ascontiguousarray(params.window, dtype=np.float64)
        sigma = np.ascontiguousarray(np.atleast_1d(params.sigma), dtype=np.float64)
        wf = params.window_freq
        wf = None if wf is None else np.ascontiguousarray(wf, dtype=np.float64)
        self._keep += [window, sigma, wf]
        prm = Dft2Params(
            radius=params.radius,
            block_size=params.block_size,
            block_step=params.block_step,
            filter_type=params.filter_type,
            zero_mean=int(bool(params.zero_mean)),
            window=window.ctypes.data_as(_dbl_p),
            window_len=window.size,
            sigma=sigma.ctypes.data_as(_dbl_p),
            sigma_len=sigma.size,
            sigma2=params.sigma2,
            pmin=params.pmin,
            pmax=params.pmax,
            window_freq=wf.ctypes.data_as(_dbl_p) if wf is not None else _dbl_p(),
            window_freq_len=0 if wf is None else wf.size,
        )
        handle = ctypes.c_void_p()
        _check(lib.dft2_plan_create(ctypes.byref(prm), sample_type, bits_per_sample, device_id, num_streams,
                                    precision, ctypes.byref(handle)))
        self._h = handle
        self.radius = params.radius
        self.device_id = device_id

    def close(self):
        if getattr(self, "_h", None):
            load().dft2_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, enable: bool):
        """Records HIP events around every fused / finalize launch while enabled."""
        _check(load().dft2_plan_set_timing(self._h, int(enable)))

    def timing(self) -> dict:
        """Summed fused / finalize kernel times (ms) and launch counts since set_timing(True)."""
        out = Dft2KernelTiming()
        _check(load().dft2_plan_get_timing(self._h, ctypes.byref(out)))
        return {name: getattr(out, name) for name, _ in Dft2KernelTiming._fields_}

    def process_batch(self, jobs: List[Dft2Job], stream: Optional[int] = None):
        arr = (Dft2Job * len(jobs))(*jobs)
        _check(load().dft2_process_batch(self._h, arr, len(jobs), _stream_arg(stream)))

    def process_batch_sigma(self, jobs: List[Dft2Job], sigma_ptr: int, sigma_scale: float,
                            stream: Optional[int] = None):
        arr = (Dft2Job * len(jobs))(*jobs)
        _check(load().dft2_process_batch_sigma(self._h, arr, len(jobs), ctypes.c_void_p(sigma_ptr),
                                               ctypes.c_float(sigma_scale), _stream_arg(stream)))

    def process_batch_noalloc(self, jobs: List[Dft2Job], stream: Optional[int] = None):
        """As process_batch, through the entry point that allocates nothing."""
        arr = (Dft2Job * len(jobs))(*jobs)
        _check(load().dft2_process_batch_noalloc(self._h, arr, len(jobs), _stream_arg(stream)))

    def process_batch_decisions(self, jobs: List[Dft2Job], decisions_ptr: int, decisions_len: int,
                                stream: Optional[int] = None):
        """As process_batch, also recording every block's hard gain decisions (filter types 1
        and 3) into decisions_len 64-bit masks at device address decisions_ptr."""
        arr = (Dft2Job * len(jobs))(*jobs)
        _check(load().dft2_process_batch_decisions(self._h, arr, len(jobs), ctypes.c_void_p(decisions_ptr),
                                                   ctypes.c_longlong(decisions_len), _stream_arg(stream)))

    def backward_batch(self, jobs: List[Dft2GradJob], sigma_ptr: int, sigma_scale: float, sigma_grad_ptr: int,
                       stream: Optional[int] = None):
        arr = (Dft2GradJob * len(jobs))(*jobs)
        _check(load().dft2_backward_batch(self._h, arr, len(jobs), ctypes.c_void_p(sigma_ptr),
                                          ctypes.c_float(sigma_scale), ctypes.c_void_p(sigma_grad_ptr),
                                          _stream_arg(stream)))

    def process_frame(self, src_ptrs: Sequence[int], src_stride: int, dst: int, dst_stride: int, width: int,
                      height: int, memory: int, stream: Optional[int] = None):
        ptrs = (ctypes.c_void_p * len(src_ptrs))(*src_ptrs)
        _check(load().dft2_process_frame(self._h, ptrs, src_stride, ctypes.c_void_p(dst), dst_stride, width, height,
                                         memory, _stream_arg(stream)))

    def process_frame_keyed(self, src_ptrs: Sequence[int], keys: Sequence[int], src_stride: int, dst: int,
                            dst_stride: int, width: int, height: int):
        """process_frame on host planes, keeping device copies of the source planes named by
        `keys` between calls (a key must always name the same plane content)."""
        ptrs = (ctypes.c_void_p * len(src_ptrs))(*src_ptrs)
        ks = (ctypes.c_ulonglong * len(keys))(*[int(k) for k in keys])
        _check(load().dft2_process_frame_keyed(self._h, ptrs, ks, src_stride, ctypes.c_void_p(dst), dst_stride,
                                               width, height))

    def forget_frames(self):
        _check(load().dft2_plan_forget_frames(self._h))


class Batch:
    """A fixed set of jobs on one plan, run as one captured graph launch (dft2_batch_*).

    autotune=True times every valid tile shape and launch form on the first
    `sample_frames` output frames of every job and keeps the fastest; otherwise tile_w /
    tile_h (0 = the plan's default) and persistent are used as given.  The job buffers
    must stay allocated, at the same addresses, while the batch exists."""

    def __init__(self, plan: Plan, jobs: List[Dft2Job], autotune: bool = True, tile_w: int = 0, tile_h: int = 0,
                 persistent: bool = False, sample_frames: int = 0):
        lib = load()
        self.plan = plan  # keeps the plan alive
        arr = (Dft2Job * len(jobs))(*jobs)
        opt = Dft2BatchOptions(autotune=int(autotune), tile_w=tile_w, tile_h=tile_h, persistent=int(persistent),
                               sample_frames=sample_frames)
        handle = ctypes.c_void_p()
        _check(lib.dft2_batch_create(plan._h, arr, len(jobs), ctypes.byref(opt), ctypes.byref(handle)))
        self._h = handle

    def launch(self, stream: Optional[int] = None):
        _check(load().dft2_batch_launch(self._h, _stream_arg(stream)))

    def info(self) -> dict:
        out = Dft2BatchInfo()
        _check(load().dft2_batch_get_info(self._h, ctypes.byref(out)))
        return {name: getattr(out, name) for name, _ in Dft2BatchInfo._fields_}

    def candidates(self) -> list:
        res = []
        for i in range(self.info()["n_candidates"]):
            c = Dft2BatchCandidate()
            _check(load().dft2_batch_get_candidate(self._h, i, ctypes.byref(c)))
            res.append({name: getattr(c, name) for name, _ in Dft2BatchCandidate._fields_})
        return res

    def close(self):
        if getattr(self, "_h", None):
            load().dft2_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
