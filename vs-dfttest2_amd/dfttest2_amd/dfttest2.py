"""DFTTest front end for the MI355X backend.

Mirrors the reference's python/dfttest2.py:24-421 (DFTTest, DFTTest2) and
python/backends.py:10-63 (Backend, select_backend) without VapourSynth: a clip is a
`VideoClip` of planes [frames, height, width] held as torch tensors on the GPU (or
numpy arrays, which are moved to the GPU and back).  Parameter translation is
params.dfttest_params / dfttest2_params; the filtering runs in the native library.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Union

import numpy as np

from . import _lib
from .params import FilterParams, LocationData, dfttest2_params, dfttest_params

__all__ = ["Backend", "VideoClip", "DFTTest", "DFTTest2", "select_backend", "filter_clip"]


class Backend:
    """Backend registry (python/backends.py:10-37) with the MI355X entry."""

    @dataclass
    class MI355X:
        device_id: int = 0
        num_streams: int = 1
        precision: str = "fp32"  # "fp32" or "split_bf16"


def select_backend(backend, sbsize: int, tbsize: int):
    """python/backends.py:40-63.  The MI355X backend serves both of the reference's
    candidate lists: sbsize 16 with tbsize 1, 3, 5 or 7 on the fused kernel (the fused
    backends' case), and other block sizes, or tbsize up to 15, on the generic block path
    (the case the reference gives its FFT-library backends)."""
    if backend is not None:
        return backend() if isinstance(backend, type) else backend
    if 1 <= sbsize <= 64 and tbsize % 2 == 1 and 1 <= tbsize <= 15:
        return Backend.MI355X()
    raise RuntimeError("No backend are available")


@dataclass
class VideoClip:
    """A constant-format clip: planes [frames, height, width] (torch tensors or numpy arrays)."""

    planes: List
    sample_type: str = "int"  # "int" or "float"
    bits_per_sample: int = 8

    @property
    def num_frames(self) -> int:
        return int(self.planes[0].shape[0])

    @staticmethod
    def from_array(arr, bits_per_sample: Optional[int] = None) -> "VideoClip":
        dt = str(arr.dtype)
        if "float32" in dt:
            return VideoClip([arr], "float", 32)
        if "uint8" in dt:
            return VideoClip([arr], "int", bits_per_sample or 8)
        if "uint16" in dt:
            return VideoClip([arr], "int", bits_per_sample or 16)
        raise ValueError("only constant format 8-16 bit integer and 32 bit float input supported")


_PLAN_CACHE: dict = {}


def _plan_key(p: FilterParams, sample_type: int, bits: int, be) -> tuple:
    h = hashlib.sha1()
    for arr in (p.window, p.sigma, p.window_freq if p.window_freq is not None else np.zeros(1)):
        h.update(np.ascontiguousarray(arr, dtype=np.float64).tobytes())
    h.update(repr((p.radius, p.block_size, p.block_step, p.filter_type, p.zero_mean, p.sigma2, p.pmin,
                   p.pmax)).encode())
    return (h.hexdigest(), sample_type, bits, be.device_id, be.num_streams, be.precision)


def get_plan(p: FilterParams, sample_type: str, bits: int, be) -> _lib.Plan:
    st = _lib.DFT2_SAMPLE_FLOAT if sample_type == "float" else _lib.DFT2_SAMPLE_INT
    prec = {"fp32": _lib.DFT2_PRECISION_FP32, "split_bf16": _lib.DFT2_PRECISION_SPLIT_BF16}[be.precision]
    key = _plan_key(p, st, bits, be)
    plan = _PLAN_CACHE.get(key)
    if plan is None:
        plan = _lib.Plan(p, st, bits, be.device_id, be.num_streams, prec)
        _PLAN_CACHE[key] = plan
    return plan


def _torch():
    import torch

    return torch


def plane_job(src, dst, frame_begin: int = 0, frame_end: Optional[int] = None) -> _lib.Dft2Job:
    """C-ABI job for one contiguous device plane [N, H, W] (torch tensors)."""
    n, h, w = (int(s) for s in src.shape)
    es = src.element_size()
    return _lib.Dft2Job(
        src=src.data_ptr(), dst=dst.data_ptr(),
        src_frame_pitch=src.stride(0) * es, src_stride=src.stride(1) * es,
        dst_frame_pitch=dst.stride(0) * es, dst_stride=dst.stride(1) * es,
        width=w, height=h, n_frames=n, frame_begin=frame_begin, frame_end=n if frame_end is None else frame_end,
    )


def filter_clip(clip: VideoClip, p: FilterParams, be=None, planes: Optional[Sequence[int]] = None) -> VideoClip:
    """Filters `planes` of `clip` (all when None); other planes are passed through."""
    torch = _torch()
    be = select_backend(be, p.block_size, p.t_len)
    plan = get_plan(p, clip.sample_type, clip.bits_per_sample, be)
    if planes is None:
        planes = p.planes
    if planes is None:
        planes = list(range(len(clip.planes)))
    for pl in planes:
        if pl < 0 or pl >= len(clip.planes):
            raise ValueError("plane index out of range")
    if len(set(planes)) != len(planes):
        raise ValueError("plane specified twice")
    device = torch.device("cuda", be.device_id)
    was_numpy = [isinstance(x, np.ndarray) for x in clip.planes]
    dev_in = []
    for x in clip.planes:
        t = torch.from_numpy(np.ascontiguousarray(x)) if isinstance(x, np.ndarray) else x
        if t.device != device:
            t = t.to(device)
        dev_in.append(t.contiguous())
    dev_out = [torch.empty_like(t) if i in planes else t for i, t in enumerate(dev_in)]
    jobs = [plane_job(dev_in[i], dev_out[i]) for i in planes]
    with torch.cuda.device(device):
        stream = torch.cuda.current_stream(device).cuda_stream
        plan.process_batch(jobs, stream)
    out = []
    for i, t in enumerate(dev_out):
        out.append(t.cpu().numpy() if was_numpy[i] else t)
    return VideoClip(out, clip.sample_type, clip.bits_per_sample)


def _as_clip(clip) -> VideoClip:
    return clip if isinstance(clip, VideoClip) else VideoClip.from_array(clip)


def DFTTest(
    clip,
    ftype: int = 0,
    sigma: float = 8.0,
    sigma2: float = 8.0,
    pmin: float = 0.0,
    pmax: float = 500.0,
    sbsize: int = 16,
    smode: int = 1,
    sosize: int = 12,
    tbsize: int = 3,
    swin: int = 0,
    twin: int = 7,
    sbeta: float = 2.5,
    tbeta: float = 2.5,
    zmean: bool = True,
    f0beta: float = 1.0,
    nlocation: Optional[Sequence[int]] = None,
    alpha: Optional[float] = None,
    slocation: Optional[LocationData] = None,
    ssx: Optional[LocationData] = None,
    ssy: Optional[LocationData] = None,
    sst: Optional[LocationData] = None,
    ssystem: int = 0,
    planes: Optional[Union[int, Sequence[int]]] = None,
    backend=None,
):
    """2D/3D frequency-domain denoiser with the reference DFTTest() surface (python/dfttest2.py:24-237).

    `clip` is a VideoClip or a single plane [frames, height, width] (torch tensor or
    numpy array of uint8 / uint16 / float32).  Returns the same kind of object.
    """
    vc = _as_clip(clip)
    if vc.sample_type == "int" and vc.bits_per_sample > 16 or vc.sample_type == "float" and vc.bits_per_sample != 32:
        raise ValueError("only constant format 8-16 bit integer and 32 bit float input supported")
    p = dfttest_params(
        ftype=ftype, sigma=sigma, sigma2=sigma2, pmin=pmin, pmax=pmax, sbsize=sbsize, smode=smode, sosize=sosize,
        tbsize=tbsize, swin=swin, twin=twin, sbeta=sbeta, tbeta=tbeta, zmean=zmean, f0beta=f0beta,
        nlocation=nlocation, alpha=alpha, slocation=slocation, ssx=ssx, ssy=ssy, sst=sst, ssystem=ssystem,
        planes=planes,
    )
    be = select_backend(backend, sbsize, tbsize)
    out = filter_clip(vc, p, be)
    return out if isinstance(clip, VideoClip) else out.planes[0]


def DFTTest2(clip, ftype: int = 0, sigma: Union[float, Callable, Sequence[Callable]] = 8.0, sigma2: float = 8.0,
             pmin: float = 0.0, pmax: float = 500.0, sbsize: int = 16, sosize: int = 12, tbsize: int = 3,
             swin: int = 0, twin: int = 7, sbeta: float = 2.5, tbeta: float = 2.5, zmean: bool = True,
             f0beta: float = 1.0, ssystem: int = 0, planes=None, backend=None):
    """Lower-level entry with sigma as a value or per-axis functions (python/dfttest2.py:240-421)."""
    vc = _as_clip(clip)
    p = dfttest2_params(ftype=ftype, sigma=sigma, sigma2=sigma2, pmin=pmin, pmax=pmax, sbsize=sbsize, sosize=sosize,
                        tbsize=tbsize, swin=swin, twin=twin, sbeta=sbeta, tbeta=tbeta, zmean=zmean, f0beta=f0beta,
                        ssystem=ssystem, planes=planes)
    out = filter_clip(vc, p, select_backend(backend, sbsize, tbsize))
    return out if isinstance(clip, VideoClip) else out.planes[0]
