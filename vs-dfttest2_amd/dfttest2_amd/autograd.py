"""Differentiable DFTTest (Wiener gain) for PyTorch on the MI355X kernels.

`dfttest_diff(x, sigma, ...)` filters float32 planes x [..., frames, H, W] on the GPU
with the per-frequency sigma array `sigma` [T, 16, 9] in user units (the plugin scales
it by the window's sum of squares, python/dfttest2.py:327-333), and backpropagates to
both x and sigma through the hand-written backward kernel (dft2_backward_batch), which
recomputes block spectra instead of storing them.  torch is used for device memory and
the current stream only.
"""

from __future__ import annotations

from typing import Optional

from . import _lib
from .dfttest2 import Backend, get_plan
from .params import FilterParams, dfttest_params


def _torch():
    import torch

    return torch


def _job(x, y):
    n, h, w = (int(v) for v in x.shape)
    return _lib.Dft2Job(src=x.data_ptr(), dst=y.data_ptr(), src_frame_pitch=x.stride(0) * 4,
                        src_stride=x.stride(1) * 4, dst_frame_pitch=y.stride(0) * 4, dst_stride=y.stride(1) * 4,
                        width=w, height=h, n_frames=n, frame_begin=0, frame_end=n)


def _grad_job(x, gy, gx):
    n, h, w = (int(v) for v in x.shape)
    return _lib.Dft2GradJob(src=x.data_ptr(), grad_out=gy.data_ptr(), grad_in=gx.data_ptr(),
                            src_frame_pitch=x.stride(0) * 4, src_stride=x.stride(1) * 4,
                            go_frame_pitch=gy.stride(0) * 4, go_stride=gy.stride(1) * 4,
                            gi_frame_pitch=gx.stride(0) * 4, gi_stride=gx.stride(1) * 4,
                            width=w, height=h, n_frames=n, frame_begin=0, frame_end=n)


def _planes(t):
    return t.reshape(-1, *t.shape[-3:])


class _DFTTestFn:  # namespace for the autograd.Function (built lazily so torch stays optional)
    fn = None


def _function():
    if _DFTTestFn.fn is not None:
        return _DFTTestFn.fn
    torch = _torch()

    class DFTTestFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, sigma, plan, sigma_scale):
            xs = _planes(x.contiguous())
            sig = sigma.contiguous()
            y = torch.empty_like(xs)
            stream = torch.cuda.current_stream(x.device).cuda_stream
            plan.process_batch_sigma([_job(xs[i], y[i]) for i in range(xs.shape[0])], sig.data_ptr(), sigma_scale,
                                     stream)
            ctx.save_for_backward(xs, sig)
            ctx.plan, ctx.sigma_scale, ctx.shape = plan, sigma_scale, x.shape
            return y.reshape(x.shape)

        @staticmethod
        def backward(ctx, gy):
            xs, sig = ctx.saved_tensors
            gys = _planes(gy.contiguous())
            gx = torch.zeros_like(xs)
            gs = torch.zeros_like(sig)
            stream = torch.cuda.current_stream(xs.device).cuda_stream
            ctx.plan.backward_batch([_grad_job(xs[i], gys[i], gx[i]) for i in range(xs.shape[0])], sig.data_ptr(),
                                    ctx.sigma_scale, gs.data_ptr(), stream)
            return gx.reshape(ctx.shape), gs, None, None

    _DFTTestFn.fn = DFTTestFunction
    return DFTTestFunction


def diff_params(tbsize: int = 3, sbsize: int = 16, sosize: int = 12, swin: int = 0, twin: int = 7,
                sbeta: float = 2.5, tbeta: float = 2.5, zmean: bool = True) -> FilterParams:
    """Plugin parameters of the differentiable Wiener filter (sigma comes from the tensor)."""
    return dfttest_params(ftype=0, sigma=1.0, sbsize=sbsize, sosize=sosize, tbsize=tbsize, swin=swin, twin=twin,
                          sbeta=sbeta, tbeta=tbeta, zmean=zmean)


def dfttest_diff(x, sigma, params: Optional[FilterParams] = None, device_id: Optional[int] = None, **kw):
    """Wiener DFTTest of float32 planes x [..., frames, H, W] with a differentiable sigma [T, 16, 9].

    Values of x are on the float scale of the plugin (0..1).  Gradients flow to x and sigma.
    """
    p = params if params is not None else diff_params(**kw)
    if x.dtype != _torch().float32 or not x.is_cuda:
        raise ValueError("x must be a float32 CUDA tensor")
    if tuple(sigma.shape) != (p.t_len, 16, 9):
        raise ValueError(f"sigma must have shape {(p.t_len, 16, 9)}")
    dev = x.device.index if device_id is None else device_id
    plan = get_plan(p, "float", 32, Backend.MI355X(device_id=dev))
    return _function().apply(x, sigma, plan, float(p.extra.get("wscale", 1.0)))
