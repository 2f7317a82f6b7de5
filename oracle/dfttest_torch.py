"""FP64 torch restatement of the DFTTest block pipeline -- TEST INFRASTRUCTURE ONLY.

An autograd graph of the same pipeline as oracle/dfttest_np.py (float planes): mirror
padding (cpu_source/getframe_impl.cpp:22-65), temporal clamp (:294-298), windowed blocks
(:105-156), 3-D real DFT, zero-mean, Wiener gain with each coefficient's own sigma,
mean re-add, centre-slice inverse with the C2R convention (cpu_source/kernel.hpp:1899-1983),
synthesis window and overlap-add (:158-175), store v/scale (:177-221).  Used by the tests
to check the GPU backward pass against autograd, and itself pinned against
oracle/dfttest_np.py (forward) and FP64 finite differences (gradients).
"""

from __future__ import annotations

import numpy as np
import torch

from .dfttest_np import pad_num, pad_size

EPS = 1e-15


def _mirror_index(n: int, pad_n: int) -> torch.Tensor:
    o = (pad_n - n) // 2
    idx = np.arange(pad_n) - o
    idx = np.where(idx < 0, -idx, idx)
    idx = np.where(idx >= n, 2 * (n - 1) - idx, idx)
    return torch.as_tensor(idx, dtype=torch.long)


def filter_plane(x: torch.Tensor, sigma_eff: torch.Tensor, p) -> torch.Tensor:
    """x: [N, H, W] float64 (values on the plugin's float scale); sigma_eff: [T, 16, 9]
    effective sigma (already multiplied by the window's sum of squares).  Wiener gain."""
    n_frames, h, w = x.shape
    bs, step, r, t_len = 16, p.block_step, p.radius, p.t_len
    scale = 255.0
    window = torch.as_tensor(np.asarray(p.window, dtype=np.float32).astype(np.float64)).reshape(t_len, bs, bs)
    ph, pw = pad_size(h, bs, step), pad_size(w, bs, step)
    ny, nx = pad_num(h, bs, step), pad_num(w, bs, step)
    oy, ox = (ph - h) // 2, (pw - w) // 2
    iy, ix = _mirror_index(h, ph), _mirror_index(w, pw)
    padded = x[:, iy][:, :, ix]  # [N, ph, pw]
    wf = None
    if p.zero_mean:
        wfa = np.asarray(p.window_freq, dtype=np.float32).astype(np.float64).reshape(t_len, bs, bs // 2 + 1, 2)
        wf = torch.complex(torch.as_tensor(wfa[..., 0]), torch.as_tensor(wfa[..., 1]))
    tw = torch.exp(2j * torch.pi * torch.arange(t_len, dtype=torch.float64) * r / t_len)
    outs = []
    for n in range(n_frames):
        idx = [min(max(n - r + t, 0), n_frames - 1) for t in range(t_len)]
        slabs = torch.stack([padded[i] for i in idx])  # [T, ph, pw]
        blocks = slabs.unfold(1, bs, step).unfold(2, bs, step)[:, :ny, :nx]  # [T, ny, nx, 16, 16]
        blocks = blocks.permute(1, 2, 0, 3, 4) * (scale * window)
        spec = torch.fft.rfftn(blocks, dim=(-3, -2, -1))
        if wf is not None:
            gf = spec[..., 0, 0, 0].real / wf[0, 0, 0].real
            mean = gf[..., None, None, None] * wf
            spec = spec - mean
        psd = spec.real ** 2 + spec.imag ** 2
        g = torch.relu((psd - sigma_eff) / (psd + EPS))
        spec = spec * g
        if wf is not None:
            spec = spec + mean
        centre = torch.tensordot(spec, tw, dims=([-3], [0]))
        centre = torch.fft.ifft(centre, dim=-2) * bs
        yc = torch.fft.irfft(centre, n=bs, dim=-1) * bs  # [ny, nx, 16, 16]
        yc = yc * window[r]
        cols = yc.reshape(ny * nx, bs * bs).t().unsqueeze(0)  # [1, 256, L]
        acc = torch.nn.functional.fold(cols, (ph, pw), kernel_size=bs, stride=step)[0, 0]
        outs.append(acc[oy:oy + h, ox:ox + w] / scale)
    return torch.stack(outs)
