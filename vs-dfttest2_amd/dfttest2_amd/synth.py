"""Seeded synthetic clips for tests and benchmarks.

frame_t(x, y) = g(x, y) + sum_k a_k sin(2 pi <u_k, (x - t, y)> / P_k + phi_k) + noise
with g = 0.25 + 0.5 (0.6 x / W + 0.4 y / H), a_k = 0.08, P_k ~ U[8, 64],
direction u_k = (cos theta_k, sin theta_k), theta_k, phi_k ~ U[0, 2 pi): a smooth
gradient plus three sinusoids translating 1 px per frame, plus Gaussian noise.
Sinusoid parameters come from numpy's PCG64(seed); the noise from a torch generator
seeded with `seed` on the target device.
"""

from __future__ import annotations

import math
from typing import List, Optional

import numpy as np


def _torch():
    import torch

    return torch


def clean_plane(n_frames: int, height: int, width: int, seed: int, device="cpu"):
    """Noise-free frames [N, H, W] (float32, values about 0.0..1.0)."""
    torch = _torch()
    rng = np.random.Generator(np.random.PCG64(seed))
    periods = rng.uniform(8.0, 64.0, 3)
    thetas = rng.uniform(0.0, 2 * math.pi, 3)
    phases = rng.uniform(0.0, 2 * math.pi, 3)
    y = torch.arange(height, device=device, dtype=torch.float32).view(1, height, 1)
    x = torch.arange(width, device=device, dtype=torch.float32).view(1, 1, width)
    t = torch.arange(n_frames, device=device, dtype=torch.float32).view(n_frames, 1, 1)
    out = 0.25 + 0.5 * (0.6 * x / width + 0.4 * y / height) + torch.zeros(n_frames, 1, 1, device=device)
    for k in range(3):
        ux, uy = math.cos(thetas[k]), math.sin(thetas[k])
        out = out + 0.08 * torch.sin((2 * math.pi / periods[k]) * (ux * (x - t) + uy * y) + phases[k])
    return out


def noise(shape, sd: float, seed: int, device="cpu"):
    torch = _torch()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    return torch.randn(shape, generator=gen, device=device, dtype=torch.float32) * sd


def float_plane(n_frames, height, width, seed, sd=0.05, clip=True, device="cpu"):
    """Float32 plane: clean + N(0, sd), clipped to [0, 1] unless clip=False."""
    f = clean_plane(n_frames, height, width, seed, device) + noise((n_frames, height, width), sd, seed, device)
    return f.clamp(0.0, 1.0) if clip else f


def int_plane(n_frames, height, width, seed, bits=8, sd_lsb=6.0, device="cpu"):
    """Limited-range integer plane: round(lo + span*clean + N(0, sd)), clipped to [lo, hi].

    8 bit: [16, 235]; 10 bit: [64, 940] (in a uint16 container).
    """
    torch = _torch()
    lo, hi = 16 << (bits - 8), 235 << (bits - 8)
    if bits == 10:
        lo, hi = 64, 940
    f = lo + (hi - lo) * clean_plane(n_frames, height, width, seed, device)
    f = f + noise((n_frames, height, width), sd_lsb, seed, device)
    f = torch.round(f).clamp(lo, hi)
    return f.to(torch.uint8) if bits == 8 else f.to(torch.int32).to(torch.uint16)


def yuv420_planes(n_frames, width, height, seed, bits=8, sd_lsb=6.0, device="cpu") -> List:
    """Three planes of a 4:2:0 clip (luma W x H, chroma W/2 x H/2), seeds seed, seed+1, seed+2."""
    return [
        int_plane(n_frames, height, width, seed, bits, sd_lsb, device),
        int_plane(n_frames, height // 2, width // 2, seed + 1, bits, sd_lsb, device),
        int_plane(n_frames, height // 2, width // 2, seed + 2, bits, sd_lsb, device),
    ]
