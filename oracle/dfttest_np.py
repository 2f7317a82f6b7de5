"""FP64 NumPy oracle of the DFTTest block pipeline -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker.  The product path never calls it.

It restates the reference CPU backend per output frame and plane:
  * padding sizes / block counts: cpu_source/getframe_impl.cpp:11-20
  * mirror padding without edge repeat: cpu_source/getframe_impl.cpp:22-65
  * temporal clamp of the T source frames: cpu_source/getframe_impl.cpp:294-298
  * block load, b = scale * w * p: cpu_source/getframe_impl.cpp:105-156
  * fused(): 3-D real DFT, zero-mean, gain, mean re-add, centre-slice inverse with the
    C2R convention (Im of kx = 0 and kx = B/2 ignored): cpu_source/kernel.hpp:1899-1983
  * gains of filter types 0..6 with the 1e-15 epsilon: cpu_source/kernel.hpp:1826-1890
  * overlap-add with the synthesis window: cpu_source/getframe_impl.cpp:158-175, 377-384
  * store, int: clamp(trunc(v/scale + 0.5), 0, peak), float: v/scale:
    cpu_source/getframe_impl.cpp:177-221
The synthesis window is the centre temporal slice w[r] as every GPU backend uses
(nvrtc_source/kernels.cu:337, hiprtc_source/kernel.hpp:324); `syn_slice="cpu"` selects
the CPU backend's slice 2r (cpu_source/getframe_impl.cpp:383) instead.

Arithmetic is float64 throughout; the parameters (window, sigma, window_freq, sigma2,
pmin, pmax) are first rounded to float32 as the plugin does when it reads them
(cpu_source/source.cpp:145-203), so the oracle measures arithmetic error only.
Parity pins: tests/test_oracle.py (textbook DFT on tiny blocks, closed-form
invariants, identity cases, and the golden vectors made from the reference's own
fused() in tests/golden/).
"""

from __future__ import annotations

import numpy as np

EPS = 1e-15


def pad_size(size: int, block_size: int, block_step: int) -> int:
    """cpu_source/getframe_impl.cpp:11-16."""
    return size + ((block_size - size % block_size) if size % block_size else 0) + max(
        block_size - block_step, block_step
    ) * 2


def pad_num(size: int, block_size: int, block_step: int) -> int:
    """cpu_source/getframe_impl.cpp:18-20."""
    return (pad_size(size, block_size, block_step) - block_size) // block_step + 1


def mirror_pad(frame: np.ndarray, block_size: int, block_step: int) -> np.ndarray:
    """Reflection padding without edge repeat (cpu_source/getframe_impl.cpp:22-65)."""
    h, w = frame.shape
    ph, pw = pad_size(h, block_size, block_step), pad_size(w, block_size, block_step)
    oy, ox = (ph - h) // 2, (pw - w) // 2
    ys = np.arange(ph) - oy
    xs = np.arange(pw) - ox
    ys = np.where(ys < 0, -ys, ys)
    ys = np.where(ys >= h, 2 * (h - 1) - ys, ys)
    xs = np.where(xs < 0, -xs, xs)
    xs = np.where(xs >= w, 2 * (w - 1) - xs, xs)
    if ys.min() < 0 or ys.max() >= h or xs.min() < 0 or xs.max() >= w:
        raise ValueError("frame too small for single reflection padding")
    return frame[np.ix_(ys, xs)]


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def gain(psd: np.ndarray, sigma: np.ndarray, sigma2: float, pmin: float, pmax: float, filter_type: int) -> np.ndarray:
    """Multiplier of filter types 0..6 (cpu_source/kernel.hpp:1826-1890)."""
    if filter_type == 0:
        return np.maximum((psd - sigma) / (psd + EPS), 0.0)
    if filter_type == 1:
        return np.where(psd < sigma, 0.0, 1.0)
    if filter_type == 2:
        return np.broadcast_to(sigma, psd.shape).astype(np.float64)
    if filter_type == 3:
        return np.where((psd >= pmin) & (psd <= pmax), sigma, sigma2)
    if filter_type == 4:
        return sigma * np.sqrt(psd * (pmax / ((psd + pmin) * (psd + pmax) + EPS)))
    if filter_type == 5:
        return np.power(np.maximum((psd - sigma) / (psd + EPS), 0.0), pmin)
    if filter_type == 6:
        return np.sqrt(np.maximum((psd - sigma) / (psd + EPS), 0.0))
    raise ValueError(f"invalid filter_type {filter_type}")


def sample_scale(bits: int, is_float: bool) -> float:
    """cpu_source/getframe_impl.cpp:117-120."""
    return 255.0 if is_float else 1.0 / (1 << (bits - 8))


def filter_blocks(blocks: np.ndarray, p) -> np.ndarray:
    """fused() on windowed blocks [..., T, B, B] -> centre-slice spatial output [..., B, B].

    cpu_source/kernel.hpp:1899-1983: unnormalised forward 3-D real DFT, zero-mean,
    gain, mean re-add, inverse over t evaluated at the centre slice, inverse over y,
    C2R over x (Im at kx = 0 and kx = B/2 ignored).
    """
    t_len, bs = p.t_len, p.block_size
    r = p.radius
    spec = np.fft.rfftn(blocks, axes=(-3, -2, -1))  # [..., T, B, B/2+1]
    sigma = f32(p.sigma_array())
    if p.zero_mean:
        wf = f32(p.window_freq).reshape(t_len, bs, bs // 2 + 1, 2)
        wf = wf[..., 0] + 1j * wf[..., 1]
        gf = spec[..., 0, 0, 0].real / wf[0, 0, 0].real
        mean = gf[..., None, None, None] * wf
        spec = spec - mean
    psd = spec.real ** 2 + spec.imag ** 2
    g = gain(psd, sigma, float(f32(p.sigma2)), float(f32(p.pmin)), float(f32(p.pmax)), p.filter_type)
    spec = spec * g
    if p.zero_mean:
        spec = spec + mean
    # inverse over t at t = r only
    kt = np.arange(t_len)
    tw = np.exp(2j * np.pi * kt * r / t_len)
    centre = np.tensordot(spec, tw, axes=([-3], [0]))  # [..., B, B/2+1]
    centre = np.fft.ifft(centre, axis=-2) * bs
    out = np.fft.irfft(centre, n=bs, axis=-1) * bs
    return out


def process_plane(frames: np.ndarray, p, bits: int, is_float: bool, syn_slice: str = "centre",
                  return_acc: bool = False):
    """Filter one plane of a clip: frames [N, H, W] -> [N, H, W] (dtype of the input).

    Per output frame: temporal clamp, mirror pad, blocks at block_step, filter_blocks,
    overlap-add with the synthesis window, store (see module docstring).
    """
    n_frames, h, w = frames.shape
    bs, step, r = p.block_size, p.block_step, p.radius
    t_len = p.t_len
    scale = sample_scale(bits, is_float)
    peak = (1 << bits) - 1
    window = f32(p.window).reshape(t_len, bs, bs)
    syn = window[r] if syn_slice == "centre" else window[2 * r]
    ph, pw = pad_size(h, bs, step), pad_size(w, bs, step)
    ny, nx = pad_num(h, bs, step), pad_num(w, bs, step)
    oy, ox = (ph - h) // 2, (pw - w) // 2
    padded = np.stack([mirror_pad(frames[i].astype(np.float64), bs, step) for i in range(n_frames)])
    out = np.empty(frames.shape, dtype=frames.dtype)
    accs = []
    for n in range(n_frames):
        idx = [min(max(n - r + t, 0), n_frames - 1) for t in range(t_len)]
        slabs = []
        for t, fi in enumerate(idx):
            view = np.lib.stride_tricks.sliding_window_view(padded[fi], (bs, bs))[::step, ::step]
            slabs.append(view[:ny, :nx])
        blocks = np.stack(slabs, axis=2) * (scale * window)  # [ny, nx, T, B, B]
        yc = filter_blocks(blocks, p) * syn  # [ny, nx, B, B]
        acc = np.zeros((ph, pw))
        for dy in range(bs):
            for dx in range(bs):
                acc[dy:dy + step * ny:step, dx:dx + step * nx:step] += yc[:, :, dy, dx]
        core = acc[oy:oy + h, ox:ox + w]
        if return_acc:
            accs.append(core.copy())
        if is_float:
            out[n] = (core / scale).astype(frames.dtype)
        else:
            q = np.trunc(core / scale + 0.5)
            out[n] = np.clip(q, 0, peak).astype(frames.dtype)
    if return_acc:
        return out, np.stack(accs) / scale
    return out
