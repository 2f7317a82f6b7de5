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


def filter_blocks(blocks: np.ndarray, p, decisions: np.ndarray = None) -> np.ndarray:
    """fused() on windowed blocks [..., T, B, B] -> centre-slice spatial output [..., B, B].

    cpu_source/kernel.hpp:1899-1983: unnormalised forward 3-D real DFT, zero-mean,
    gain, mean re-add, inverse over t evaluated at the centre slice, inverse over y,
    C2R over x (Im at kx = 0 and kx = B/2 ignored).
    decisions: for filter types 1 and 3, boolean [..., T, B, B/2+1] replacing the
    comparisons of psd with the thresholds (type 1: True = dropped, psd < sigma; type 3:
    True = pmin <= psd <= pmax, gain sigma), so that another arithmetic's decisions can
    be replayed; the gains are otherwise those of `gain`.
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
    if decisions is not None and p.filter_type == 1:
        g = np.where(decisions, 0.0, 1.0)
    elif decisions is not None and p.filter_type == 3:
        g = np.where(decisions, sigma, float(f32(p.sigma2)))
    else:
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


def _frame_blocks(padded: np.ndarray, n: int, p, scale: float, window: np.ndarray, ny: int, nx: int) -> np.ndarray:
    """Windowed blocks [ny, nx, T, B, B] of output frame n: temporal clamp
    (cpu_source/getframe_impl.cpp:294-298), block origins at block_step, b = scale * w * p
    (cpu_source/getframe_impl.cpp:105-156)."""
    n_frames = padded.shape[0]
    bs, step, r = p.block_size, p.block_step, p.radius
    idx = [min(max(n - r + t, 0), n_frames - 1) for t in range(p.t_len)]
    slabs = []
    for fi in idx:
        view = np.lib.stride_tricks.sliding_window_view(padded[fi], (bs, bs))[::step, ::step]
        slabs.append(view[:ny, :nx])
    return np.stack(slabs, axis=2) * (scale * window)


def _geometry(frames: np.ndarray, p):
    _, h, w = frames.shape
    bs, step = p.block_size, p.block_step
    ph, pw = pad_size(h, bs, step), pad_size(w, bs, step)
    ny, nx = pad_num(h, bs, step), pad_num(w, bs, step)
    return ph, pw, ny, nx, (ph - h) // 2, (pw - w) // 2


def _overlap_add(values: np.ndarray, ph: int, pw: int, step: int) -> np.ndarray:
    """Sums per-block [ny, nx, B, B] values into the padded plane (block origins at step)."""
    ny, nx, bs, _ = values.shape
    acc = np.zeros((ph, pw))
    for dy in range(bs):
        for dx in range(bs):
            acc[dy:dy + step * ny:step, dx:dx + step * nx:step] += values[:, :, dy, dx]
    return acc


def process_plane(frames: np.ndarray, p, bits: int, is_float: bool, syn_slice: str = "centre",
                  return_acc: bool = False, decisions=None):
    """Filter one plane of a clip: frames [N, H, W] -> [N, H, W] (dtype of the input).

    Per output frame: temporal clamp, mirror pad, blocks at block_step, filter_blocks,
    overlap-add with the synthesis window, store (see module docstring).
    decisions: per output frame, the [ny, nx, T, B, B/2+1] hard decisions filter_blocks
    replays (filter types 1 and 3), or None.
    """
    n_frames, h, w = frames.shape
    bs, step, r = p.block_size, p.block_step, p.radius
    t_len = p.t_len
    scale = sample_scale(bits, is_float)
    peak = (1 << bits) - 1
    window = f32(p.window).reshape(t_len, bs, bs)
    syn = window[r] if syn_slice == "centre" else window[2 * r]
    ph, pw, ny, nx, oy, ox = _geometry(frames, p)
    padded = np.stack([mirror_pad(frames[i].astype(np.float64), bs, step) for i in range(n_frames)])
    out = np.empty(frames.shape, dtype=frames.dtype)
    accs = []
    for n in range(n_frames):
        blocks = _frame_blocks(padded, n, p, scale, window, ny, nx)  # [ny, nx, T, B, B]
        yc = filter_blocks(blocks, p, None if decisions is None else decisions[n]) * syn  # [ny, nx, B, B]
        acc = _overlap_add(yc, ph, pw, step)
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


# Error bound of one block DFT coefficient as the GPU forms compute it, as a multiple of
# the block's L1 norm sum|b| (which bounds every coefficient).  With u = 2^-24 (FP32) and
# u' = 2^-9 (BF16):
#   fp32: each of the forward stages (x contraction with exact products or FP32 MFMA, y
#     contraction on the FP32 MFMA, T-term temporal sums) adds at most (n - 1) u of the
#     sum of its |terms|, and the FP32 bases add u each: about 40 u = 2^-18.7 < 2^-18.
#   split_bf16: a value a = hi + lo + e with |e| <= u'^2 |a| = 2^-18 |a|, and a product
#     drops lo_a lo_b (<= 2^-18 |ab|): at most 3 * 2^-18 per split contraction, two
#     contractions, plus the FP32 sums: 6 * 2^-18 + 40 u = 2^-15.3 < 2^-15.
COEF_ERROR_BOUND = {"fp32": 2.0 ** -18, "split_bf16": 2.0 ** -15}


def hard_decisions(frames: np.ndarray, p, bits: int, is_float: bool, kappa: float):
    """FP64 hard decisions of filter types 1 and 3 and where another arithmetic may differ.

    Returns, per output frame, (decisions, near, closeness, psd, margin), each
    [ny, nx, T, B, B/2+1]:
    decisions as filter_blocks takes them (type 1: psd < sigma; type 3: pmin <= psd <=
    pmax, cpu_source/kernel.hpp:1826-1890), with psd = |X'|^2 after the zero-mean step;
    near: the coefficient's psd lies within 2 |X'| d + d^2 of a threshold, where d bounds
    the coefficient error of a form whose error is at most kappa * sum|b| per block
    (COEF_ERROR_BOUND), times 1 + |WF|/|WF0| for the zero-mean step (it subtracts gf * WF
    with gf = X0/WF0); closeness: |psd - threshold| over that margin (<= 1 where near);
    psd and margin themselves.  Only near coefficients can be decided otherwise by that
    form.
    """
    n_frames = frames.shape[0]
    if p.filter_type not in (1, 3):
        raise ValueError("hard decisions exist for filter types 1 and 3")
    bs, step, t_len = p.block_size, p.block_step, p.t_len
    scale = sample_scale(bits, is_float)
    window = f32(p.window).reshape(t_len, bs, bs)
    ph, pw, ny, nx, oy, ox = _geometry(frames, p)
    padded = np.stack([mirror_pad(frames[i].astype(np.float64), bs, step) for i in range(n_frames)])
    sigma = f32(p.sigma_array())
    res = []
    for n in range(n_frames):
        blocks = _frame_blocks(padded, n, p, scale, window, ny, nx)
        spec = np.fft.rfftn(blocks, axes=(-3, -2, -1))
        d = kappa * np.abs(blocks).sum(axis=(-3, -2, -1))[..., None, None, None]
        if p.zero_mean:
            wf = f32(p.window_freq).reshape(t_len, bs, bs // 2 + 1, 2)
            wf = wf[..., 0] + 1j * wf[..., 1]
            spec = spec - (spec[..., 0, 0, 0].real / wf[0, 0, 0].real)[..., None, None, None] * wf
            d = d * (1.0 + np.abs(wf) / abs(wf[0, 0, 0].real))
        mag = np.abs(spec)
        psd = mag * mag
        margin = 2.0 * mag * d + d * d
        if p.filter_type == 1:
            dec = psd < sigma
            dist = np.abs(psd - sigma)
        else:
            lo, hi = float(f32(p.pmin)), float(f32(p.pmax))
            dec = (psd >= lo) & (psd <= hi)
            dist = np.minimum(np.abs(psd - lo), np.abs(psd - hi))
        closeness = dist / np.maximum(margin, 1e-300)
        res.append((dec, closeness <= 1.0, closeness, psd, margin))
    return res
