"""Pins the oracle's hard-decision tools (oracle/dfttest_np.py: filter_blocks /
process_plane with `decisions`, hard_decisions, COEF_ERROR_BOUND) on the CPU.

* Replay semantics: replaying the FP64 decisions reproduces process_plane exactly; all
  coefficients kept (type 1) or all in range (type 3) is the fixed-gain filter (type 2)
  with gain 1, sigma or sigma2, which a swapped meaning of True would fail.
* Error margins: the block spectrum computed in two emulated arithmetics (FP32 matrix
  products, and hi+lo BF16 operands with three products and FP32 sums): every
  coefficient's psd must lie within the margin hard_decisions states for that form's
  bound (too small a bound fails here), and every decision the emulation takes otherwise
  than FP64 must be on a coefficient marked near; the split emulation must differ
  somewhere, so that the check has something to catch.
"""

import numpy as np
import pytest

from dfttest2_amd import synth
from dfttest2_amd.params import dfttest_params
from oracle import dfttest_np as onp


def clip(fmt, shape, seed):
    n, h, w = shape
    if fmt == "f32":
        return synth.float_plane(n, h, w, seed, 0.05).numpy(), 32, True
    bits = int(fmt[1:])
    return synth.int_plane(n, h, w, seed, bits, {8: 6.0, 10: 12.0, 16: 700.0}[bits]).numpy(), bits, False


@pytest.mark.parametrize("kw", [dict(ftype=1, sigma=4.0), dict(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16]),
                                dict(ftype=3, sigma=3.0, sigma2=0.5, pmin=2.0, pmax=60.0)])
def test_replaying_fp64_decisions_is_the_filter(kw):
    p = dfttest_params(**kw)
    x, bits, fl = clip("u8", (4, 40, 44), 3)
    dec = [d for d, *_ in onp.hard_decisions(x, p, bits, fl, onp.COEF_ERROR_BOUND["fp32"])]
    _, acc = onp.process_plane(x, p, bits, fl, return_acc=True)
    _, acc_r = onp.process_plane(x, p, bits, fl, return_acc=True, decisions=dec)
    np.testing.assert_array_equal(acc_r, acc)


def test_replay_meaning_of_true_and_false():
    x, bits, fl = clip("f32", (3, 36, 40), 4)
    _, _, ny, nx, _, _ = onp._geometry(x, dfttest_params())
    for tbsize in (1, 3):
        p1 = dfttest_params(ftype=1, sigma=5.0, tbsize=tbsize)
        shape = (ny, nx, p1.t_len, 16, 9)
        keep_all = [np.zeros(shape, bool)] * x.shape[0]
        ident = dfttest_params(ftype=2, sigma=1.0, tbsize=tbsize)
        np.testing.assert_allclose(onp.process_plane(x, p1, bits, fl, decisions=keep_all),
                                   onp.process_plane(x, ident, bits, fl), atol=1e-12)
        p3 = dfttest_params(ftype=3, sigma=1.7, sigma2=0.4, pmin=1.0, pmax=9.0, tbsize=tbsize)
        for flag, s in ((True, 1.7), (False, 0.4)):
            dec = [np.full(shape, flag)] * x.shape[0]
            np.testing.assert_allclose(onp.process_plane(x, p3, bits, fl, decisions=dec),
                                       onp.process_plane(x, dfttest_params(ftype=2, sigma=s, tbsize=tbsize), bits, fl),
                                       atol=1e-12)


def _bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _mm_fp32(a, b):
    return np.matmul(a.astype(np.float32), b.astype(np.float32))


def _mm_split(a, b):
    a = a.astype(np.float32)
    b = b.astype(np.float32)
    ah, bh = _bf16(a), _bf16(b)
    al, bl = _bf16(a - ah), _bf16(b - bh)
    return np.matmul(ah, bh) + np.matmul(al, bh) + np.matmul(ah, bl)


def emulated_decisions(x, p, bits, fl, mm):
    """Spectrum of every block by separable real-basis contractions in the arithmetic of
    `mm` (x transform, y transform, then the T-term temporal sums and the zero-mean step
    in FP32), and its type 1 / type 3 decisions."""
    bs, step, t_len = 16, p.block_step, p.t_len
    scale = onp.sample_scale(bits, fl)
    window = onp.f32(p.window).reshape(t_len, bs, bs)
    _, _, ny, nx, _, _ = onp._geometry(x, p)
    padded = np.stack([onp.mirror_pad(x[i].astype(np.float64), bs, step) for i in range(x.shape[0])])
    xs, ks = np.arange(16), np.arange(9)
    cx, sx = np.cos(2 * np.pi * np.outer(xs, ks) / 16), -np.sin(2 * np.pi * np.outer(xs, ks) / 16)
    ky = np.arange(16)
    cy, sy = np.cos(2 * np.pi * np.outer(ky, xs) / 16), -np.sin(2 * np.pi * np.outer(ky, xs) / 16)
    kt = np.arange(t_len)
    ct = np.cos(2 * np.pi * np.outer(kt, kt) / t_len).astype(np.float32)
    st = -np.sin(2 * np.pi * np.outer(kt, kt) / t_len).astype(np.float32)
    sigma = onp.f32(p.sigma_array()).astype(np.float32)
    out = []
    for n in range(x.shape[0]):
        b = onp._frame_blocks(padded, n, p, scale, window, ny, nx).astype(np.float32)
        re1, im1 = mm(b, cx), mm(b, sx)  # [..., T, y, kx]
        re2 = mm(cy, re1) - mm(sy, im1)
        im2 = mm(cy, im1) + mm(sy, re1)  # [..., T, ky, kx]
        re = np.einsum("kt,...tyx->...kyx", ct, re2) - np.einsum("kt,...tyx->...kyx", st, im2)
        im = np.einsum("kt,...tyx->...kyx", ct, im2) + np.einsum("kt,...tyx->...kyx", st, re2)
        re, im = re.astype(np.float32), im.astype(np.float32)
        if p.zero_mean:
            wf = onp.f32(p.window_freq).reshape(t_len, bs, 9, 2).astype(np.float32)
            gf = re[..., 0, 0, 0] / wf[0, 0, 0, 0]
            re = re - gf[..., None, None, None] * wf[..., 0]
            im = im - gf[..., None, None, None] * wf[..., 1]
        psd = re * re + im * im
        if p.filter_type == 1:
            out.append((psd < sigma, psd))
        else:
            out.append(((psd >= np.float32(p.pmin)) & (psd <= np.float32(p.pmax)), psd))
    return out


MARGIN_CASES = [
    ("u8", dict(ftype=1, sigma=4.0)),
    ("u10", dict(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])),
    ("u16", dict(ftype=1, sigma=3.0, sosize=8)),
    ("f32", dict(ftype=1, sigma=2.0, swin=4)),
    ("f32", dict(ftype=3, sigma=5.0, pmin=0.5, pmax=80.0)),
    ("u8", dict(ftype=3, sigma=2.0, sigma2=0.3, pmin=1.0, pmax=40.0, zmean=False)),
]


@pytest.mark.parametrize("form", ["fp32", "split_bf16"])
def test_margins_cover_emulated_arithmetic(form):
    mm = _mm_fp32 if form == "fp32" else _mm_split
    differing, worst = 0, 0.0
    for i, (fmt, kw) in enumerate(MARGIN_CASES):
        p = dfttest_params(**kw)
        x, bits, fl = clip(fmt, (4, 48, 56), 20 + i)
        fp64 = onp.hard_decisions(x, p, bits, fl, onp.COEF_ERROR_BOUND[form])
        emu = emulated_decisions(x, p, bits, fl, mm)
        for (dec, near, _, psd, margin), (d_emu, psd_emu) in zip(fp64, emu):
            ratio = np.abs(psd_emu.astype(np.float64) - psd) / margin
            worst = max(worst, float(ratio.max()))
            assert ratio.max() <= 1.0, f"case {i}: psd error {ratio.max():.3g} x the {form} margin"
            miss = dec != d_emu
            differing += int(miss.sum())
            assert not np.any(miss & ~near), f"case {i}: a decision outside the {form} margin differs"
    print(f"{form}: largest psd error / margin {worst:.3g}, {differing} decisions differ from FP64")
    if form == "split_bf16":
        assert differing > 0  # the emulated arithmetic does flip near-threshold decisions


@pytest.mark.parametrize("t_len", [1, 3, 5, 7])
def test_decision_record_layout_round_trip(t_len):
    # Every coefficient is decided by some lane (directly, or for kx = 0, 8 and kt > r by
    # its conjugate), coefficients held twice agree, and decoding inverts the layout.
    from dfttest2_amd.decisions import decode, lane_positions

    r = t_len // 2
    rng = np.random.default_rng(t_len)
    dec = rng.integers(0, 2, (2, 3, t_len, 16, 9)).astype(bool)
    for kt in range(r + 1, t_len):
        for ky in range(16):
            for kx in (0, 8):
                dec[..., kt, ky, kx] = dec[..., (t_len - kt) % t_len, (16 - ky) % 16, kx]
    masks = np.zeros((2, 3, r + 1, 4), np.uint64)
    held = np.zeros((t_len, 16, 9), int)
    for g, slot, lane, kt, ky, kx in lane_positions(t_len):
        masks[..., g, slot] |= dec[..., kt, ky, kx].astype(np.uint64) << np.uint64(lane)
        held[kt, ky, kx] += 1
    np.testing.assert_array_equal(decode(masks, t_len), dec)
    assert held[: r + 1].min() >= 1 and held[:, :, 1:8].min() >= 1
    assert held.max() <= 2
    bad = masks.copy()
    g, slot, lane, kt, ky, kx = next(t for t in lane_positions(t_len) if t[0] == 0 and t[5] not in (0, 8))
    bad[0, 0, g, slot] ^= np.uint64(1) << np.uint64(lane)  # one copy of a doubly held coefficient
    with pytest.raises(AssertionError, match="decided differently"):
        decode(bad, t_len)


@pytest.mark.parametrize("t_len,block_size", [(3, 8), (1, 9), (9, 16), (5, 12)])
def test_generic_decision_record_decoding(t_len, block_size):
    # Generic-path record: bit i of a block's mask sequence is coefficient i in
    # [kt][ky][kx] order; the last mask is padded with zero bits.
    from dfttest2_amd.decisions import decode_generic

    h = block_size // 2 + 1
    rng = np.random.default_rng(block_size)
    dec = rng.integers(0, 2, (2, 3, t_len, block_size, h)).astype(bool)
    flat = dec.reshape(2, 3, -1)
    words = (flat.shape[-1] + 63) // 64
    masks = np.zeros((2, 3, words), np.uint64)
    for i in range(flat.shape[-1]):
        masks[..., i // 64] |= flat[..., i].astype(np.uint64) << np.uint64(i % 64)
    np.testing.assert_array_equal(decode_generic(masks, t_len, block_size), dec)
