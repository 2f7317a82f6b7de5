"""The exact three-part BF16 split behind the FP32 path's first contraction (CPU).

With 8-bit samples the FP32 fused kernel runs the x rDFT of every slice on the BF16 MFMA:
each FP32 basis value b is split as b = hi + mid + lo, each part a BF16 value
(make_opnd3 in vs-dfttest2_amd/csrc/dft2_device.h: hi = round-to-nearest-even BF16 of b,
mid the same of b - hi, lo = (b - hi) - mid), and sample * part products are exact in
FP32.  These tests check the arithmetic facts that make the contraction as accurate as an
FP32 one, on the device's own operation sequence in float32: the remainders are computed
without rounding, lo is a BF16 value, the parts sum back to b exactly, and an 8-bit
integer times a part is exact.
"""

import numpy as np
import pytest


def bf16_rne(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), as float32 (finite inputs)."""
    b = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def is_bf16(x: np.ndarray) -> np.ndarray:
    return (x.astype(np.float32).view(np.uint32) & 0xFFFF) == 0


def split3(b: np.ndarray):
    b = b.astype(np.float32)
    hi = bf16_rne(b)
    r = (b - hi).astype(np.float32)
    mid = bf16_rne(r)
    lo_f = (r - mid).astype(np.float32)
    return hi, mid, lo_f, r


def basis_values() -> np.ndarray:
    """Real cos/sin basis of a 16-point transform times a few window shapes, as the
    separable-window bases hold them (scaled, both signs)."""
    n = np.arange(16)
    k = np.arange(16)
    ang = 2 * np.pi * np.outer(n, k) / 16
    base = np.concatenate([np.cos(ang).ravel(), np.sin(ang).ravel()])
    hann = 0.5 - 0.5 * np.cos(2 * np.pi * (n + 0.5) / 16)
    vals = [base]
    for w in (hann, np.ones(16), np.linspace(0.05, 1.0, 16)):
        for scale in (1.0, 1 / 16, 1 / np.sqrt(3), 255.0, 0.25):
            vals.append((np.outer(w, np.ones(16)).ravel().repeat(2) * base * scale))
    return np.concatenate(vals).astype(np.float32)


@pytest.mark.parametrize("source", ["basis", "random"])
def test_three_parts_sum_back_exactly(source):
    if source == "basis":
        b = basis_values()
    else:
        rng = np.random.default_rng(7)
        b = (rng.standard_normal(200_000) * np.exp2(rng.integers(-30, 30, 200_000))).astype(np.float32)
    hi, mid, lo, r = split3(b)
    b64 = b.astype(np.float64)
    # remainders computed in float32 are exact
    assert np.array_equal(r.astype(np.float64), b64 - hi.astype(np.float64))
    assert np.array_equal(lo.astype(np.float64), r.astype(np.float64) - mid.astype(np.float64))
    # every part is a BF16 value, so the device's packing of lo loses nothing
    assert is_bf16(hi).all() and is_bf16(mid).all() and is_bf16(lo).all()
    # and the three parts are b itself
    assert np.array_equal(hi.astype(np.float64) + mid.astype(np.float64) + lo.astype(np.float64), b64)


def test_products_with_8bit_samples_are_exact():
    b = basis_values()
    hi, mid, lo, _ = split3(b)
    s = np.arange(256, dtype=np.float32)
    for part in (hi, mid, lo):
        p32 = (s[:, None] * part[None, :]).astype(np.float32)
        p64 = s[:, None].astype(np.float64) * part[None, :].astype(np.float64)
        assert np.array_equal(p32.astype(np.float64), p64)


def test_a_dropped_part_is_detected():
    # the checks above fail for a two-part split (lo dropped): b != hi + mid for most b
    b = basis_values()
    hi, mid, _, _ = split3(b)
    assert not np.array_equal(hi.astype(np.float64) + mid.astype(np.float64), b.astype(np.float64))
