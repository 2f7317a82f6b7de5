"""Pins the FP64 oracle (oracle/dfttest_np.py) before it is trusted as the checker.

Checks, each chosen so that a plausible slip (sign of an exponent, a dropped C2R
rule, a wrong padding offset, a wrong synthesis slice) fails one of them:
  * padding sizes / block counts against the values derived for the benchmark sizes;
  * mirror padding against numpy.pad(mode="reflect");
  * filter_blocks against a brute-force textbook 3-D DFT / C2R written with explicit sums;
  * identity (ftype 2, sigma 1) and constant-frame invariants;
  * the golden vectors made by the reference's own fused() (tests/golden/);
  * the reference harness itself (oracle/_ref) where it is built.
"""

import numpy as np
import pytest

from dfttest2_amd.params import dfttest_params
from golden_util import fixture_names, load_fixture, max_err
from oracle import dfttest_np as onp
from oracle import ref_lib


@pytest.mark.parametrize(
    "size,pad,num",
    [(1920, 1944, 483), (1080, 1112, 275), (960, 984, 243), (540, 568, 139), (3840, 3864, 963),
     (2160, 2184, 543), (640, 664, 163), (480, 504, 123), (256, 280, 67), (64, 88, 19)],
)
def test_pad_sizes_at_benchmark_shapes(size, pad, num):
    assert onp.pad_size(size, 16, 4) == pad
    assert onp.pad_num(size, 16, 4) == num


def test_pad_num_covers_padded_plane():
    for size in range(20, 90):
        for step in (1, 2, 3, 4, 8, 12, 16):
            ps, n = onp.pad_size(size, 16, step), onp.pad_num(size, 16, step)
            assert (n - 1) * step + 16 <= ps < n * step + 16


@pytest.mark.parametrize("h,w,step", [(29, 37, 4), (40, 33, 8), (30, 30, 2), (45, 52, 16)])
def test_mirror_pad_matches_numpy_reflect(h, w, step):
    rng = np.random.default_rng(0)
    f = rng.integers(0, 255, (h, w))
    ph, pw = onp.pad_size(h, 16, step), onp.pad_size(w, 16, step)
    oy, ox = (ph - h) // 2, (pw - w) // 2
    ref = np.pad(f, ((oy, ph - h - oy), (ox, pw - w - ox)), mode="reflect")
    np.testing.assert_array_equal(onp.mirror_pad(f, 16, step), ref)


def brute_force_block(blk, p):
    """Textbook transform of one windowed block with explicit sums (no FFT library)."""
    T, B = p.t_len, p.block_size
    H = B // 2 + 1
    r = p.radius
    t = np.arange(T)[:, None, None]
    y = np.arange(B)[None, :, None]
    x = np.arange(B)[None, None, :]
    X = np.zeros((T, B, H), dtype=complex)
    for kt in range(T):
        for ky in range(B):
            for kx in range(H):
                ph = np.exp(-2j * np.pi * (kt * t / T + ky * y / B + kx * x / B))
                X[kt, ky, kx] = np.sum(blk * ph)
    sigma = onp.f32(p.sigma_array())
    if p.zero_mean:
        wf = onp.f32(p.window_freq).reshape(T, B, H, 2)
        wf = wf[..., 0] + 1j * wf[..., 1]
        gf = X[0, 0, 0].real / wf[0, 0, 0].real
        X = X - gf * wf
    psd = np.abs(X) ** 2
    Y = X * onp.gain(psd, sigma, onp.f32(p.sigma2), onp.f32(p.pmin), onp.f32(p.pmax), p.filter_type)
    if p.zero_mean:
        Y = Y + gf * wf
    # Inverse over t at the centre slice, inverse over y, C2R over x: the real part of
    # each kx column with weight 1 at kx = 0 and (B even) kx = B/2, whose imaginary parts
    # are dropped, and 2 otherwise.
    out = np.zeros((B, B))
    yy = np.arange(B)[:, None]
    xx = np.arange(B)[None, :]
    for kt in range(T):
        for ky in range(B):
            for kx in range(H):
                zy = Y[kt, ky, kx] * np.exp(2j * np.pi * (kt * r / T + ky * yy / B))
                if kx == 0 or 2 * kx == B:
                    out += zy.real * np.cos(2 * np.pi * kx * xx / B)
                else:
                    out += 2.0 * (zy * np.exp(2j * np.pi * kx * xx / B)).real
    return out


@pytest.mark.parametrize(
    "kw",
    [dict(), dict(tbsize=1), dict(tbsize=5, ftype=1), dict(ftype=3, pmin=100.0, pmax=1e5), dict(ftype=4, pmin=3.0),
     dict(f0beta=0.7), dict(f0beta=0.5, zmean=False), dict(ftype=2, sigma=0.3)],
)
def test_filter_blocks_matches_textbook_dft(kw):
    p = dfttest_params(**kw)
    rng = np.random.default_rng(5)
    T, B = p.t_len, p.block_size
    blk = rng.integers(0, 256, (T, B, B)) * onp.f32(p.window).reshape(T, B, B)
    np.testing.assert_allclose(onp.filter_blocks(blk, p), brute_force_block(blk, p), rtol=0, atol=1e-8)


@pytest.mark.parametrize("kw", [dict(sbsize=8, sosize=4), dict(sbsize=9, sosize=3, ftype=1, sigma=2.0),
                                dict(sbsize=12, sosize=8, tbsize=5, ftype=3, pmin=1.0, pmax=300.0),
                                dict(sbsize=24, sosize=18, tbsize=1, f0beta=0.5),
                                dict(sbsize=8, sosize=4, tbsize=9, swin=2, twin=0, ftype=4, pmin=2.0)])
def test_filter_blocks_matches_textbook_dft_other_block_sizes(kw):
    # Odd block sizes have no Nyquist column; tbsize 9 has four temporal pairs.
    test_filter_blocks_matches_textbook_dft(kw)


@pytest.mark.parametrize("sb,so,tb", [(8, 4, 3), (9, 6, 1), (12, 8, 5), (24, 16, 3), (8, 4, 9)])
def test_identity_other_block_sizes(sb, so, tb):
    # ftype 2 with sigma 1 returns the plane for any block size and overlap (the window is
    # normalised per residue of the block step), as at 16.
    p = dfttest_params(ftype=2, sigma=1.0, sbsize=sb, sosize=so, tbsize=tb)
    rng = np.random.default_rng(sb + tb)
    x = rng.integers(0, 256, (4, 37, 41)).astype(np.uint8)
    np.testing.assert_array_equal(onp.process_plane(x, p, 8, False), x)


@pytest.mark.parametrize("tbsize", [1, 3, 5, 7])
def test_unit_gain_inverse_is_scaled_centre_slice(tbsize):
    # With gain 1 the unnormalised inverse returns N * b[r] (N = T * 256) for any block.
    p = dfttest_params(ftype=2, sigma=1.0, zmean=False, tbsize=tbsize)
    rng = np.random.default_rng(9)
    blk = rng.normal(size=(p.t_len, 16, 16))
    np.testing.assert_allclose(onp.filter_blocks(blk, p), p.t_len * 256 * blk[p.radius], rtol=0, atol=1e-9)


@pytest.mark.parametrize("bits,dtype", [(8, np.uint8), (10, np.uint16), (16, np.uint16)])
def test_identity_ftype2_sigma1(bits, dtype):
    # ftype 2 (multiplier) with sigma 1 is not rescaled by wscale; with a rectangular
    # temporal window, sum of w^2 over overlapping blocks * N = 1, so the plane returns.
    p = dfttest_params(ftype=2, sigma=1.0)
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1 << bits, (4, 37, 41)).astype(dtype)
    np.testing.assert_array_equal(onp.process_plane(x, p, bits, False), x)


def test_constant_frames_pass_through():
    for kw in (dict(), dict(ftype=1, tbsize=5), dict(ftype=4, pmin=1.0)):
        p = dfttest_params(**kw)
        x = np.full((3, 30, 34), 117, dtype=np.uint8)
        np.testing.assert_array_equal(onp.process_plane(x, p, 8, False), x)


@pytest.mark.parametrize("name", fixture_names())
def test_oracle_reproduces_golden_vectors(name):
    p, x, y, bits, is_float = load_fixture(name)
    out = onp.process_plane(x, p, bits, is_float)
    if is_float:
        assert max_err(out, y) <= 1e-5
    else:
        assert max_err(out, y) <= 1


@pytest.mark.skipif(not ref_lib.available(), reason="reference harness (oracle/_ref) not built here")
@pytest.mark.parametrize("kw", [dict(), dict(tbsize=7, twin=0), dict(ftype=1, tbsize=5), dict(sosize=8, ftype=3)])
def test_oracle_matches_reference_harness(kw):
    p = dfttest_params(**kw)
    rng = np.random.default_rng(4)
    x = rng.integers(0, 256, (5, 35, 43)).astype(np.uint8)
    assert max_err(onp.process_plane(x, p, 8, False), ref_lib.filter_plane(x, p, 8, False)) <= 1
    xf = rng.random((5, 35, 43)).astype(np.float32)
    assert max_err(onp.process_plane(xf, p, 32, True), ref_lib.filter_plane(xf, p, 32, True)) <= 1e-5


@pytest.mark.skipif(not ref_lib.available(), reason="reference harness (oracle/_ref) not built here")
def test_reference_cpu_synthesis_slice_differs_for_hanning_twin():
    # The CPU backend weights the overlap-add with slice 2r of the window, the GPU
    # backends with slice r; with a Hanning temporal window the results differ, and the
    # oracle's syn_slice="cpu" switch reproduces the CPU backend.
    p = dfttest_params(tbsize=3, twin=0)
    rng = np.random.default_rng(2)
    x = rng.integers(0, 256, (4, 33, 33)).astype(np.uint8)
    cpu = ref_lib.filter_plane(x, p, 8, False, syn_slice="cpu")
    assert max_err(onp.process_plane(x, p, 8, False, syn_slice="cpu"), cpu) <= 1
    assert max_err(onp.process_plane(x, p, 8, False), cpu) > 10
