"""Parameter layer: windows, sigma arrays, wscale and RDFT (CPU only).

Pins: scipy.signal window routines for the cosine-sum / Bartlett windows sampled at
i + 0.5, scipy.special.i0 for Kaiser-Bessel, the fixture written from the reference's
own helpers for the windows nothing else pins (tests/golden/windows.npz), closed forms
for wscale and the config-2 sigma curve, numpy.fft.rfftn for RDFT.
"""

import math
import os

import numpy as np
import pytest
import scipy.signal.windows as sw
import scipy.special

from dfttest2_amd import helpers
from dfttest2_amd.params import dfttest_params, flatten, rdft, to_func

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def half_samples(full):
    # scipy windows of length 2N (periodic) or 2N+1 sampled at odd indices are our
    # windows sampled at i + 0.5 on [0, N].
    return np.asarray(full)[1::2]


@pytest.mark.parametrize("n", [1, 3, 5, 7, 16])
def test_cosine_windows_match_scipy(n):
    ours = lambda mode: np.array([helpers.get_window_value(i + 0.5, n, mode, 2.5) for i in range(n)])
    np.testing.assert_allclose(ours(0), half_samples(sw.hann(2 * n, sym=False)), atol=1e-15)
    np.testing.assert_allclose(ours(1), half_samples(sw.general_hamming(2 * n, 0.53836, sym=False)), atol=1e-15)
    np.testing.assert_allclose(ours(2), half_samples(sw.blackman(2 * n, sym=False)), atol=1e-15)
    np.testing.assert_allclose(ours(3), half_samples(sw.blackmanharris(2 * n, sym=False)), atol=1e-15)
    np.testing.assert_allclose(ours(11), half_samples(sw.nuttall(2 * n, sym=False)), atol=1e-15)
    np.testing.assert_allclose(ours(8), half_samples(sw.bartlett(2 * n + 1)), atol=1e-15)
    assert np.all(ours(7) == 1.0)


@pytest.mark.parametrize("beta", [1.0, 2.5])
def test_kaiser_matches_scipy_i0(beta):
    # The reference's I0 series stops after 15 terms, which converges for these betas;
    # for large beta (e.g. 6) it deviates by about 1% and only the vectors from the
    # reference's own helpers (test_windows_match_reference_helpers_vectors) pin it.
    n = 16
    ours = np.array([helpers.get_window_value(i + 0.5, n, 4, beta) for i in range(n)])
    v = 2 * (np.arange(n) + 0.5) / n - 1
    ref = scipy.special.i0(np.pi * beta * np.sqrt(1 - v * v)) / scipy.special.i0(np.pi * beta)
    np.testing.assert_allclose(ours, ref, rtol=2e-7)


def test_windows_match_reference_helpers_vectors():
    rows = np.load(os.path.join(GOLDEN, "windows.npz"), allow_pickle=False)["rows"]
    for mode, size, beta, loc, val in rows:
        assert helpers.get_window_value(loc, int(size), int(mode), beta) == pytest.approx(val, rel=0, abs=1e-15)


@pytest.mark.parametrize("step", [1, 2, 4, 8, 12, 16])
@pytest.mark.parametrize("swin", [0, 1, 2, 4, 8, 10])
def test_normalize_residue_classes_have_unit_energy(step, swin):
    ws = helpers.normalize([helpers.get_window_value(i + 0.5, 16, swin, 2.5) for i in range(16)], 16, step)
    for q in range(step):
        assert sum(ws[h] ** 2 for h in range(q, 16, step)) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("tbsize", [1, 3, 5, 7])
@pytest.mark.parametrize("swin", [0, 2, 5, 11])
def test_wscale_is_one_sixteenth_at_step4(tbsize, swin):
    # sum w^2 = S^2 * sum wt^2 / (T * 256) = 1/16 for S = 4 and a rectangular twin.
    p = dfttest_params(tbsize=tbsize, swin=swin)
    assert p.extra["wscale"] == pytest.approx(1.0 / 16.0, abs=1e-14)
    assert p.sigma[0] == pytest.approx(8.0 / 16.0, abs=1e-14)


def test_location_of_x_axis_uses_length_nine():
    # kx = 5..8 map to 1.0, 0.75, 0.5, 0.25 because the x axis length is B/2+1 = 9.
    assert [helpers.get_location(k, 9) for k in range(9)] == [0, 0.25, 0.5, 0.75, 1.0, 1.0, 0.75, 0.5, 0.25]
    assert [helpers.get_location(k, 3) for k in range(3)] == [0.0, 1.0, 1.0]


def test_sigma_curve_of_precision_config():
    p = dfttest_params(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])
    s = p.sigma.reshape(5, 16, 9) * 16.0  # undo wscale = 1/16
    assert s[0, 0, 0] == pytest.approx(4.0, rel=1e-12)
    # kx = 5 sits at location 1.0: cbrt(4) * cbrt(4) * cbrt(16)
    assert s[0, 0, 5] == pytest.approx((4 * 4 * 16) ** (1 / 3), rel=1e-12)
    # ky = 8 is location 1.0 on the y axis, kt = 2 is location 1.0 on the t axis (T = 5)
    assert s[2, 8, 4] == pytest.approx(16.0, rel=1e-12)
    # symmetric in t and y
    np.testing.assert_array_equal(s[1], s[4])
    np.testing.assert_array_equal(s[:, 3], s[:, 13])


def test_to_func_interpolates_and_rejects_out_of_range():
    f = to_func(flatten([(0.0, 1.0), (1.0, 10.0)]), lambda v: v, 5.0)
    assert f(0.5) == pytest.approx(5.5)
    assert f(1.0) == pytest.approx(10.0)
    with pytest.raises(ValueError):
        f(1.5)
    assert flatten([0.0, 1.0, 1.0, 10.0]) == [0.0, 1.0, 1.0, 10.0]


def test_filter_type_translation():
    assert dfttest_params(f0beta=1.0).filter_type == 0
    assert dfttest_params(f0beta=0.5).filter_type == 6
    p = dfttest_params(f0beta=0.7)
    assert p.filter_type == 5 and p.pmin == pytest.approx(0.7)
    assert dfttest_params(ftype=3).filter_type == 3


@pytest.mark.parametrize("bad", [dict(ftype=5), dict(sbsize=0), dict(smode=0), dict(sosize=13), dict(tbsize=0),
                                 dict(swin=12), dict(twin=-1), dict(nlocation=[1]), dict(slocation=[0.0, 1.0, 1.0]),
                                 dict(ssystem=2), dict(sbsize=8, sosize=5)])
def test_argument_errors(bad):
    with pytest.raises(ValueError):
        dfttest_params(**bad)


@pytest.mark.parametrize("kw", [dict(sbsize=8, sosize=4), dict(sbsize=12, sosize=8, tbsize=1), dict(tbsize=9),
                                dict(sbsize=24, sosize=18, tbsize=5), dict(sbsize=9, sosize=3)])
def test_other_block_sizes_and_radii(kw):
    # The reference's DFTTest() takes any sbsize >= 1 and tbsize >= 1 (python/dfttest2.py:
    # 176-190); the layouts scale with them: window [T][B][B], sigma [T][B][B/2+1].
    p = dfttest_params(**kw)
    B, T = kw.get("sbsize", 16), kw.get("tbsize", 3)
    assert p.block_size == B and p.t_len == T and p.block_step == B - kw.get("sosize", 12)
    assert np.asarray(p.window).size == T * B * B
    assert np.asarray(p.window_freq).size == T * B * (B // 2 + 1) * 2
    # wscale: the window's sum of squares, S^2 sum(wt^2) / (T B^2) with a rectangular twin
    assert p.extra["wscale"] == pytest.approx(p.block_step ** 2 / B ** 2, rel=1e-12)


@pytest.mark.parametrize("shape", [(16,), (16, 16), (3, 16, 16), (5, 16, 16), (4, 6, 10)])
def test_rdft_matches_numpy_rfftn(shape):
    rng = np.random.default_rng(3)
    data = rng.normal(size=shape)
    spec = np.fft.rfftn(data)
    ours = rdft(data.reshape(-1), shape).reshape(spec.shape + (2,))
    np.testing.assert_allclose(ours[..., 0], spec.real, atol=1e-10)
    np.testing.assert_allclose(ours[..., 1], spec.imag, atol=1e-10)


def test_window_freq_dc_is_255_sum_of_window():
    p = dfttest_params()
    assert p.window_freq[0] == pytest.approx(255 * math.fsum(p.window), rel=1e-12)
