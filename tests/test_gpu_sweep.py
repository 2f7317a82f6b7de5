"""Seeded random parameter sweep of the fused kernel against the FP64 oracle (GPU).

Each case draws a sample format, precision, tbsize, sosize, window pair, zero-mean, filter
type and its parameters, and a small clip shape, then compares the plugin-path output with
oracle/dfttest_np.  What is asserted follows the filter type:

* continuous gains (filter types 0, 2, 4, 5, 6): the FP32 form within the parity bars (1 LSB,
  1e-5 for float32); the split-BF16 form within its bars for types 0, 2, 5, 6;
* hard decisions (type 1: psd against sigma; type 3: sigma or sigma2 by pmin <= psd <= pmax):
  a coefficient whose psd lies within the arithmetic's error of a threshold is kept or
  dropped differently than in FP64 (the reference's own FP32 arithmetic is exposed the same
  way), so the FP32 form is held to at least 95 % of samples within the bar (the share a
  flip touches grows with the overlap: 1.7 % at block step 2 in this sweep; the reference's
  own golden vectors for these types stay within 1 LSB).  The split-BF16
  form's absolute error, relative to the block's largest coefficients, flips such decisions
  far more often, and with type 4 (whose gain is steep near pmin) exceeds the bar on a small
  share of samples: those cases only have to run.
"""

import numpy as np
import pytest

from dfttest2_amd import synth
from dfttest2_amd.dfttest2 import Backend, VideoClip, filter_clip
from dfttest2_amd.params import dfttest_params
from oracle import dfttest_np as onp

pytestmark = pytest.mark.gpu


def draw_case(rng, i):
    fmt = str(rng.choice(["u8", "u10", "u12", "u16", "f32"]))
    prec = str(rng.choice(["fp32", "split_bf16"]))
    so = int(rng.choice([0, 2, 4, 6, 8, 12, 14, 15]))  # 10: overlap > 50 % with a non-divisor step
    ftype = int(rng.choice([0, 1, 2, 3, 4]))
    kw = dict(tbsize=int(rng.choice([1, 3, 5, 7])), sosize=so, ftype=ftype, swin=int(rng.integers(0, 12)),
              twin=int(rng.integers(0, 12)), zmean=bool(rng.integers(0, 2)))
    if ftype == 2:
        kw["sigma"] = float(rng.uniform(0.5, 1.5))
    elif ftype in (3, 4):
        kw["sigma"] = float(rng.uniform(1, 20))
        kw["pmin"] = float(rng.uniform(0.5, 5))
        kw["pmax"] = float(rng.uniform(50, 500))
    else:
        kw["sigma"] = float(rng.uniform(1, 30))
    if ftype == 0 and rng.integers(0, 3) == 0:
        kw["f0beta"] = float(rng.choice([0.5, 0.8]))  # filter types 6 and 5
    shape = (int(rng.integers(1, 7)), int(rng.integers(30, 70)), int(rng.integers(30, 80)))
    return fmt, prec, kw, shape


CASES = [draw_case(np.random.default_rng(5150 + i), i) for i in range(120)]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_random_parameters(case):
    fmt, prec, kw, shape = CASES[case]
    print(f"\n[case] sweep {case}: {fmt} {prec} {shape} {kw}", flush=True)
    bits = {"u8": 8, "u10": 10, "u12": 12, "u16": 16, "f32": 32}[fmt]
    is_float = fmt == "f32"
    p = dfttest_params(**kw)
    if p.zero_mean and abs(float(np.asarray(p.window_freq)[0])) == 0.0:
        pytest.skip("window spectrum DC is 0: zero-mean is undefined (the plan rejects it)")
    n, h, w = shape
    if is_float:
        x = synth.float_plane(n, h, w, 300 + case, 0.05).numpy()
    else:
        x = synth.int_plane(n, h, w, 300 + case, bits, {8: 6.0, 10: 12.0, 12: 50.0, 16: 700.0}[bits]).numpy()
    out = filter_clip(VideoClip([x], "float" if is_float else "int", bits), p, Backend.MI355X(precision=prec))
    ref = onp.process_plane(x, p, bits, is_float)
    diff = np.abs(out.planes[0].astype(np.float64) - ref.astype(np.float64))
    if prec == "fp32":
        bar = 1e-5 if is_float else 1.0
    else:
        bar = 5e-5 if is_float else (4.0 if bits > 12 else 1.0)
    hard = p.filter_type in (1, 3)
    if prec == "fp32" and not hard:
        assert diff.max() <= bar
    elif prec == "fp32":
        assert np.mean(diff > bar) <= 0.05, f"max {diff.max():.3g}"
    elif p.filter_type in (0, 2, 5, 6):
        assert diff.max() <= bar
    else:
        assert np.isfinite(out.planes[0].astype(np.float64)).all()
