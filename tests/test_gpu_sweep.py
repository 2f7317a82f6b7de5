"""Seeded random parameter sweep of the fused kernel against the FP64 oracle (GPU).

Each case draws a sample format, precision, tbsize, sosize, window pair, zero-mean, filter
type and its parameters, and a small clip shape, then compares the plugin-path output with
oracle/dfttest_np on every sample.  Bars: FP32 1 LSB (integers) and 1e-5 (float32);
split BF16 1 LSB up to 12 bits, 4 LSB at 16 bits, 5e-5 for float32, times sigma for the
fixed-gain filter (type 2 multiplies the coefficients, and their error, by sigma).

* Continuous gains (filter types 0, 2, 4, 5, 6): every sample within the bar of FP64.
* Hard decisions (type 1: psd < sigma drops a coefficient; type 3: sigma or sigma2 by
  pmin <= psd <= pmax): the kernel records every block's decisions
  (dft2_process_batch_decisions).  Every decision that differs from FP64 must be on a
  coefficient whose FP64 psd lies within the form's error margin of its threshold
  (oracle hard_decisions), and the FP64 oracle replaying the kernel's own decisions must
  match the output on every sample within the bar.  The share of samples that differ from
  plain FP64 by more than the bar (those near-threshold decisions) is printed.
* Split BF16 with types 3 and 4 (gains above 1, type 3 jumping between them) is rejected
  at plan creation: the case checks the error.
"""

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import Backend, VideoClip, filter_clip, get_plan
from dfttest2_amd.params import dfttest_params
from parity_util import bar_of, check_hard_decisions
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
    if prec == "split_bf16" and p.filter_type in (3, 4):
        with pytest.raises(_lib.Dft2Error, match="split BF16"):
            get_plan(p, "float" if is_float else "int", bits, Backend.MI355X(precision=prec))
        return
    n, h, w = shape
    if is_float:
        x = synth.float_plane(n, h, w, 300 + case, 0.05).numpy()
    else:
        x = synth.int_plane(n, h, w, 300 + case, bits, {8: 6.0, 10: 12.0, 12: 50.0, 16: 700.0}[bits]).numpy()
    bar = bar_of(prec, bits, is_float, p)
    if p.filter_type in (1, 3):
        check_hard_decisions(x, p, bits, is_float, prec, bar, f"sweep {case}")
        return
    out = filter_clip(VideoClip([x], "float" if is_float else "int", bits), p, Backend.MI355X(precision=prec))
    ref = onp.process_plane(x, p, bits, is_float)
    assert np.abs(out.planes[0].astype(np.float64) - ref.astype(np.float64)).max() <= bar


# The precision workload's filter (YUV420P10, ftype 1, tbsize 5, sigma curve 0.0:4 0.5:8
# 1.0:16) at more seeds and sizes, both precisions, by decision replay.
CONFIG_PRECISION_CASES = [(prec, shape, seed) for prec in ("fp32", "split_bf16")
                          for shape, seed in (((6, 120, 160), 901), ((4, 200, 264), 902), ((3, 360, 640), 903),
                                              ((8, 90, 130), 904))]


@pytest.mark.parametrize("prec,shape,seed", CONFIG_PRECISION_CASES)
def test_precision_workload_filter_by_decision_replay(prec, shape, seed):
    print(f"\n[case] u10 ftype 1 tbsize 5 sigma curve {prec} {shape} seed {seed}", flush=True)
    p = dfttest_params(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])
    n, h, w = shape
    x = synth.int_plane(n, h, w, seed, 10, 12.0).numpy()
    check_hard_decisions(x, p, 10, False, prec, 1.0, f"{prec} {shape}")
