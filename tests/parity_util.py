"""Shared GPU parity checks: the bars per precision, and the decision-replay check of the
hard-threshold filters (filter types 1 and 3) against the FP64 oracle."""

import numpy as np

from dfttest2_amd.decisions import run_with_decisions
from dfttest2_amd.dfttest2 import Backend, VideoClip, filter_clip, get_plan
from oracle import dfttest_np as onp


def bar_of(prec, bits, is_float, p):
    if prec == "fp32":
        return 1e-5 if is_float else 1.0
    bar = 5e-5 if is_float else (4.0 if bits > 12 else 1.0)
    if p.filter_type == 2:
        bar *= max(1.0, float(np.max(np.abs(p.sigma))))
    return bar


def check_hard_decisions(x, p, bits, is_float, prec, bar, label):
    """Decision-replay check of one clip (filter types 1 and 3); returns the output."""
    plan = get_plan(p, "float" if is_float else "int", bits, Backend.MI355X(precision=prec))
    out, dec_gpu = run_with_decisions(plan, x, p)
    # The product path (for tbsize >= 5 the cross-frame form, which computes every slice
    # spectrum and decision with the same code, on other tiles) must match the replay too.
    plain = filter_clip(VideoClip([x], "float" if is_float else "int", bits), p,
                        Backend.MI355X(precision=prec)).planes[0]
    fp64 = onp.hard_decisions(x, p, bits, is_float, onp.COEF_ERROR_BOUND[prec])
    flips, worst_close = 0, 0.0
    for n, ((dec, near, close, _, _), dg) in enumerate(zip(fp64, dec_gpu)):
        miss = dec != dg
        flips += int(miss.sum())
        if miss.any():
            worst_close = max(worst_close, float(close[miss].max()))
        assert not np.any(miss & ~near), f"frame {n}: a decision outside the {prec} error margin differs from FP64"
    replay = onp.process_plane(x, p, bits, is_float, decisions=dec_gpu).astype(np.float64)
    d_replay = np.abs(out.astype(np.float64) - replay)
    d_plain = np.abs(plain.astype(np.float64) - replay)
    d_fp64 = np.abs(plain.astype(np.float64) - onp.process_plane(x, p, bits, is_float).astype(np.float64))
    print(f"[{label}] decisions differing from FP64: {flips} (largest |psd - threshold| / margin "
          f"{worst_close:.3g}); samples beyond the bar vs plain FP64: {np.mean(d_fp64 > bar):.4%} "
          f"(max {d_fp64.max():.3g}); vs replay: recording run max {d_replay.max():.3g}, "
          f"product path max {d_plain.max():.3g}", flush=True)
    assert d_replay.max() <= bar
    assert d_plain.max() <= bar
    return plain
