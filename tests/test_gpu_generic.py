"""The generic block path on the GPU (block sizes other than 16, tbsize above 7), through the
C ABI, against the FP64 oracle, which handles any block size (pinned on the CPU by the
textbook DFT at block sizes 8, 9, 12 and 24 and tbsize 9, tests/test_oracle.py).  The
reference serves these sizes only with its FFT-library backends, which need VapourSynth and
a GPU runtime of their own, so the oracle is the yardstick here.

Bars as everywhere: integer formats within 1 LSB, float32 within 1e-5.  Hard-threshold
filters by decision replay (tests/parity_util.py)."""

import sys

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import Backend, VideoClip, filter_clip, get_plan, plane_job
from dfttest2_amd.params import dfttest_params
from oracle import dfttest_np as onp
from parity_util import check_hard_decisions

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def make_plane(fmt, shape, seed):
    n, h, w = shape
    if fmt == "f32":
        return synth.float_plane(n, h, w, seed, 0.05).numpy(), 32, True
    bits = int(fmt[1:])
    return synth.int_plane(n, h, w, seed, bits, {8: 6.0, 10: 12.0, 16: 700.0}[bits]).numpy(), bits, False


CASES = [
    ("u8", (5, 50, 60), dict(sbsize=8, sosize=4)),
    ("f32", (4, 45, 52), dict(sbsize=12, sosize=8, tbsize=5)),
    ("u10", (4, 40, 48), dict(sbsize=9, sosize=3, tbsize=1, f0beta=0.5)),
    ("u16", (3, 70, 66), dict(sbsize=24, sosize=18, ftype=2, sigma=0.8)),
    ("u8", (3, 80, 90), dict(sbsize=32, sosize=24, ftype=4, sigma=2.0, pmin=2.0, pmax=300.0)),
    ("f32", (11, 40, 40), dict(tbsize=9)),                               # block size 16, radius 4
    ("u8", (9, 44, 50), dict(sbsize=8, sosize=4, tbsize=15, swin=1, twin=0)),  # radius 7
    ("u8", (4, 39, 41), dict(sbsize=6, sosize=2, zmean=False, f0beta=0.8)),
    ("f32", (3, 36, 40), dict(sbsize=5, sosize=0, tbsize=3, swin=2)),     # odd, no overlap
    ("u8", (3, 100, 120), dict(sbsize=48, sosize=36, tbsize=1)),
]


@pytest.mark.parametrize("fmt,shape,kw", CASES, ids=[f"{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_generic_block_path_vs_oracle(fmt, shape, kw):
    announce("generic", fmt, shape, kw)
    p = dfttest_params(**kw)
    x, bits, is_float = make_plane(fmt, shape, seed=400 + len(kw))
    out = filter_clip(VideoClip([x], "float" if is_float else "int", bits), p).planes[0]
    ref = onp.process_plane(x, p, bits, is_float)
    assert np.abs(out.astype(np.float64) - ref.astype(np.float64)).max() <= (1e-5 if is_float else 1.0)


HARD = [
    ("u8", (4, 48, 56), dict(sbsize=8, sosize=4, ftype=1, sigma=6.0)),
    ("f32", (4, 44, 50), dict(sbsize=12, sosize=8, ftype=3, sigma=2.0, sigma2=0.5, pmin=1.0, pmax=200.0)),
    ("u10", (10, 40, 44), dict(tbsize=9, ftype=1, slocation=[0.0, 4, 0.5, 8, 1.0, 16])),
]


@pytest.mark.parametrize("fmt,shape,kw", HARD, ids=[f"{c[0]}-{i}" for i, c in enumerate(HARD)])
def test_generic_hard_thresholds_by_decision_replay(fmt, shape, kw):
    announce("generic hard decisions", fmt, shape, kw)
    p = dfttest_params(**kw)
    x, bits, is_float = make_plane(fmt, shape, seed=500 + len(kw))
    check_hard_decisions(x, p, bits, is_float, "fp32", 1e-5 if is_float else 1.0, f"generic {fmt}")


@pytest.mark.parametrize("sb,so,tb", [(8, 4, 3), (12, 6, 5), (9, 3, 1)])
def test_generic_identity_is_exact(sb, so, tb):
    announce("generic identity", sb, so, tb)
    p = dfttest_params(ftype=2, sigma=1.0, sbsize=sb, sosize=so, tbsize=tb)
    x = synth.int_plane(4, 120, 160, 9, 8, 6.0).numpy()
    out = filter_clip(VideoClip([x], "int", 8), p).planes[0]
    np.testing.assert_array_equal(out, x)


def test_generic_frame_entry_host_and_device():
    # The GetFrame boundary on the generic path: host and device planes, one output frame.
    announce("generic process_frame")
    p = dfttest_params(sbsize=8, sosize=4, tbsize=5)
    x, bits, _ = make_plane("u8", (6, 46, 58), 43)
    full = filter_clip(VideoClip([x], "int", 8), p).planes[0]
    plan = get_plan(p, "int", 8, Backend.MI355X())
    n, h, w = x.shape
    for fr in (0, 3, 5):
        srcs = [np.ascontiguousarray(x[min(max(fr - 2 + t, 0), n - 1)]) for t in range(5)]
        out = np.zeros((h, w), np.uint8)
        plan.process_frame([s.ctypes.data for s in srcs], w, out.ctypes.data, w, w, h, memory=1)
        np.testing.assert_array_equal(out, full[fr])
        dev = [torch.from_numpy(s).cuda() for s in srcs]
        dout = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        plan.process_frame([d.data_ptr() for d in dev], w, dout.data_ptr(), w, w, h, memory=0)
        np.testing.assert_array_equal(dout.cpu().numpy(), full[fr])


def test_generic_plans_reject_fused_only_entries():
    announce("generic plans: fused-only entries")
    p = dfttest_params(sbsize=8, sosize=4)
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32)
    x = torch.zeros((2, 32, 32), dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(_lib.Dft2Error, match="allocation-free"):
        plan.process_batch_noalloc([plane_job(x, y)], None)
    with pytest.raises(_lib.Dft2Error, match="batches"):
        _lib.Batch(plan, [plane_job(x, y)])
    sigma = torch.ones(3 * 8 * 5, device="cuda")
    with pytest.raises(_lib.Dft2Error, match="device array"):
        plan.process_batch_sigma([plane_job(x, y)], sigma.data_ptr(), 1.0, None)
    plan.close()
