"""The forward's two block-loop forms on the GPU: the cross-frame form (K output frames per
work item, each input slice transformed once per item) and the one-frame form, forced per
plan with DFT2_REUSE = 1, 4, 8 at tbsize 5 and 7 for 8-bit, 10-bit and float samples, both
against the FP64 oracle and against each other; frame chunks that the scratch size cuts
(rounded to whole groups of K frames); and a per-frame call of a cross-frame plan (which
takes the one-frame form).  Bars: integer formats within 1 LSB, float32 within 1e-5."""

import sys

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import plane_job
from dfttest2_amd.params import dfttest_params
from golden_util import max_err
from oracle import dfttest_np as onp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def plane(fmt, shape, seed):
    n, h, w = shape
    if fmt == "u8":
        return synth.int_plane(n, h, w, seed, 8, 6.0).numpy(), 8, False
    if fmt == "u10":
        return synth.int_plane(n, h, w, seed, 10, 12.0).numpy(), 10, False
    return synth.float_plane(n, h, w, seed, 0.05).numpy(), 32, True


def run(x, p, bits, is_float, precision=_lib.DFT2_PRECISION_FP32):
    st = _lib.DFT2_SAMPLE_FLOAT if is_float else _lib.DFT2_SAMPLE_INT
    plan = _lib.Plan(p, st, bits, 0, 1, precision)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    plan.process_batch([plane_job(xd, yd)])
    torch.cuda.synchronize()
    plan.close()
    return yd.cpu().numpy()


@pytest.mark.parametrize("fmt,tbsize,precision", [("u8", 5, "fp32"), ("u8", 7, "fp32"), ("u10", 5, "fp32"),
                                                  ("u10", 7, "split_bf16"), ("f32", 5, "split_bf16"),
                                                  ("f32", 7, "fp32")])
def test_forms_match_oracle_and_each_other(fmt, tbsize, precision, monkeypatch):
    # 19 frames: groups of 8, 8 and 3 at K = 8, of 4 (and a group of 3) at K = 4
    announce("cross-frame forms", fmt, tbsize, precision)
    p = dfttest_params(tbsize=tbsize)
    x, bits, is_float = plane(fmt, (19, 60, 76), 41)
    prec = _lib.DFT2_PRECISION_SPLIT_BF16 if precision == "split_bf16" else _lib.DFT2_PRECISION_FP32
    ref = onp.process_plane(x, p, bits, is_float)
    bar = 1e-5 if is_float else 1.0
    if precision == "split_bf16" and is_float:
        bar = 5e-5
    outs = {}
    for k in ("1", "4", "8"):
        monkeypatch.setenv("DFT2_REUSE", k)
        outs[k] = run(x, p, bits, is_float, prec)
        err = max_err(outs[k], ref)
        print(f"K = {k}: max error {err:.3g} against FP64", flush=True)
        assert err <= bar
    for k in ("4", "8"):
        assert max_err(outs[k], outs["1"]) <= bar


def test_chunks_hold_whole_groups(monkeypatch):
    # A 1 MiB scratch holds fewer than the clip's 19 frames per launch in the cross-frame
    # form (about 14 with its tile regions); the launches hold whole groups of K = 8.
    announce("cross-frame form, chunked launches")
    p = dfttest_params(tbsize=5)
    x, bits, is_float = plane("u8", (19, 60, 72), 43)
    monkeypatch.setenv("DFT2_REUSE", "8")
    whole = run(x, p, bits, is_float)
    monkeypatch.setenv("DFT2_SCRATCH_MB", "1")
    chunked = run(x, p, bits, is_float)
    assert np.array_equal(chunked, whole)
    assert max_err(chunked, onp.process_plane(x, p, bits, is_float)) <= 1.0


def test_per_frame_call_of_cross_frame_plan(monkeypatch):
    # dft2_process_frame on device planes (one output frame per call) on a plan whose
    # batches take the cross-frame form: the call takes the one-frame form.
    announce("per-frame call, cross-frame plan")
    monkeypatch.setenv("DFT2_REUSE", "8")
    p = dfttest_params(tbsize=5)
    x, bits, is_float = plane("u8", (7, 44, 52), 47)
    ref = onp.process_plane(x, p, bits, is_float)
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    h, w = x.shape[1:]
    dev = [torch.from_numpy(x[i]).cuda() for i in range(7)]
    for n in range(7):
        idx = [min(max(n - 2 + t, 0), 6) for t in range(5)]
        out = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        plan.process_frame([dev[i].data_ptr() for i in idx], w, out.data_ptr(), w, w, h, memory=0)
        assert max_err(out.cpu().numpy(), ref[n]) <= 1.0, n
    plan.close()
