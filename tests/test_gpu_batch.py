"""GPU parity of batches (dft2_batch_*): clips of uneven size and length set up once and
run as one captured graph launch, with the tile shape and launch form picked by the
autotuner or forced.  Every result is held to the oracle's bar (1 LSB for integer
formats, 1e-5 absolute for float32), and graph replays are checked against eager runs
of the same inputs, with eager work between the replays.
"""

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

SHAPES = [(5, 40, 72), (9, 66, 48), (3, 30, 30), (7, 100, 36), (4, 52, 90), (6, 36, 120)]


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def u8_clips():
    p = dfttest_params()
    xs = [synth.int_plane(*s, seed=300 + i, bits=8, sd_lsb=6.0).numpy() for i, s in enumerate(SHAPES)]
    refs = [onp.process_plane(x, p, 8, False) for x in xs]
    return p, xs, refs


@pytest.fixture(scope="module")
def f32_clips():
    p = dfttest_params()
    xs = [synth.float_plane(*s, seed=400 + i).numpy() for i, s in enumerate(SHAPES)]
    refs = [onp.process_plane(x, p, 32, True) for x in xs]
    return p, xs, refs


def _run(plan, xs, **kw):
    dev = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.empty_like(d) for d in dev]
    batch = _lib.Batch(plan, [plane_job(d, o) for d, o in zip(dev, outs)], **kw)
    batch.launch()
    return batch, dev, outs


def test_autotuned_batch_of_uneven_float_clips(f32_clips):
    announce("batch autotune f32", SHAPES)
    p, xs, refs = f32_clips
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32)
    batch, dev, outs = _run(plan, xs, autotune=True, sample_frames=1)
    for o, ref in zip(outs, refs):
        assert max_err(o.cpu().numpy(), ref) <= 1e-5
    info = batch.info()
    cands = batch.candidates()
    assert info["n_candidates"] == len(cands) > 2
    best = min(cands, key=lambda c: c["ms"])
    assert (info["tile_w"], info["tile_h"], info["persistent"]) == (best["tile_w"], best["tile_h"], best["persistent"])
    assert info["tuned_ms"] == pytest.approx(best["ms"])
    assert info["work_items"] > 0 and info["n_chunks"] >= 1
    batch.close()
    plan.close()


@pytest.mark.parametrize("tile", [(4, 4), (16, 8), (8, 16), (12, 12), (16, 16), (6, 12)])
@pytest.mark.parametrize("persistent", [False, True])
def test_forced_tile_shapes_u8(u8_clips, tile, persistent):
    announce("batch u8 tile", tile, "persistent", persistent)
    p, xs, refs = u8_clips
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    batch, dev, outs = _run(plan, xs, autotune=False, tile_w=tile[0], tile_h=tile[1], persistent=persistent)
    info = batch.info()
    assert (info["tile_w"], info["tile_h"], info["persistent"]) == (tile[0], tile[1], int(persistent))
    assert info["n_candidates"] == 0
    for o, ref in zip(outs, refs):
        assert max_err(o.cpu().numpy(), ref) <= 1
    batch.close()
    plan.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_graph_replays_follow_the_inputs(persistent):
    # Replays read the buffers' current contents: change the inputs with eager torch work
    # between replays and compare every replay with an eager process_batch of the same data.
    announce("graph replays, persistent", persistent)
    p = dfttest_params()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32)
    dev = [synth.float_plane(*s, seed=500 + i).cuda() for i, s in enumerate(SHAPES[:4])]
    outs = [torch.empty_like(d) for d in dev]
    refs = [torch.empty_like(d) for d in dev]
    batch = _lib.Batch(plan, [plane_job(d, o) for d, o in zip(dev, outs)], autotune=False, tile_w=8, tile_h=8,
                       persistent=persistent)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(4):
        if rep:
            for d in dev:
                d.mul_(0.7).add_(0.05 * rep)
        for o in outs:
            o.fill_(-1.0)
        batch.launch(stream if rep % 2 else None)
        torch.cuda.synchronize()
        plan.process_batch([plane_job(d, r) for d, r in zip(dev, refs)], stream)
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            assert torch.max(torch.abs(o - r)).item() <= 1e-6
    batch.close()
    plan.close()


def test_chunked_persistent_batch_matches_oracle(u8_clips, monkeypatch):
    announce("batch chunked persistent")
    p, xs, refs = u8_clips
    monkeypatch.setenv("DFT2_SCRATCH_MB", "1")
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    batch, dev, outs = _run(plan, xs, autotune=False, tile_w=16, tile_h=16, persistent=True)
    assert batch.info()["n_chunks"] > 1
    for o, ref in zip(outs, refs):
        assert max_err(o.cpu().numpy(), ref) <= 1
    batch.close()
    plan.close()


def test_batch_argument_checks():
    announce("batch argument checks")
    p = dfttest_params()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    x = torch.zeros((3, 40, 40), dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(_lib.Dft2Error, match="at most 2 tiles"):
        _lib.Batch(plan, [plane_job(x, y)], autotune=False, tile_w=2, tile_h=2)
    with pytest.raises(_lib.Dft2Error, match="LDS"):
        _lib.Batch(plan, [plane_job(x, y)], autotune=False, tile_w=48, tile_h=48)
    small = torch.zeros((2, 8, 8), dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.Dft2Error, match="too small"):
        _lib.Batch(plan, [plane_job(small, torch.empty_like(small))])
    empty = _lib.Batch(plan, [], autotune=False)
    empty.launch()
    assert empty.info()["work_items"] == 0
    empty.close()
    plan.close()


def test_persistent_form_needs_separable_window():
    # The persistent launch form is built for separable windows only: forcing it on another
    # window is an argument error, and the autotuner does not offer it.
    announce("persistent form, non-separable window")
    p = dfttest_params()
    w = np.asarray(p.window, dtype=np.float64).reshape(p.t_len, 16, 16)
    w[1, 3, 5] *= 1.5
    p.window = w.reshape(-1)
    from dfttest2_amd.params import rdft
    p.window_freq = np.asarray(rdft([v * 255 for v in p.window], (p.t_len, 16, 16)))
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    x = torch.from_numpy(synth.int_plane(3, 40, 48, seed=7, bits=8, sd_lsb=6.0).numpy()).cuda()
    y = torch.empty_like(x)
    with pytest.raises(_lib.Dft2Error, match="separable"):
        _lib.Batch(plan, [plane_job(x, y)], autotune=False, persistent=True)
    batch = _lib.Batch(plan, [plane_job(x, y)], autotune=True, sample_frames=1)
    assert all(c["persistent"] == 0 for c in batch.candidates())
    batch.launch()
    assert max_err(y.cpu().numpy(), onp.process_plane(x.cpu().numpy(), p, 8, False)) <= 1
    batch.close()
    plan.close()
