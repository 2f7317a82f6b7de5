"""GPU parity of the fused gfx950 kernel, called through the C ABI.

Bars: integer formats within 1 LSB, float32 within 1e-5 absolute (values in [0, 1]),
against (a) the golden vectors made by the reference's own fused(), (b) the FP64
oracle on the same seeded inputs, and (c) size-independent properties at full HD.
"""

import dataclasses
import os
import sys

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import Backend, VideoClip, filter_clip, get_plan, plane_job
from dfttest2_amd.params import dfttest_params, rdft
from golden_util import fixture_names, load_fixture, max_err
from oracle import dfttest_np as onp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def tol(is_float):
    return 1e-5 if is_float else 1.0


def gpu_filter(x, p, bits, is_float, be=None):
    clip = VideoClip([x], "float" if is_float else "int", bits)
    return filter_clip(clip, p, be).planes[0]


def make_plane(fmt, shape, seed):
    n, h, w = shape
    if fmt == "u8":
        return synth.int_plane(n, h, w, seed, 8, 6.0).numpy(), 8, False
    if fmt == "u10":
        return synth.int_plane(n, h, w, seed, 10, 12.0).numpy(), 10, False
    if fmt == "u12":
        return synth.int_plane(n, h, w, seed, 12, 50.0).numpy(), 12, False
    if fmt == "u16":
        return synth.int_plane(n, h, w, seed, 16, 700.0).numpy(), 16, False
    return synth.float_plane(n, h, w, seed, 0.05).numpy(), 32, True


@pytest.mark.parametrize("name", fixture_names())
def test_golden_vectors(name):
    announce("golden", name)
    p, x, y, bits, is_float = load_fixture(name)
    out = gpu_filter(x, p, bits, is_float)
    assert out.dtype == y.dtype and out.shape == y.shape
    assert max_err(out, y) <= tol(is_float)
    assert max_err(out, onp.process_plane(x, p, bits, is_float)) <= tol(is_float)


CASES = [
    ("u8", (6, 70, 90), dict()),
    ("u8", (5, 64, 64), dict(tbsize=1)),
    ("u10", (7, 66, 80), dict(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])),
    ("u16", (8, 60, 72), dict(tbsize=7, swin=4, twin=1)),
    ("f32", (6, 75, 61), dict()),
    ("f32", (5, 48, 100), dict(ftype=3, pmin=1.0, pmax=300.0)),
    ("f32", (5, 52, 52), dict(ftype=4, pmin=2.0, pmax=400.0)),
    ("u8", (5, 50, 70), dict(f0beta=0.5, sosize=8)),
    ("u8", (5, 50, 70), dict(f0beta=0.8, sigma=16.0)),
    ("u8", (5, 40, 50), dict(ftype=2, sigma=0.7, zmean=False)),
    ("u8", (5, 47, 53), dict(sosize=14)),
    ("u8", (5, 47, 53), dict(sosize=0)),
    ("f32", (4, 40, 40), dict(sosize=4, swin=9)),
    ("u8", (4, 38, 46), dict(sosize=15)),  # block step 1: 16 blocks per axis cover a pixel
    ("f32", (4, 45, 50), dict(sosize=7, tbsize=5)),  # block step 9
    ("u8", (4, 44, 44), dict(ssystem=1, slocation=[0.0, 2.0, 1.0, 30.0])),
]


@pytest.mark.parametrize("fmt,shape,kw", CASES, ids=[f"{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_random_params_vs_oracle(fmt, shape, kw):
    announce("oracle", fmt, shape, kw)
    p = dfttest_params(**kw)
    x, bits, is_float = make_plane(fmt, shape, seed=100 + len(kw))
    out = gpu_filter(x, p, bits, is_float)
    assert max_err(out, onp.process_plane(x, p, bits, is_float)) <= tol(is_float)


@pytest.mark.parametrize("precision", ["fp32", "split_bf16"])
@pytest.mark.parametrize("step", [4, 5, 6])
@pytest.mark.parametrize("fmt,shape", [("u8", (3, 52, 180)), ("f32", (3, 46, 44))])
def test_block_steps_4_to_6(fmt, shape, step, precision):
    # Block steps 4..6 take the one-frame kernel's per-lane gather (one pass column per
    # lane, at most four covering blocks); 5 and 6 are reachable through the C ABI only
    # (the front end rejects more than 50 % overlap with a step that does not divide 16).
    # 180 columns make tiles whose last pass has fewer than eight blocks, 44 a tile
    # narrower than one pass.
    announce("block step", step, fmt, shape, precision)
    p = dataclasses.replace(dfttest_params(), block_step=step)
    x, bits, is_float = make_plane(fmt, shape, seed=300 + step)
    out = gpu_filter(x, p, bits, is_float, Backend.MI355X(precision=precision))
    bar = tol(is_float) * (5 if precision == "split_bf16" and is_float else 1)
    assert max_err(out, onp.process_plane(x, p, bits, is_float)) <= bar


@pytest.mark.parametrize("precision", ["fp32", "split_bf16"])
@pytest.mark.parametrize("pattern", ["checker", "binary_noise", "impulses"])
def test_full_scale_8bit_content(pattern, precision):
    # With 8-bit samples both precisions take the samples into the first contraction as
    # exact BF16 values (FP32: against the exact three-part split of the basis); full-scale
    # content gives the spectra their widest dynamic range.
    announce("full-scale u8", pattern, precision)
    n, h, w = 5, 52, 68
    rng = np.random.default_rng(11)
    if pattern == "checker":
        yy, xx = np.indices((h, w))
        x = np.stack([((yy + xx + t) % 2) * 255 for t in range(n)])
    elif pattern == "binary_noise":
        x = rng.integers(0, 2, (n, h, w)) * 255
    else:
        x = np.zeros((n, h, w), np.int64)
        x.reshape(-1)[rng.choice(x.size, 60, replace=False)] = 255
    x = x.astype(np.uint8)
    p = dfttest_params()
    out = gpu_filter(x, p, 8, False, Backend.MI355X(precision=precision))
    assert max_err(out, onp.process_plane(x, p, 8, False)) <= 1


@pytest.mark.parametrize("precision", ["fp32", "split_bf16"])
@pytest.mark.parametrize("bits,pattern", [(16, "checker"), (16, "binary_noise"), (10, "binary_noise")])
def test_full_scale_16bit_content(bits, pattern, precision):
    # 9- to 16-bit samples enter the first contraction as their two bytes, each exact in
    # BF16 (both precisions); full-scale content exercises both bytes at their extremes.
    announce("full-scale u16", bits, pattern, precision)
    n, h, w = 5, 52, 68
    peak = (1 << bits) - 1
    rng = np.random.default_rng(13)
    if pattern == "checker":
        yy, xx = np.indices((h, w))
        x = np.stack([((yy + xx + t) % 2) * peak for t in range(n)])
    else:
        x = rng.integers(0, 2, (n, h, w)) * peak
    x = x.astype(np.uint16)
    p = dfttest_params(tbsize=5)
    out = gpu_filter(x, p, bits, False, Backend.MI355X(precision=precision))
    bound = 1 if precision == "fp32" or bits <= 12 else 4
    assert max_err(out, onp.process_plane(x, p, bits, False)) <= bound


def test_asymmetric_sigma_array():
    announce("asymmetric sigma")
    p = dfttest_params()
    rng = np.random.default_rng(7)
    p.sigma = np.abs(rng.normal(1.0, 3.0, p.t_len * 16 * 9)) * 2.0
    x, bits, is_float = make_plane("u8", (5, 50, 60), 31)
    assert max_err(gpu_filter(x, p, bits, is_float), onp.process_plane(x, p, bits, is_float)) <= 1


def test_yuv420_planes_and_passthrough():
    announce("yuv420 planes")
    planes = [t.numpy() for t in synth.yuv420_planes(4, 96, 64, seed=5)]
    clip = VideoClip(planes, "int", 8)
    p = dfttest_params(planes=[0, 2])
    out = filter_clip(clip, p)
    np.testing.assert_array_equal(out.planes[1], planes[1])
    for i in (0, 2):
        assert max_err(out.planes[i], onp.process_plane(planes[i], p, 8, False)) <= 1


def test_frame_entry_host_and_device_match_batch():
    announce("process_frame host/device")
    p = dfttest_params(tbsize=5)
    x, bits, _ = make_plane("u8", (6, 58, 66), 41)
    full = gpu_filter(x, p, bits, False)
    plan = get_plan(p, "int", 8, Backend.MI355X())
    n, h, w = x.shape
    for fr in (0, 2, 5):
        idx = [min(max(fr - 2 + t, 0), n - 1) for t in range(5)]
        srcs = [np.ascontiguousarray(x[i]) for i in idx]
        out = np.zeros((h, w), np.uint8)
        plan.process_frame([s.ctypes.data for s in srcs], w, out.ctypes.data, w, w, h, memory=1)
        np.testing.assert_array_equal(out, full[fr])
        dev = [torch.from_numpy(s).cuda() for s in srcs]
        dout = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()  # the plan's own stream does not wait for torch's
        plan.process_frame([d.data_ptr() for d in dev], w, dout.data_ptr(), w, w, h, memory=0)
        np.testing.assert_array_equal(dout.cpu().numpy(), full[fr])


@pytest.mark.parametrize("step", [4, 5, 6])
@pytest.mark.parametrize("tbsize", [1, 3, 7])
def test_frame_entry_block_steps_match_batch(tbsize, step):
    # Per-frame calls (one output frame per launch, host planes as GetFrame passes them)
    # at the block steps whose overlap-add takes the per-lane gather in the 8-bit
    # instantiations: the same bits as the whole-clip batch, within 1 LSB of the oracle.
    announce("process_frame block step", step, "tbsize", tbsize)
    p = dataclasses.replace(dfttest_params(tbsize=tbsize), block_step=step)
    x, bits, _ = make_plane("u8", (5, 50, 140), 60 + step)
    full = gpu_filter(x, p, bits, False)
    assert max_err(full, onp.process_plane(x, p, bits, False)) <= 1
    plan = get_plan(p, "int", 8, Backend.MI355X())
    n, h, w = x.shape
    r = p.radius
    for fr in (0, n // 2, n - 1):
        srcs = [np.ascontiguousarray(x[min(max(fr - r + t, 0), n - 1)]) for t in range(p.t_len)]
        out = np.zeros((h, w), np.uint8)
        plan.process_frame([s.ctypes.data for s in srcs], w, out.ctypes.data, w, w, h, memory=1)
        np.testing.assert_array_equal(out, full[fr])


@pytest.mark.parametrize("fmt", ["u8", "u10", "f32"])
def test_frame_entry_pinned_host_planes(fmt):
    # Pinned host planes take the direct copies, pageable ones the pinned staging buffer:
    # both give the batch result, also with row strides wider than the plane (strides in
    # bytes, samples of 1, 2 and 4 bytes).
    announce("process_frame pinned host planes", fmt)
    p = dfttest_params()
    x, bits, is_float = make_plane(fmt, (5, 46, 58), 43)
    full = gpu_filter(x, p, bits, is_float)
    plan = get_plan(p, "float" if is_float else "int", bits, Backend.MI355X())
    n, h, w = x.shape
    pitch = 64
    tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.int16, np.dtype(np.float32): torch.float32}
    wide = torch.zeros((n, h, pitch), dtype=tdt[x.dtype]).pin_memory()
    wide[:, :, :w] = torch.from_numpy(x.view(np.int16) if x.dtype == np.uint16 else x)
    stride = pitch * x.itemsize
    for fr in (0, 2, 4):
        idx = [min(max(fr - 1 + t, 0), n - 1) for t in range(3)]
        out = torch.zeros((h, pitch), dtype=wide.dtype).pin_memory()
        plan.process_frame([wide[i].data_ptr() for i in idx], stride, out.data_ptr(), stride, w, h, memory=1)
        np.testing.assert_array_equal(out[:, :w].numpy().view(x.dtype), full[fr])
        paged = np.zeros((h, pitch), x.dtype)
        srcs = [np.ascontiguousarray(wide[i].numpy().view(x.dtype)) for i in idx]
        plan.process_frame([s_.ctypes.data for s_ in srcs], stride, paged.ctypes.data, stride, w, h, memory=1)
        np.testing.assert_array_equal(paged[:, :w], full[fr])


def test_frame_entry_concurrent_threads():
    # fmParallel: VapourSynth calls GetFrame for different frames from several threads at once
    # (hiprtc_source/source.cpp:445-472, tickets over num_streams at :559-566).  Six threads
    # share one plan with two stream slots, each filtering its own frames from pageable host
    # planes and from device planes; every output equals the clip result.
    import threading

    announce("process_frame from 6 threads, num_streams 2")
    p = dfttest_params()
    x, bits, _ = make_plane("u8", (12, 70, 90), 47)
    full = gpu_filter(x, p, bits, False)
    plan = get_plan(p, "int", 8, Backend.MI355X(num_streams=2))
    n, h, w = x.shape
    dev_x = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    errors = []

    def worker(k):
        try:
            torch.cuda.set_device(0)
            for rep in range(3):
                for fr in range(k, n, 6):
                    idx = [min(max(fr - 1 + t, 0), n - 1) for t in range(3)]
                    out = np.zeros((h, w), np.uint8)
                    srcs = [np.ascontiguousarray(x[i]) for i in idx]
                    plan.process_frame([s.ctypes.data for s in srcs], w, out.ctypes.data, w, w, h, memory=1)
                    if not np.array_equal(out, full[fr]):
                        errors.append(f"host frame {fr} thread {k} rep {rep}")
                    dout = torch.empty((h, w), dtype=torch.uint8, device="cuda")  # no fill kernel to wait for
                    plan.process_frame([dev_x[i].data_ptr() for i in idx], w, dout.data_ptr(), w, w, h, memory=0)
                    if not np.array_equal(dout.cpu().numpy(), full[fr]):
                        errors.append(f"device frame {fr} thread {k} rep {rep}")
        except Exception as e:  # reported below, in the main thread
            errors.append(f"thread {k}: {e!r}")

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=90)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors[:5]


def test_batch_of_uneven_clips_matches_single_runs():
    announce("batch uneven clips")
    p = dfttest_params()
    shapes = [(5, 40, 72), (9, 66, 48), (3, 30, 30), (7, 100, 36)]
    xs = [synth.float_plane(*s, seed=60 + i).cuda() for i, s in enumerate(shapes)]
    outs = [torch.empty_like(x) for x in xs]
    plan = get_plan(p, "float", 32, Backend.MI355X())
    plan.process_batch([plane_job(x, o) for x, o in zip(xs, outs)], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for x, o in zip(xs, outs):
        ref = onp.process_plane(x.cpu().numpy(), p, 32, True)
        assert max_err(o.cpu().numpy(), ref) <= 1e-5


def test_strided_rows_and_frame_range():
    announce("strided rows, frame range")
    p = dfttest_params()
    x, bits, _ = make_plane("u8", (6, 45, 51), 71)
    big = torch.zeros((6, 45, 64), dtype=torch.uint8, device="cuda")
    big[:, :, :51] = torch.from_numpy(x).cuda()
    src = big[:, :, :51]
    dst_big = torch.full((6, 45, 64), 7, dtype=torch.uint8, device="cuda")
    dst = dst_big[:, :, :51]
    plan = get_plan(p, "int", 8, Backend.MI355X())
    plan.process_batch([plane_job(src, dst, 2, 5)], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ref = onp.process_plane(x, p, 8, False)
    got = dst_big.cpu().numpy()
    assert max_err(got[2:5, :, :51], ref[2:5]) <= 1
    assert np.all(got[:2] == 7) and np.all(got[5:] == 7) and np.all(got[:, :, 51:] == 7)


def test_short_clip_temporal_clamp():
    announce("1-frame clip, tbsize 5")
    p = dfttest_params(tbsize=5)
    x, bits, _ = make_plane("u8", (1, 40, 40), 81)
    assert max_err(gpu_filter(x, p, bits, False), onp.process_plane(x, p, bits, False)) <= 1


def test_scratch_chunking_gives_same_result(monkeypatch):
    announce("scratch chunking")
    p = dfttest_params()
    x = synth.int_plane(12, 90, 120, 91).cuda()
    o1 = torch.empty_like(x)
    o2 = torch.empty_like(x)
    plan1 = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    monkeypatch.setenv("DFT2_SCRATCH_MB", "1")
    plan2 = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    plan1.process_batch([plane_job(x, o1)])
    plan2.process_batch([plane_job(x, o2)])
    assert max_err(o1.cpu().numpy(), o2.cpu().numpy()) <= 1
    plan1.close()
    plan2.close()


def test_too_small_plane_is_rejected():
    announce("too small plane")
    p = dfttest_params()
    x = torch.zeros((2, 8, 8), dtype=torch.uint8, device="cuda")
    plan = get_plan(p, "int", 8, Backend.MI355X())
    with pytest.raises(_lib.Dft2Error, match="too small"):
        plan.process_batch([plane_job(x, torch.empty_like(x))])


@pytest.mark.parametrize("fmt", ["u8", "u10"])
def test_full_hd_identity_is_exact(fmt):
    # ftype 2 with sigma 1 is the identity (rectangular twin): integer output must equal
    # the input sample for sample at the full 1080p luma size.
    announce("1080p identity", fmt)
    p = dfttest_params(ftype=2, sigma=1.0)
    bits = 8 if fmt == "u8" else 10
    x = synth.int_plane(3, 1080, 1920, 3, bits, 6.0, device="cuda")
    out = torch.empty_like(x)
    plan = get_plan(p, "int", bits, Backend.MI355X())
    plan.process_batch([plane_job(x, out)], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.equal(out.to(torch.int32), x.to(torch.int32))


@pytest.mark.parametrize("bits", [8, 16])
def test_8k_identity_is_exact(bits, monkeypatch):
    # The identity at 7680x4320 with the scratch limited to 256 MiB: many tiles per frame,
    # launches split into chunks, and (16-bit) the byte-split x transform at full scale.
    announce("8K identity", bits)
    monkeypatch.setenv("DFT2_SCRATCH_MB", "256")
    p = dfttest_params(ftype=2, sigma=1.0)
    x = synth.int_plane(3, 4320, 7680, 5, bits, 300.0 if bits == 16 else 6.0, device="cuda")
    out = torch.empty_like(x)
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, bits)
    plan.process_batch([plane_job(x, out)], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(out.to(torch.int32), x.to(torch.int32))


def test_full_hd_constant_frames_pass_through():
    announce("1080p constant")
    p = dfttest_params(ftype=1, tbsize=5)
    x = torch.full((5, 1080, 1920), 101, dtype=torch.uint8, device="cuda")
    out = torch.empty_like(x)
    plan = get_plan(p, "int", 8, Backend.MI355X())
    plan.process_batch([plane_job(x, out)], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.equal(out, x)


SPLIT_CASES = [
    ("u8", (5, 64, 80), dict()),
    ("u10", (6, 60, 72), dict(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])),
    ("u8", (8, 50, 66), dict(tbsize=7, swin=1)),
    ("u12", (4, 48, 48), dict(tbsize=1, f0beta=0.5)),
    ("u16", (4, 48, 48), dict(tbsize=1, ftype=2, sigma=0.8)),
    ("f32", (5, 55, 64), dict()),
]


@pytest.mark.parametrize("ftype", [3, 4])
def test_split_bf16_rejects_gains_above_one(ftype):
    announce("split-bf16 plan rejects filter type", ftype)
    p = dfttest_params(ftype=ftype, sigma=4.0, pmin=1.0, pmax=100.0)
    with pytest.raises(_lib.Dft2Error, match="split BF16"):
        get_plan(p, "int", 10, Backend.MI355X(precision="split_bf16"))


@pytest.mark.parametrize("fmt,shape,kw", SPLIT_CASES, ids=[f"{c[0]}-{i}" for i, c in enumerate(SPLIT_CASES)])
def test_split_bf16_vs_oracle(fmt, shape, kw):
    # Split BF16 (hi + lo operands, three products, FP32 accumulation) keeps about 16
    # significant bits per operand: integer formats up to 12 bits stay within 1 LSB, 16-bit
    # samples within 4 LSB; float32 output is held to 5e-5 absolute (FP32 meets 1e-5).
    announce("split-bf16", fmt, shape, kw)
    p = dfttest_params(**kw)
    x, bits, is_float = make_plane(fmt, shape, seed=200 + len(kw))
    out = gpu_filter(x, p, bits, is_float, Backend.MI355X(precision="split_bf16"))
    bound = 5e-5 if is_float else (4.0 if bits > 12 else 1.0)
    assert max_err(out, onp.process_plane(x, p, bits, is_float)) <= bound


def test_split_bf16_golden_defaults():
    announce("split-bf16 golden u8_defaults")
    p, x, y, bits, is_float = load_fixture("u8_defaults")
    out = gpu_filter(x, p, bits, is_float, Backend.MI355X(precision="split_bf16"))
    assert max_err(out, y) <= 1


def test_repeated_and_alternating_job_tables():
    # Identical job tables are reused from device memory; the results must still follow
    # the buffers' current contents, and a table evicted from the cache must be rebuilt.
    announce("job table reuse")
    p = dfttest_params()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32)
    fresh = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32)
    shapes = [(3, 40, 48), (4, 36, 52), (3, 44, 40), (2, 40, 40), (3, 50, 36), (4, 38, 46)]
    xs = [synth.float_plane(*s, seed=900 + i).cuda() for i, s in enumerate(shapes)]
    ys = [torch.empty_like(x) for x in xs]
    refs = [torch.empty_like(x) for x in xs]
    for rep in range(3):
        for i in list(range(len(xs))) + [0, 1, 0]:
            xs[i].mul_(0.9).add_(0.03)
            plan.process_batch([plane_job(xs[i], ys[i])])
            fresh.process_batch([plane_job(xs[i], refs[i])])
            # the overlap-add sums in a fixed order: bit-identical results
            assert torch.equal(ys[i], refs[i]), (rep, i)
    plan.close()
    fresh.close()


@pytest.mark.parametrize("tbsize", [5, 7])
def test_forward_is_bit_reproducible(tbsize):
    # Fixed-order overlap-add at every radius: two runs give the same bits.
    announce("bit-reproducible forward, tbsize", tbsize)
    p = dfttest_params(tbsize=tbsize)
    x = synth.float_plane(6, 70, 90, seed=31).cuda()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32)
    outs = []
    for _ in range(2):
        y = torch.empty_like(x)
        plan.process_batch([plane_job(x, y)])
        torch.cuda.synchronize()
        outs.append(y)
    plan.close()
    assert torch.equal(outs[0], outs[1])
    assert max_err(outs[0].cpu().numpy(), onp.process_plane(x.cpu().numpy(), p, 32, True)) <= 1e-5


@pytest.mark.parametrize("fmt", ["u8", "f32"])
def test_separable_window_with_zero_centre_slice(fmt):
    # a[t] b[y] c[x] with a[r] = 0: the separable form scales the temporal basis by 1 / a[r],
    # so such a window takes the per-lane form; its output is the oracle's (all zero here: the
    # synthesis window is the centre slice), with no division by zero.
    announce("separable window, zero centre slice", fmt)
    p = dfttest_params(twin=0)
    w = np.asarray(p.window, dtype=np.float64).reshape(p.t_len, 16, 16).copy()
    w[p.radius] = 0.0
    p.window = w.reshape(-1)
    p.window_freq = np.asarray(rdft([v * 255 for v in p.window], (p.t_len, 16, 16)))
    x, bits, is_float = make_plane(fmt, (5, 40, 44), 19)
    out = gpu_filter(x, p, bits, is_float)
    ref = onp.process_plane(x, p, bits, is_float)
    assert np.isfinite(out.astype(np.float64)).all()
    assert max_err(out, ref) <= tol(is_float)


@pytest.mark.parametrize("fmt,precision,asym", [("u8", "fp32", False), ("f32", "fp32", True),
                                                ("u10", "split_bf16", False)])
def test_non_separable_window(fmt, precision, asym):
    # A window that is not a[t] b[y] c[x] takes the kernel form that holds the window per
    # lane (separable windows, the ones DFTTest builds, fold it into the bases instead).
    announce("non-separable window", fmt, precision, asym)
    p = dfttest_params()
    rng = np.random.default_rng(11)
    w = np.asarray(p.window, dtype=np.float64).reshape(p.t_len, 16, 16)
    p.window = (w * rng.uniform(0.7, 1.3, w.shape)).reshape(-1)
    p.window_freq = np.asarray(rdft([v * 255 for v in p.window], (p.t_len, 16, 16)))
    if asym:
        p.sigma = np.abs(rng.normal(1.0, 3.0, p.t_len * 16 * 9)) * 2.0
    x, bits, is_float = make_plane(fmt, (5, 52, 60), 17)
    out = gpu_filter(x, p, bits, is_float, Backend.MI355X(precision=precision))
    bound = tol(is_float) if precision == "fp32" else (5e-5 if is_float else 1.0)
    assert max_err(out, onp.process_plane(x, p, bits, is_float)) <= bound
