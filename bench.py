"""Throughput of the MI355X DFTTest path on its headline workload, with every part of the
metric measured in the same run.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|fp32_1080p|...]

Headline (`value`): YUV420 8-bit, 1920x1080 luma + 2 x 960x540 chroma, 240 frames, DFTTest
defaults (ftype 0, sigma 8, sbsize 16, sosize 12, tbsize 3, swin 0, twin 7, zmean), all
three planes, synthetic seeded clip resident in HBM before timing.  One step = the whole
240-frame clip through the fused kernel (one dft2_process_batch call).

The same JSON line also carries (default --workload all):
  split_bf16_1080p  the headline workload in the split-BF16 form;
  precision4k       YUV420P10 3840x2160 + 2 x 1920x1080, 120 frames, ftype 1, tbsize 5,
                    sigma curve 0.0:4 0.5:8 1.0:16, FP32 and split BF16;
  diff              8 x 256x256x16 float32 clips, Wiener with sigma [3, 16, 9] as a
                    tensor, MSE loss: forward vs forward + backward;
  batch             64 float32 clips of uneven size, one autotuned graph launch;
  roofline          the headline fused kernel against its MFMA peak;
  cpu_baseline      the reference's own block core (fused() of the GCC backend, built
                    from the reference sources into oracle/_ref) timed on this host's CPU
                    share on a bounded sample, and 1 core;
  parity            the GPU outputs against the reference's outputs on that sample (full
                    HD, all planes), and both 4K forms against the FP64 oracle on a check
                    clip, hard decisions replayed; exit status 1 if a bar is missed.

Other workloads run alone, never in the default line: frame_host (per-frame calls on host
planes, as a VapourSynth GetFrame makes them, PCIe included; plain and keyed entries,
pageable and pinned planes, one and four caller threads).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "vs-dfttest2_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mpx/s at 1080p & 4K, plain FP32 vs split-BF16 MFMA at ≤1 LSB; fwd+bwd Mpx/s"
WIDTH, HEIGHT, FRAMES = 1920, 1080, 240
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0  # dense
CLOCK_GHZ = 2.4                 # MI355X peak engine clock
SIMDS = 256 * 4                 # 256 CUs x 4 SIMDs
# HBM bytes per fused launch on the headline workload (one launch per step): rocprofv3
# --pmc FETCH_SIZE x 2 + WRITE_SIZE x 1 (KiB per dispatch, 240-frame launch), the factors
# being what the same counters read for a plain copy kernel of known size
# (profiles/fused_kernel_1080p_yuv420_u8_fp32.md).  A fixed figure of this build's kernel,
# not measured inside the timed run.
FUSED_TRAFFIC_BYTES = (1_019_360 * 2 + 4_797_214) * 1024
# MFMA pipe cycles per instruction on gfx950 (wave64): the FP32 16x16x4 and the BF16
# 16x16x16 measured, the BF16 16x16x32 at the published dense BF16 rate.
MFMA_CYCLES = {"f32_16x16x4": 32, "bf16_16x16x16": 16, "bf16_16x16x32": 16}
MFMA_FLOP = {"f32_16x16x4": 2 * 16 * 16 * 4, "bf16_16x16x16": 2 * 16 * 16 * 16, "bf16_16x16x32": 2 * 16 * 16 * 32}


def flop_per_block(t_len: int) -> int:
    # Dense separable DFT contractions per block: forward x rDFT + y cDFT per slice,
    # temporal forward and centre-slice inverse, spatial inverse of the centre slice.
    return 2 * (t_len * 13824 + 576 * t_len * t_len + 576 * t_len + 13824)


def mfma_peak(t_len: int, precision: str, exact_input: bool):
    """Published MFMA peak for one block's algorithmic flops, each stage priced at the
    rate of the MFMA type it runs on: with 8-bit samples the FP32 form computes the first
    forward contraction (the x rDFT of every slice, 2 * 4,608 flop per slice) on the BF16
    MFMA (exact products), the rest on the FP32 MFMA; the split-BF16 form runs every
    contraction on the BF16 MFMA."""
    flops = flop_per_block(t_len)
    if precision == "split_bf16":
        return PEAK_BF16_MFMA_TFLOPS, "bf16 MFMA"
    bf16 = 2 * 4608 * t_len if exact_input else 0
    if not bf16:
        return PEAK_FP32_MFMA_TFLOPS, "fp32 MFMA"
    seconds = bf16 / (PEAK_BF16_MFMA_TFLOPS * 1e12) + (flops - bf16) / (PEAK_FP32_MFMA_TFLOPS * 1e12)
    return round(flops / seconds / 1e12, 1), (f"fp32 MFMA for {flops - bf16} flop/block, bf16 MFMA "
                                              f"(exact products) for the {bf16} flop/block of the x rDFT")


def mfma_per_block(slices: float, precision: str, sample_bytes: int) -> dict:
    """MFMA instructions one wave issues per output block in the separable-window kernel,
    given the slice spectra it computes per block (T in the one-frame form; in the
    cross-frame form a group of K frames computes K + T - 1, which the library's timing
    record counts): the x transform of each slice (8-bit samples: 16x16x16 BF16 over the
    basis parts, three in FP32, two in split BF16; 16-bit samples: 16x16x32 BF16 over their
    two bytes, three or two), the y transform of each slice and the two inverse contractions
    (four FP32 16x16x4 or three split-BF16 16x16x16 each)."""
    split = precision == "split_bf16"
    parts = 2 if split else 3
    x_kind = "bf16_16x16x16" if sample_bytes == 1 else "bf16_16x16x32"
    counts = {x_kind: round(parts * slices, 3)}
    other = 2 + slices  # y per slice + two inverse contractions
    if split:
        counts["bf16_16x16x16"] = round(counts.get("bf16_16x16x16", 0) + 3 * other, 3)
    else:
        counts["f32_16x16x4"] = round(4 * other, 3)
    return counts


def slices_per_block(kt: dict, t_len: int) -> float:
    """Slice spectra computed per output block by the timed launches (library count)."""
    return kt["slice_spectra"] / kt["blocks"] if kt.get("blocks") else float(t_len)


def executed(counts: dict, blocks: float, kernel_ms: float) -> dict:
    """Executed MFMA work of a launch: flops per block, the executed rate, and the share of
    the SIMDs' cycles the MFMA pipe is busy with it (issue cycles per instruction above)."""
    flop = round(sum(MFMA_FLOP[k] * n for k, n in counts.items()))
    cycles = sum(MFMA_CYCLES[k] * n for k, n in counts.items())
    s = kernel_ms / 1e3
    return {"mfma_per_block": counts, "flop_per_block": flop, "tflops": round(blocks * flop / s / 1e12, 2),
            "mfma_pipe_busy": round(blocks * cycles / (s * SIMDS * CLOCK_GHZ * 1e9), 4)}


def blocks_per_plane(w: int, h: int, step: int = 4) -> int:
    def pad_num(size):
        ps = size + ((16 - size % 16) if size % 16 else 0) + max(16 - step, step) * 2
        return (ps - 16) // step + 1

    return pad_num(w) * pad_num(h)


def cpu_threads() -> int:
    """This job's CPU share: the CPUs it may run on, capped by OMP_NUM_THREADS where the
    host sets it (a GPU box gives one GPU's job a fixed share of its cores)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _time_steps(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def run_clip(plan, jobs, steps, warmup, stream):
    """Times `steps` process_batch calls after `warmup`, with the library's per-launch HIP
    events on the launch stream; returns (ms per step, timing totals)."""
    import torch

    handle = stream.cuda_stream
    for _ in range(warmup):
        plan.process_batch(jobs, handle)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    plan.set_timing(True)
    ev0.record(stream)
    for _ in range(steps):
        plan.process_batch(jobs, handle)
    ev1.record(stream)
    torch.cuda.synchronize()
    kt = plan.timing()
    plan.set_timing(False)
    return ev0.elapsed_time(ev1) / steps, kt


def bench_1080p(args, precision, planes, keep_frames):
    """The headline workload in one precision; returns (summary, first keep_frames output
    frames of every plane on the host)."""
    import torch

    from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
    from dfttest2_amd.params import dfttest_params

    params = dfttest_params()
    outs = [torch.empty_like(p) for p in planes]
    plan = get_plan(params, "int", 8, Backend.MI355X(precision=precision))
    jobs = [plane_job(p, o) for p, o in zip(planes, outs)]
    stream = torch.cuda.current_stream()
    ms, kt = run_clip(plan, jobs, args.steps, args.warmup, stream)
    fused_ms = kt["fused_ms"] / max(1, kt["fused_launches"])
    fin_ms = kt["finalize_ms"] / max(1, kt["finalize_launches"])
    log(f"1080p {precision}: {ms:.3f} ms/step; fused {fused_ms:.3f} ms/launch over {kt['fused_launches']} "
        f"launches, finalize {fin_ms:.3f} ms")
    blocks = (blocks_per_plane(WIDTH, HEIGHT) + 2 * blocks_per_plane(WIDTH // 2, HEIGHT // 2)) * FRAMES
    launches = kt["fused_launches"] / args.steps
    logical = blocks * flop_per_block(params.t_len) / launches / (fused_ms / 1e3) / 1e12
    summary = {
        "ms_per_step": round(ms, 4), "fps": round(FRAMES / (ms / 1e3), 1),
        "mpx_s": round(WIDTH * HEIGHT * FRAMES / (ms / 1e3) / 1e6, 2),
        "kernel": "fused_kernel", "kernel_ms": round(fused_ms, 4), "finalize_ms": round(fin_ms, 4),
        "launches_per_step": launches, "tflops_logical": round(logical, 3),
        "executed": executed(mfma_per_block(slices_per_block(kt, params.t_len), precision, 1), blocks / launches,
                             fused_ms),
    }
    summary["executed"]["slices_per_block"] = round(slices_per_block(kt, params.t_len), 4)
    host = [o[:keep_frames].cpu().numpy() for o in outs]
    del outs
    return summary, host


def bench_precision4k(args):
    """4K YUV420P10, ftype 1, tbsize 5, sigma curve: FP32 vs split BF16 throughput, and both
    forms on a 640x360x8 check clip of the same generator (outputs returned for the CPU
    leg's FP64 comparison, with the kernel's decision records)."""
    import torch

    from dfttest2_amd import synth
    from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
    from dfttest2_amd.params import dfttest_params

    w, h, frames = 3840, 2160, 120
    params = dfttest_params(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])
    log(f"generating {frames} frames of {w}x{h} YUV420P10")
    planes = synth.yuv420_planes(frames, w, h, seed=2, bits=10, sd_lsb=12.0, device="cuda")
    outs = [torch.empty_like(p) for p in planes]
    jobs = [plane_job(p, o) for p, o in zip(planes, outs)]
    stream = torch.cuda.current_stream()
    blocks = (blocks_per_plane(w, h) + 2 * blocks_per_plane(w // 2, h // 2)) * frames
    res = {"workload": "YUV420P10 3840x2160 + 2x1920x1080, 120 frames, ftype 1, tbsize 5, "
                       "slocation 0.0:4 0.5:8 1.0:16"}
    for prec in ("fp32", "split_bf16"):
        plan = get_plan(params, "int", 10, Backend.MI355X(precision=prec))
        ms, kt = run_clip(plan, jobs, args.steps, args.warmup, stream)
        fused_ms = kt["fused_ms"] / max(1, kt["fused_launches"])
        launches = kt["fused_launches"] / args.steps
        logical = blocks * flop_per_block(params.t_len) / launches / (fused_ms / 1e3) / 1e12
        peak = mfma_peak(params.t_len, prec, exact_input=True)[0]  # 10-bit samples: exact bytes on BF16
        res[prec] = {"ms_per_step": round(ms, 3), "fps": round(frames / (ms / 1e3), 1),
                     "mpx_s": round(w * h * frames / (ms / 1e3) / 1e6, 1), "kernel_ms": round(fused_ms, 3),
                     "launches_per_step": launches, "tflops_logical": round(logical, 2),
                     "frac_logical": round(logical / peak, 4),
                     "executed": executed(mfma_per_block(slices_per_block(kt, params.t_len), prec, 2),
                                          blocks / launches, fused_ms)}
        res[prec]["executed"]["slices_per_block"] = round(slices_per_block(kt, params.t_len), 4)
        log(f"4K {prec}: {ms:.2f} ms per {frames} frames, fused {fused_ms:.2f} ms/launch")
    res["speedup_split_bf16"] = round(res["fp32"]["ms_per_step"] / res["split_bf16"]["ms_per_step"], 3)
    del outs, jobs, planes
    torch.cuda.empty_cache()
    # check clip: outputs and decision records of both forms
    from dfttest2_amd.decisions import run_with_decisions

    check = [pl.numpy() for pl in synth.yuv420_planes(8, 640, 360, seed=3, bits=10, sd_lsb=12.0)]
    gpu = {}
    for prec in ("fp32", "split_bf16"):
        plan = get_plan(params, "int", 10, Backend.MI355X(precision=prec))
        runs = []
        for pl in check:
            out_rec, dec = run_with_decisions(plan, pl, params)
            # the product path (the cross-frame form at this tbsize) on the same clip
            x = torch.from_numpy(pl).cuda()
            y = torch.empty_like(x)
            plan.process_batch([plane_job(x, y)], None)
            runs.append((out_rec, dec, y.cpu().numpy()))
        gpu[prec] = runs
    return res, (params, check, gpu)


def bench_diff(args):
    """8 float32 clips of 256x256x16 frames: noisy = clean + N(0, 0.05), sigma [3, 16, 9] = 8
    with requires_grad, loss = MSE(DFTTest(noisy), clean): forward vs forward+backward."""
    import torch

    from dfttest2_amd import synth
    from dfttest2_amd.autograd import dfttest_diff, diff_params

    dev = torch.device("cuda", 0)
    p = diff_params(tbsize=3)
    clean = torch.stack([synth.clean_plane(16, 256, 256, 50 + i, device=dev) for i in range(8)])
    noisy = clean + torch.stack([synth.noise((16, 256, 256), 0.05, 60 + i, device=dev) for i in range(8)])
    x = noisy.clone().requires_grad_(True)
    sigma = torch.full((3, 16, 9), 8.0, device=dev, requires_grad=True)
    stream = torch.cuda.current_stream(dev)

    def fwd():
        with torch.no_grad():
            dfttest_diff(x, sigma, p)

    def fwd_bwd():
        # a training step: gradients released first (optimizer.zero_grad(set_to_none=True),
        # torch's default), then forward, MSE loss and backward
        x.grad = None
        sigma.grad = None
        loss = torch.nn.functional.mse_loss(dfttest_diff(x, sigma, p), clean)
        loss.backward()

    ms_f = _time_steps(fwd, args.steps, args.warmup, stream)
    ms_fb = _time_steps(fwd_bwd, args.steps, args.warmup, stream)
    mpx = 8 * 16 * 256 * 256 / 1e6
    log(f"diff: fwd {ms_f:.3f} ms, fwd+bwd {ms_fb:.3f} ms")
    return {"workload": "8 x 256x256x16 float32 clips, Wiener, sigma [3,16,9] tensor, MSE loss",
            "fwd_mpx_s": round(mpx / (ms_f / 1e3), 1), "fwd_bwd_mpx_s": round(mpx / (ms_fb / 1e3), 1),
            "ratio_fwd_bwd_to_fwd": round(ms_f / ms_fb, 3), "ms_fwd": round(ms_f, 3), "ms_fwd_bwd": round(ms_fb, 3)}


def batch_clips(n_clips: int = 64, seed: int = 3):
    """Uneven clips: width ~ U{320..1920}, height ~ U{240..1080} (even), 8..32 frames,
    noise sd ~ U[0.03, 0.08], clip i from generator seed 1000 + i."""
    import numpy as np

    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_clips):
        w = 2 * int(rng.integers(160, 961))
        h = 2 * int(rng.integers(120, 541))
        n = int(rng.integers(8, 33))
        sd = float(rng.uniform(0.03, 0.08))
        out.append((n, h, w, sd, 1000 + i))
    return out


def bench_batch(args):
    """64 float32 single-plane clips of uneven size and length, ftype 0, sbsize 16,
    sosize 12, tbsize 3: one autotuned batch (dft2_batch_*) run as one graph launch per
    step, beside the same jobs through dft2_process_batch with the plan's default tile."""
    import torch

    from dfttest2_amd import _lib, synth
    from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
    from dfttest2_amd.params import dfttest_params

    params = dfttest_params()
    clips = batch_clips()
    log(f"generating {len(clips)} clips ({sum(c[0] for c in clips)} frames)")
    xs = [synth.float_plane(n, h, w, seed, sd, device="cuda") for n, h, w, sd, seed in clips]
    outs = [torch.empty_like(x) for x in xs]
    jobs = [plane_job(x, o) for x, o in zip(xs, outs)]
    pixels = sum(n * h * w for n, h, w, _, _ in clips)
    plan = get_plan(params, "float", 32, Backend.MI355X())
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    batch = _lib.Batch(plan, jobs, autotune=True)
    tune_s = time.perf_counter() - t0
    info = batch.info()
    ms_graph = _time_steps(lambda: batch.launch(stream.cuda_stream), args.steps, args.warmup, stream)
    ms_plain = _time_steps(lambda: plan.process_batch(jobs, stream.cuda_stream), args.steps, args.warmup, stream)
    blocks = sum(n * blocks_per_plane(w, h) for n, h, w, _, _ in clips)
    log(f"batch: graph {ms_graph:.2f} ms, plain {ms_plain:.2f} ms, tile {info['tile_w']}x{info['tile_h']} "
        f"persistent {info['persistent']}")
    batch.close()
    return {"workload": "64 float32 clips, W U{320..1920}, H U{240..1080}, 8-32 frames, ftype 0, tbsize 3, "
                        "one autotuned graph launch per step",
            "mpx_s": round(pixels / (ms_graph / 1e3) / 1e6, 1), "ms_per_step": round(ms_graph, 3),
            "plain_process_batch_mpx_s": round(pixels / (ms_plain / 1e3) / 1e6, 1),
            "tflops_logical": round(blocks * flop_per_block(params.t_len) / (ms_graph / 1e3) / 1e12, 2),
            "autotune": {"seconds": round(tune_s, 2), "tile": [info["tile_w"], info["tile_h"]],
                         "persistent": info["persistent"], "candidates": info["n_candidates"]}}


def cpu_leg(planes_1080p, gpu_1080p, check4k, do_time=True):
    """The CPU leg: the reference's fused() (oracle/_ref) on a bounded sample of the
    headline clip, timed on this job's CPU share and on 1 core, its output compared with
    the GPU outputs of the same frames; and the FP64 oracle on the 4K check clip against
    both GPU forms, with the kernel's own hard decisions replayed."""
    import numpy as np

    from oracle import dfttest_np as onp
    from oracle import ref_lib
    from dfttest2_amd.params import dfttest_params

    out = {}
    params = dfttest_params()
    n = planes_1080p[0].shape[0]
    r = params.radius
    threads = cpu_threads()
    if ref_lib.available():
        if do_time:
            t0 = time.perf_counter()
            ref_lib.filter_plane(planes_1080p[0][:2], params, 8, False, threads=1)
            ref_lib.filter_plane(planes_1080p[1][:2], params, 8, False, threads=1)
            ref_lib.filter_plane(planes_1080p[2][:2], params, 8, False, threads=1)
            dt1 = time.perf_counter() - t0
        log(f"CPU baseline: reference fused() on {n} frames, {threads} threads")
        t0 = time.perf_counter()
        refs = [ref_lib.filter_plane(pl, params, 8, False, threads=threads) for pl in planes_1080p]
        dt = time.perf_counter() - t0
        if do_time:
            out["cpu_baseline"] = {
                "value": round(WIDTH * HEIGHT * n / dt / 1e6, 4), "unit": "Mpx/s", "cores": threads,
                "kind": "reference",
                "sample": f"{n} frames of the same 1080p YUV420 clip (3 planes), {dt:.2f} s wall",
                "value_1core": round(WIDTH * HEIGHT * 2 / dt1 / 1e6, 4),
                "host_cpus": os.cpu_count(),
            }
        # The sample is a clip of n frames: its last r frames see the temporal clamp at
        # frame n - 1, the full clip's do not, so frames [0, n - r) are compared.
        par = {}
        for prec, outs in gpu_1080p.items():
            worst, differ, total = 0, 0, 0
            for got, ref in zip(outs, refs):
                d = np.abs(got[: n - r].astype(np.int64) - ref[: n - r].astype(np.int64))
                worst = max(worst, int(d.max()))
                differ += int((d > 0).sum())
                total += d.size
            par[prec] = {"max_lsb": worst, "samples_differing": differ, "samples": total, "frames": n - r,
                         "planes": len(refs), "against": "reference fused() (oracle/_ref)"}
        out["parity_1080p"] = par
    if check4k is not None:
        params4k, check, gpu = check4k
        par = {}
        for prec, runs in gpu.items():
            worst, worst_replay, differ, total, flips, close = 0, 0, 0, 0, 0, 0.0
            for pl, (got_rec, dec_gpu, got) in zip(check, runs):
                ref = onp.process_plane(pl, params4k, 10, False)
                d = np.abs(got.astype(np.int64) - ref.astype(np.int64))
                worst = max(worst, int(d.max()))
                differ += int((d > 0).sum())
                total += d.size
                rep = onp.process_plane(pl, params4k, 10, False, decisions=dec_gpu).astype(np.int64)
                worst_replay = max(worst_replay, int(np.abs(got.astype(np.int64) - rep).max()),
                                   int(np.abs(got_rec.astype(np.int64) - rep).max()))
                for (dec, near, c, _, _), dg in zip(onp.hard_decisions(pl, params4k, 10, False,
                                                                       onp.COEF_ERROR_BOUND[prec]), dec_gpu):
                    miss = dec != dg
                    flips += int(miss.sum())
                    if miss.any():
                        close = max(close, float(c[miss].max()))
            par[prec] = {"max_lsb_vs_fp64": worst, "samples_differing": differ, "samples": total,
                         "max_lsb_vs_replay": worst_replay, "decisions_differing": flips,
                         "compared": "product path (process_batch) vs FP64; it and the decision-recording run "
                                     "vs the FP64 replay of the recorded decisions",
                         "largest_psd_distance_over_margin": round(close, 4),
                         "clip": "YUV420P10 640x360 + 2x320x180, 8 frames, same generator and filter"}
        out["parity_4k_check"] = par
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    ap.add_argument("--workload", choices=["all", "fp32_1080p", "split_1080p", "precision4k", "batch", "diff",
                                           "frame_host"], default="all",
                    help="all: the headline line with every part of the metric; the others run one part alone "
                         "(for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the path has no multi-GPU form here (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mb:
        import torch

        n = args.copy_probe_mb << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "dispatches": 3}), flush=True)
        return 0

    # Whole clips in one fused launch: scratch for the tile regions of every frame (about
    # 5.2 GB for the 240-frame 1080p clip, 10 GB for the 120-frame 4K clip, per plan).
    os.environ.setdefault("DFT2_SCRATCH_MB", "12288")
    import torch

    from dfttest2_amd import synth

    torch.cuda.set_device(0)
    if args.workload == "frame_host":
        return bench_frame_host(args)
    if args.workload == "precision4k":
        print(json.dumps({"metric": METRIC, "precision4k": bench_precision4k(args)[0]}), flush=True)
        return 0
    if args.workload == "diff":
        print(json.dumps({"metric": METRIC, "diff": bench_diff(args)}), flush=True)
        return 0
    if args.workload == "batch":
        print(json.dumps({"metric": METRIC, "batch": bench_batch(args)}), flush=True)
        return 0

    log(f"generating {FRAMES} frames of {WIDTH}x{HEIGHT} YUV420 u8 on cuda:0")
    planes = synth.yuv420_planes(FRAMES, WIDTH, HEIGHT, seed=1, bits=8, sd_lsb=6.0, device="cuda")
    keep = 2 * cpu_threads() if not args.no_cpu_baseline else 0
    keep = min(FRAMES, max(keep, 31))
    headline_prec = "split_bf16" if args.workload == "split_1080p" else "fp32"
    head, head_out = bench_1080p(args, headline_prec, planes, keep)
    gpu_1080p = {headline_prec: head_out}
    extras = {}
    check4k = None
    if args.workload == "all":
        split, split_out = bench_1080p(args, "split_bf16", planes, keep)
        gpu_1080p["split_bf16"] = split_out
        split["speedup_vs_fp32"] = round(head["ms_per_step"] / split["ms_per_step"], 3)
        split["frac_logical"] = round(split["tflops_logical"] / PEAK_BF16_MFMA_TFLOPS, 4)
        extras["split_bf16_1080p"] = split
    planes_cpu = [p[:keep].cpu().numpy() for p in planes]
    del planes
    torch.cuda.empty_cache()
    if args.workload == "all":
        extras["precision4k"], check4k = bench_precision4k(args)
        torch.cuda.empty_cache()
        extras["diff"] = bench_diff(args)
        extras["batch"] = bench_batch(args)
        torch.cuda.empty_cache()

    t_len = 3
    blocks = (blocks_per_plane(WIDTH, HEIGHT) + 2 * blocks_per_plane(WIDTH // 2, HEIGHT // 2)) * FRAMES
    peak, peak_basis = mfma_peak(t_len, headline_prec, exact_input=True)
    achieved = head["tflops_logical"]
    result = {
        "metric": METRIC,
        "value": head["mpx_s"],
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gradient + 3 moving sinusoids + N(0, 6 LSB)), resident in HBM",
        "config": {
            "workload": "DFTTest FP32 drop-in: YUV420 u8 1920x1080 + 2x960x540, 240 frames, all planes, "
                        "ftype 0 sigma 8 sbsize 16 sosize 12 tbsize 3 swin 0 twin 7 zmean",
            "frames": FRAMES,
            "fps": head["fps"],
            "precision": headline_prec,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "peak_basis": peak_basis,
            "traffic": FUSED_TRAFFIC_BYTES if headline_prec == "fp32" else None,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE on this kernel (fixed figure, "
                              "profiles/fused_kernel_1080p_yuv420_u8_fp32.md)",
            "kernel": "fused_kernel",
            "kernel_ms": head["kernel_ms"],
            "finalize_ms": head["finalize_ms"],
            "launches_per_step": head["launches_per_step"],
            "executed": head["executed"],
            "blocks_per_launch": blocks / head["launches_per_step"],
        },
    }
    result.update(extras)
    if not args.no_cpu_baseline:
        cl = cpu_leg(planes_cpu, gpu_1080p, check4k)
        if "cpu_baseline" in cl:
            result["cpu_baseline"] = cl.pop("cpu_baseline")
        result["parity"] = cl
    print(json.dumps(result), flush=True)
    bad = []
    for prec, v in result.get("parity", {}).get("parity_1080p", {}).items():
        if v["max_lsb"] > 1:
            bad.append(f"1080p {prec}: {v['max_lsb']} LSB against the reference")
    for prec, v in result.get("parity", {}).get("parity_4k_check", {}).items():
        if v["max_lsb_vs_replay"] > 1:
            bad.append(f"4K check {prec}: {v['max_lsb_vs_replay']} LSB against the decision replay")
    if bad:
        log("parity bar missed: " + "; ".join(bad))
        return 1
    return 0


def bench_frame_host(args):
    """The reference's own call pattern: one dft2_process_frame call per output frame and
    plane on host planes (the T source frames copied in, the output copied back), as
    DFTTestGetFrame runs it (hiprtc_source/source.cpp:445-692).  1080p YUV420 u8, 48 frames;
    one caller thread on one stream slot, and four threads on four slots (VapourSynth's
    fmParallel); pageable (numpy) and pinned (hipHostMalloc through torch) host memory; and
    the same through dft2_process_frame_keyed, which uploads each source plane once.  The
    timed run goes four times through the clip's frames, as 192 output frames whose keys
    are all distinct.  PCIe-inclusive frames per second; never the headline value."""
    import threading

    import numpy as np
    import torch

    from dfttest2_amd import _lib, synth
    from dfttest2_amd.params import dfttest_params

    frames = 48
    passes = 4
    p = dfttest_params()
    t_len, r = p.t_len, p.t_len // 2
    base = synth.yuv420_planes(frames, WIDTH, HEIGHT, seed=1, bits=8, sd_lsb=6.0)
    results = {}
    for memory in ("pageable", "pinned"):
        if memory == "pinned":
            planes = [b.pin_memory() for b in base]
            outs = [torch.empty_like(b).pin_memory() for b in base]
            ptr = lambda a, n: a[n].data_ptr()  # noqa: E731
        else:
            planes = [np.ascontiguousarray(b.numpy()) for b in base]
            outs = [np.empty_like(a) for a in planes]
            ptr = lambda a, n: a[n].ctypes.data  # noqa: E731

        def run(plan, frame_ids, keyed):
            for m in frame_ids:  # output frame m: frame n of pass m // frames
                n, base = m % frames, m - m % frames
                for k, (pl, out) in enumerate(zip(planes, outs)):
                    _, h, w = pl.shape
                    idx = [min(max(n - r + t, 0), frames - 1) for t in range(t_len)]
                    srcs = [ptr(pl, i) for i in idx]
                    if keyed:  # dft2_process_frame_keyed: key = (frame, plane)
                        plan.process_frame_keyed(srcs, [4 * (base + i) + k for i in idx], w, ptr(out, n), w, w, h)
                    else:
                        plan.process_frame(srcs, w, ptr(out, n), w, w, h, memory=1)

        for keyed in (False, True):
            for threads in (1, 4):
                plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, threads, 0)
                run(plan, range(4), keyed)
                if keyed:
                    plan.forget_frames()
                t0 = time.perf_counter()
                total = frames * passes
                if threads == 1:
                    run(plan, range(total), keyed)
                else:
                    pool = [threading.Thread(target=run, args=(plan, range(i, total, threads), keyed))
                            for i in range(threads)]
                    for th in pool:
                        th.start()
                    for th in pool:
                        th.join()
                dt = time.perf_counter() - t0
                plan.close()
                name = f"{memory}_{threads}thread" + ("_keyed" if keyed else "")
                results[name] = {"fps": round(total / dt, 1), "ms_per_frame": round(dt / total * 1e3, 3)}
                log(f"frame_host {memory}, {threads} thread(s){', keyed' if keyed else ''}: {total / dt:.1f} fps")
    print(json.dumps({"metric": METRIC, "workload": "frame_host: dft2_process_frame (and _keyed) per frame and plane "
                      "on host planes, 1080p YUV420 u8, 4 passes over 48 frames, DFTTest defaults", "unit": "fps",
                      "results": results}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
