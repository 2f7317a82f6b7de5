"""Throughput of the MI355X DFTTest path on its headline workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: YUV420 8-bit, 1920x1080 luma + 2 x 960x540 chroma, 240 frames, DFTTest
defaults (ftype 0, sigma 8, sbsize 16, sosize 12, tbsize 3, swin 0, twin 7, zmean),
all three planes, synthetic seeded clip resident in HBM before timing.  One step = the
whole 240-frame clip through the fused kernel (one dft2_process_batch call).
Prints one JSON line (metric, value in luma Mpx/s, roofline of the fused kernel, CPU
baseline of the reference block core timed on this host).
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
# HBM bytes per fused launch on the headline workload (one launch per step): rocprofv3
# --pmc FETCH_SIZE 4,562,310 KiB x 2 + WRITE_SIZE 4,797,210 KiB x 1, the factors being what
# the same counters read for a plain copy kernel of known size
# (profiles/fused_kernel_1080p_yuv420_u8_fp32.md).  Valid for the FP32 default workload.
FUSED_TRAFFIC_BYTES = (4_562_310 * 2 + 4_797_210) * 1024


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


def blocks_per_plane(w: int, h: int, step: int = 4) -> int:
    def pad_num(size):
        ps = size + ((16 - size % 16) if size % 16 else 0) + max(16 - step, step) * 2
        return (ps - 16) // step + 1

    return pad_num(w) * pad_num(h)


def cpu_baseline(planes_cpu, params, threads):
    """Reference GCC fused() (oracle/_ref) on a bounded sample, or the FP64 oracle."""
    from oracle import ref_lib

    n = planes_cpu[0].shape[0]
    if ref_lib.available():
        t0 = time.perf_counter()
        for pl in planes_cpu:
            ref_lib.filter_plane(pl, params, 8, False, threads=threads)
        dt = time.perf_counter() - t0
        kind, used = "reference", threads
    else:
        from oracle import dfttest_np

        planes_cpu = [pl[:2] for pl in planes_cpu]
        n = 2
        t0 = time.perf_counter()
        for pl in planes_cpu:
            dfttest_np.process_plane(pl, params, 8, False)
        dt = time.perf_counter() - t0
        kind, used = "port", 1
    return {
        "value": round(WIDTH * HEIGHT * n / dt / 1e6, 4),
        "unit": "Mpx/s",
        "cores": used,
        "kind": kind,
        "sample": f"{n} frames of the same 1080p YUV420 clip (3 planes), {dt:.2f} s wall",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["fp32", "split_bf16"], default="fp32")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    ap.add_argument("--workload", choices=["fp32_1080p", "precision4k", "batch", "diff", "frame_host"],
                    default="fp32_1080p",
                    help="fp32_1080p: the headline line; precision4k: 4K 10-bit FP32 vs split-BF16 with "
                         "errors vs FP64; batch: 64 uneven float clips as one autotuned graph launch; "
                         "diff: forward vs forward+backward of the differentiable filter; frame_host: "
                         "per-frame calls on host planes (PCIe included)")
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
        return

    if args.workload == "precision4k":
        return bench_precision4k(args)
    if args.workload == "diff":
        return bench_diff(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "frame_host":
        return bench_frame_host(args)

    # Whole clip in one fused launch: scratch for 240 frames of tile regions.
    os.environ.setdefault("DFT2_SCRATCH_MB", "8192")
    import torch

    from dfttest2_amd import synth
    from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
    from dfttest2_amd.params import dfttest_params

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    params = dfttest_params()
    print(f"[bench] generating {FRAMES} frames of {WIDTH}x{HEIGHT} YUV420 u8 on {dev}", file=sys.stderr, flush=True)
    planes = synth.yuv420_planes(FRAMES, WIDTH, HEIGHT, seed=1, bits=8, sd_lsb=6.0, device=dev)
    outs = [torch.empty_like(p) for p in planes]
    plan = get_plan(params, "int", 8, Backend.MI355X(precision=args.precision))
    jobs = [plane_job(p, o) for p, o in zip(planes, outs)]
    stream = torch.cuda.current_stream(dev)
    handle = stream.cuda_stream

    for _ in range(args.warmup):
        plan.process_batch(jobs, handle)
    torch.cuda.synchronize()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    # Step time from events on the launch stream; the library additionally records HIP
    # events around each fused / finalize launch (same stream) for the kernel roofline.
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    plan.set_timing(True)
    ev0.record(stream)
    for _ in range(args.steps):
        plan.process_batch(jobs, handle)
    ev1.record(stream)
    torch.cuda.synchronize()
    kt = plan.timing()
    plan.set_timing(False)
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    fused_ms = kt["fused_ms"] / max(1, kt["fused_launches"])
    print(f"[bench] {args.steps} steps, {ms_per_step:.3f} ms/step; fused kernel {fused_ms:.3f} ms/launch over "
          f"{kt['fused_launches']} launches, finalize {kt['finalize_ms'] / max(1, kt['finalize_launches']):.3f} ms",
          file=sys.stderr, flush=True)

    luma_mpx = WIDTH * HEIGHT * FRAMES / 1e6
    value = luma_mpx / (ms_per_step / 1e3)
    t_len = params.t_len
    blocks = (blocks_per_plane(WIDTH, HEIGHT) + 2 * blocks_per_plane(WIDTH // 2, HEIGHT // 2)) * FRAMES
    launches_per_step = kt["fused_launches"] / args.steps
    flops_per_launch = blocks * flop_per_block(t_len) / launches_per_step
    achieved = flops_per_launch / (fused_ms / 1e3) / 1e12
    peak, peak_basis = mfma_peak(t_len, args.precision, exact_input=True)  # 8-bit samples, separable window

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gradient + 3 moving sinusoids + N(0, 6 LSB)), resident in HBM",
        "config": {
            "workload": "DFTTest FP32 drop-in: YUV420 u8 1920x1080 + 2x960x540, 240 frames, all planes, "
                        "ftype 0 sigma 8 sbsize 16 sosize 12 tbsize 3 swin 0 twin 7 zmean",
            "frames": FRAMES,
            "fps": round(FRAMES / (ms_per_step / 1e3), 1),
            "precision": args.precision,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "peak_basis": peak_basis,
            "traffic": FUSED_TRAFFIC_BYTES if args.precision == "fp32" else None,
            "kernel": "fused_kernel",
            "kernel_ms": round(fused_ms, 4),
            "launches_per_step": launches_per_step,
        },
    }
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        n_cpu = 2 * threads
        planes_cpu = [p[:n_cpu].cpu().numpy() for p in planes]
        print(f"[bench] CPU baseline on {n_cpu} frames, {threads} threads", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(planes_cpu, params, threads)
    print(json.dumps(result), flush=True)


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


def bench_precision4k(args):
    """4K YUV420P10 (3840x2160 + 2x1920x1080, 120 frames), ftype 1, tbsize 5, sigma curve
    0.0:4 0.5:8 1.0:16: FP32 vs split-BF16 throughput, and both forms' error against the FP64
    oracle on a 640x360x8 clip of the same generator and parameters."""
    os.environ.setdefault("DFT2_SCRATCH_MB", "8192")
    import numpy as np
    import torch

    from dfttest2_amd import synth
    from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
    from dfttest2_amd.params import dfttest_params
    from oracle import dfttest_np

    w, h, frames = 3840, 2160, 120
    params = dfttest_params(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])
    dev = torch.device("cuda", 0)
    print(f"[bench] generating {frames} frames of {w}x{h} YUV420P10", file=sys.stderr, flush=True)
    planes = synth.yuv420_planes(frames, w, h, seed=2, bits=10, sd_lsb=12.0, device=dev)
    outs = [torch.empty_like(p) for p in planes]
    jobs = [plane_job(p, o) for p, o in zip(planes, outs)]
    stream = torch.cuda.current_stream(dev)
    res = {}
    for prec in ("fp32", "split_bf16"):
        plan = get_plan(params, "int", 10, Backend.MI355X(precision=prec))
        ms = _time_steps(lambda: plan.process_batch(jobs, stream.cuda_stream), args.steps, args.warmup, stream)
        res[prec] = {"ms_per_step": round(ms, 3), "fps": round(frames / (ms / 1e3), 1),
                     "mpx_s": round(w * h * frames / (ms / 1e3) / 1e6, 1)}
        print(f"[bench] {prec}: {ms:.2f} ms per {frames} frames", file=sys.stderr, flush=True)
    # error yardstick: FP64 oracle on a smaller clip of the same generator and parameters
    small = [pl.numpy() for pl in synth.yuv420_planes(8, 640, 360, seed=3, bits=10, sd_lsb=12.0)]
    for prec in ("fp32", "split_bf16"):
        plan = get_plan(params, "int", 10, Backend.MI355X(precision=prec))
        worst, flips, total = 0, 0, 0
        for pl in small:
            x = torch.from_numpy(pl).to(dev)
            o = torch.empty_like(x)
            plan.process_batch([plane_job(x, o)], stream.cuda_stream)
            torch.cuda.synchronize()
            got = o.cpu().numpy().astype(np.int64)
            ref = dfttest_np.process_plane(pl, params, 10, False).astype(np.int64)
            worst = max(worst, int(np.abs(got - ref).max()))
            flips += int((got != ref).sum())
            total += got.size
        res[prec].update({"max_lsb_vs_fp64": worst, "samples_differing_vs_fp64": flips, "samples": total})
    res["speedup_split_bf16"] = round(res["fp32"]["ms_per_step"] / res["split_bf16"]["ms_per_step"], 3)
    print(json.dumps({"metric": METRIC, "workload": "precision4k: YUV420P10 3840x2160 + 2x1920x1080, 120 frames, "
                      "ftype 1, tbsize 5, slocation 0.0:4 0.5:8 1.0:16", "unit": "fps", "results": res}), flush=True)


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
    os.environ.setdefault("DFT2_SCRATCH_MB", "8192")
    import torch

    from dfttest2_amd import _lib, synth
    from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
    from dfttest2_amd.params import dfttest_params

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    params = dfttest_params()
    clips = batch_clips()
    print(f"[bench] generating {len(clips)} clips ({sum(c[0] for c in clips)} frames)", file=sys.stderr, flush=True)
    xs = [synth.float_plane(n, h, w, seed, sd, device=dev) for n, h, w, sd, seed in clips]
    outs = [torch.empty_like(x) for x in xs]
    jobs = [plane_job(x, o) for x, o in zip(xs, outs)]
    pixels = sum(n * h * w for n, h, w, _, _ in clips)
    plan = get_plan(params, "float", 32, Backend.MI355X())
    stream = torch.cuda.current_stream(dev)

    t0 = time.perf_counter()
    batch = _lib.Batch(plan, jobs, autotune=True)
    tune_s = time.perf_counter() - t0
    info = batch.info()
    print(f"[bench] autotune {tune_s:.2f} s: tile {info['tile_w']}x{info['tile_h']} persistent {info['persistent']}",
          file=sys.stderr, flush=True)
    ms_graph = _time_steps(lambda: batch.launch(stream.cuda_stream), args.steps, args.warmup, stream)
    ms_plain = _time_steps(lambda: plan.process_batch(jobs, stream.cuda_stream), args.steps, args.warmup, stream)
    blocks = sum(n * blocks_per_plane(w, h) for n, h, w, _, _ in clips)
    achieved = blocks * flop_per_block(params.t_len) / (ms_graph / 1e3) / 1e12
    cands = sorted(batch.candidates(), key=lambda c: c["ms"])
    print(json.dumps({
        "metric": METRIC, "workload": "batch: 64 float32 clips, W U{320..1920}, H U{240..1080}, 8-32 frames, "
        "ftype 0 sbsize 16 sosize 12 tbsize 3, one autotuned graph launch per step",
        "unit": "Mpx/s", "value": round(pixels / (ms_graph / 1e3) / 1e6, 1), "ms_per_step": round(ms_graph, 3),
        "pixels_per_step": pixels, "frames_per_step": sum(c[0] for c in clips),
        "plain_process_batch": {"ms_per_step": round(ms_plain, 3),
                                "mpx_s": round(pixels / (ms_plain / 1e3) / 1e6, 1)},
        "speedup_vs_plain": round(ms_plain / ms_graph, 3),
        "tflops_step": round(achieved, 2),
        "autotune": {"seconds": round(tune_s, 2), "chosen": {k: info[k] for k in ("tile_w", "tile_h", "persistent",
                     "workgroups_per_cu", "lds_bytes", "n_chunks", "work_items")},
                     "candidates": [[c["tile_w"], c["tile_h"], c["persistent"], c["workgroups_per_cu"], round(c["ms"], 3)]
                                    for c in cands]},
    }), flush=True)
    batch.close()


def bench_frame_host(args):
    """The reference's own call pattern: one dft2_process_frame call per output frame and
    plane on host planes (the T source frames copied in, the output copied back), as
    DFTTestGetFrame runs it (hiprtc_source/source.cpp:445-692).  1080p YUV420 u8, 48 frames;
    one caller thread on one stream slot, and four threads on four slots (VapourSynth's
    fmParallel); pageable (numpy) and pinned (hipHostMalloc through torch) host memory.
    PCIe-inclusive frames per second; never the headline value."""
    import threading

    import numpy as np
    import torch

    from dfttest2_amd import _lib, synth
    from dfttest2_amd.params import dfttest_params

    frames = 48
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

        def run(plan, frame_ids):
            for n in frame_ids:
                for pl, out in zip(planes, outs):
                    _, h, w = pl.shape
                    srcs = [ptr(pl, min(max(n - r + t, 0), frames - 1)) for t in range(t_len)]
                    plan.process_frame(srcs, w, ptr(out, n), w, w, h, memory=1)

        for threads in (1, 4):
            plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, threads, 0)
            run(plan, range(4))
            t0 = time.perf_counter()
            if threads == 1:
                run(plan, range(frames))
            else:
                pool = [threading.Thread(target=run, args=(plan, range(i, frames, threads))) for i in range(threads)]
                for th in pool:
                    th.start()
                for th in pool:
                    th.join()
            dt = time.perf_counter() - t0
            plan.close()
            results[f"{memory}_{threads}thread"] = {"fps": round(frames / dt, 1), "ms_per_frame": round(dt / frames * 1e3, 3)}
            print(f"[bench] frame_host {memory}, {threads} thread(s): {frames / dt:.1f} fps", file=sys.stderr, flush=True)
    print(json.dumps({"metric": METRIC, "workload": "frame_host: dft2_process_frame per frame and plane on host "
                      "planes, 1080p YUV420 u8, 48 frames, DFTTest defaults", "unit": "fps",
                      "results": results}), flush=True)


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
        loss = torch.mean((dfttest_diff(x, sigma, p) - clean) ** 2)
        loss.backward()

    ms_f = _time_steps(fwd, args.steps, args.warmup, stream)
    ms_fb = _time_steps(fwd_bwd, args.steps, args.warmup, stream)
    mpx = 8 * 16 * 256 * 256 / 1e6
    print(json.dumps({"metric": METRIC, "workload": "diff: 8 x 256x256x16 float32 clips, Wiener, MSE loss",
                      "unit": "Mpx/s", "fwd_mpx_s": round(mpx / (ms_f / 1e3), 1),
                      "fwd_bwd_mpx_s": round(mpx / (ms_fb / 1e3), 1),
                      "ratio_fwd_bwd_to_fwd": round(ms_f / ms_fb, 3), "ms_fwd": round(ms_f, 3),
                      "ms_fwd_bwd": round(ms_fb, 3)}), flush=True)


if __name__ == "__main__":
    main()
