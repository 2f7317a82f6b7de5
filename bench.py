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


def flop_per_block(t_len: int) -> int:
    # Dense separable DFT contractions per block: forward x rDFT + y cDFT per slice,
    # temporal forward and centre-slice inverse, spatial inverse of the centre slice.
    return 2 * (t_len * 13824 + 576 * t_len * t_len + 576 * t_len + 13824)


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

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.process_batch(jobs, handle)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    print(f"[bench] {args.steps} steps, {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    luma_mpx = WIDTH * HEIGHT * FRAMES / 1e6
    value = luma_mpx / (ms_per_step / 1e3)
    t_len = params.t_len
    blocks = (blocks_per_plane(WIDTH, HEIGHT) + 2 * blocks_per_plane(WIDTH // 2, HEIGHT // 2)) * FRAMES
    flops = blocks * flop_per_block(t_len)
    achieved = flops / (ms_per_step / 1e3) / 1e12

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
            "peak": PEAK_FP32_MFMA_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_MFMA_TFLOPS, 4),
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        n_cpu = 2 * threads
        planes_cpu = [p[:n_cpu].cpu().numpy() for p in planes]
        print(f"[bench] CPU baseline on {n_cpu} frames, {threads} threads", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(planes_cpu, params, threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
