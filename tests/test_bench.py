"""bench.py's contract and measurement arithmetic (CPU)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("t_len,flops", [(3, 124_416), (5, 200_448)])
def test_algorithmic_flops_per_block(t_len, flops):
    # 2 * (T * 13,824 + 576 T^2 + 576 T + 13,824): the dense separable DFT contractions
    assert bench.flop_per_block(t_len) == flops


def test_blocks_at_1080p_yuv420():
    # pad_num(1920) x pad_num(1080) and the two 960x540 chroma planes at block step 4
    assert bench.blocks_per_plane(1920, 1080) == 483 * 275 == 132_825
    assert bench.blocks_per_plane(960, 540) == 243 * 139 == 33_777
    assert bench.blocks_per_plane(1920, 1080) + 2 * bench.blocks_per_plane(960, 540) == 200_379


def test_roofline_peaks():
    # all-FP32 kernels price at the FP32 MFMA peak, the split form at the BF16 one; the FP32
    # form with 8-bit samples prices its x rDFT (2 * 4,608 flop per slice) at BF16
    assert bench.mfma_peak(3, "fp32", False)[0] == bench.PEAK_FP32_MFMA_TFLOPS
    assert bench.mfma_peak(3, "split_bf16", True)[0] == bench.PEAK_BF16_MFMA_TFLOPS
    peak = bench.mfma_peak(3, "fp32", True)[0]
    bf16 = 2 * 4608 * 3
    seconds = bf16 / 2500e12 + (124_416 - bf16) / 157.3e12
    assert peak == pytest.approx(124_416 / seconds / 1e12, rel=1e-3)
    assert bench.PEAK_FP32_MFMA_TFLOPS < peak < bench.PEAK_BF16_MFMA_TFLOPS


def test_mfma_counts_per_block():
    # FP32, 8-bit: 3 BF16 16x16x16 per slice (exact split of the x basis), 4 FP32 16x16x4
    # per y transform and per inverse contraction; split BF16: 2 and 3 x 16x16x16; 16-bit
    # samples take the 16x16x32 for the x transform.
    assert bench.mfma_per_block(3, "fp32", 1) == {"bf16_16x16x16": 9, "f32_16x16x4": 20}
    assert bench.mfma_per_block(3, "split_bf16", 1) == {"bf16_16x16x16": 21}
    assert bench.mfma_per_block(5, "fp32", 2) == {"bf16_16x16x32": 15, "f32_16x16x4": 28}
    assert bench.mfma_per_block(5, "split_bf16", 2) == {"bf16_16x16x32": 10, "bf16_16x16x16": 21}
    # Cross-frame form: 8 output frames of tbsize 5 compute 12 slices, 1.5 per block.
    assert bench.slices_per_block({"blocks": 8, "slice_spectra": 12}, 5) == 1.5
    assert bench.slices_per_block({"blocks": 0, "slice_spectra": 0}, 5) == 5.0
    assert bench.mfma_per_block(1.5, "fp32", 2) == {"bf16_16x16x32": 4.5, "f32_16x16x4": 14}
    e = bench.executed({"f32_16x16x4": 20, "bf16_16x16x16": 9}, 1e6, 1.0)
    assert e["flop_per_block"] == 20 * 2048 + 9 * 8192
    # 1e6 blocks x 784 cycles in 1 ms on 1024 SIMDs at 2.4 GHz
    assert e["mfma_pipe_busy"] == pytest.approx(1e6 * 784 / (1e-3 * 1024 * 2.4e9), rel=1e-3)


def test_cpu_share_respects_omp_threads(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    assert bench.cpu_threads() == min(3, len(os.sched_getaffinity(0)))
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.cpu_threads() == len(os.sched_getaffinity(0))
