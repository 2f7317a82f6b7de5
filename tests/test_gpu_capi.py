"""C ABI behaviour on the GPU: stream-slot ordering across streams, argument checks of
the host-frame entry, and the bounded kernel-timing pool."""

import sys

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import Backend, get_plan, plane_job
from dfttest2_amd.params import dfttest_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def test_host_frame_after_side_stream_batch_waits_for_the_slot():
    # One stream slot: a batch call queued on a torch side stream is still running when a
    # host-frame call on the same plan starts.  The host call must not reuse the slot's
    # scratch and job table under the batch; both outputs must equal separate runs.
    announce("slot ordering: side-stream batch, then host frame, num_streams 1")
    p = dfttest_params()
    x = synth.int_plane(48, 540, 960, seed=71, bits=8, sd_lsb=6.0).cuda()
    small = synth.int_plane(3, 64, 80, seed=72, bits=8, sd_lsb=6.0).numpy()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, 1, 0)

    ref_big = torch.empty_like(x)
    plan.process_batch([plane_job(x, ref_big)], None)
    _, h, w = small.shape
    ref_small = np.zeros((h, w), np.uint8)
    srcs = [np.ascontiguousarray(small[i]) for i in range(3)]
    plan.process_frame([s.ctypes.data for s in srcs], w, ref_small.ctypes.data, w, w, h, memory=1)

    side = torch.cuda.Stream()
    for rep in range(3):
        big = torch.empty_like(x)
        torch.cuda.synchronize()
        plan.process_batch([plane_job(x, big)], side.cuda_stream)  # queued, not waited for
        out = np.zeros((h, w), np.uint8)
        plan.process_frame([s.ctypes.data for s in srcs], w, out.ctypes.data, w, w, h, memory=1)
        side.synchronize()
        np.testing.assert_array_equal(out, ref_small)
        assert torch.equal(big, ref_big), f"rep {rep}: side-stream batch output changed"
    plan.close()


def test_device_frame_after_side_stream_batch_on_plan_stream():
    # The same with a device-plane frame call on the plan's own stream (stream = NULL).
    announce("slot ordering: side-stream batch, then device frame on the plan stream")
    p = dfttest_params()
    x = synth.int_plane(48, 540, 960, seed=73, bits=8, sd_lsb=6.0).cuda()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, 1, 0)
    ref_big = torch.empty_like(x)
    plan.process_batch([plane_job(x, ref_big)], None)
    n, h, w = x.shape
    ref_frame = ref_big[1].clone()
    side = torch.cuda.Stream()
    big = torch.empty_like(x)
    out = torch.empty((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    plan.process_batch([plane_job(x, big)], side.cuda_stream)
    plan.process_frame([x[i].data_ptr() for i in range(3)], w, out.data_ptr(), w, w, h, memory=0)
    side.synchronize()
    assert torch.equal(out, ref_frame)
    assert torch.equal(big, ref_big)
    plan.close()


@pytest.mark.parametrize(
    "src_stride,dst_stride,width,height,memory,msg",
    [
        (64, 32, 24, 40, 1, "row stride"),      # dst stride smaller than the row (48 bytes)
        (32, 64, 24, 40, 1, "row stride"),      # src stride smaller than the row
        (64, -64, 24, 40, 1, "row stride"),     # negative stride
        (64, 64, -5, 40, 1, "positive"),        # negative width
        (64, 64, 24, 0, 1, "positive"),         # zero height
        (65, 64, 24, 40, 1, "multiples"),       # 16-bit samples: odd byte stride
        (64, 64, 4, 40, 1, "too small"),        # narrower than the mirror padding
        (64, 64, 24, 40, 2, "memory"),
    ],
)
def test_frame_entry_rejects_bad_geometry(src_stride, dst_stride, width, height, memory, msg):
    announce("process_frame geometry", src_stride, dst_stride, width, height, memory)
    p = dfttest_params()
    plan = get_plan(p, "int", 10, Backend.MI355X())
    buf = np.zeros(64 * 64, np.uint16)
    out = np.zeros(64 * 64, np.uint16)
    with pytest.raises(_lib.Dft2Error, match=msg):
        plan.process_frame([buf.ctypes.data] * 3, src_stride, out.ctypes.data, dst_stride, width, height, memory)
    assert not out.any()  # nothing was written


@pytest.mark.parametrize("tbsize,frames,per_position", [(3, 10, 30), (5, 10, 18), (5, 16, 24), (5, 3, 7)])
def test_timing_counts_blocks_and_slice_spectra(tbsize, frames, per_position):
    # One-frame form (tbsize 3): T slices per output block.  Cross-frame form (tbsize 5,
    # 8 frames per group): frames + 4 per group of consecutive frames at each block
    # position, so 10 frames = 12 + 6 slices, 16 = 12 + 12, 3 = 7.
    announce("timing: blocks and slice spectra", tbsize, frames)
    p = dfttest_params(tbsize=tbsize)
    x = synth.float_plane(frames, 48, 64, seed=75).cuda()
    o = torch.empty_like(x)
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32, 0, 1, 0)
    plan.set_timing(True)
    plan.process_batch([plane_job(x, o)], torch.cuda.current_stream().cuda_stream)
    t = plan.timing()
    positions = ((48 + 24) // 4 - 3) * ((64 + 24) // 4 - 3)  # padded by 12 on each side
    assert t["blocks"] == positions * frames
    assert t["slice_spectra"] == positions * per_position
    plan.close()


def test_timing_totals_over_many_launches():
    # More launch pairs than the pending pool holds: the oldest are folded into the totals
    # as new ones come, and the totals still count every launch.
    announce("timing pool: 150 launch pairs")
    p = dfttest_params()
    x = synth.float_plane(2, 48, 64, seed=74).cuda()
    o = torch.empty_like(x)
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32, 0, 1, 0)
    stream = torch.cuda.current_stream().cuda_stream
    plan.set_timing(True)
    for _ in range(150):
        plan.process_batch([plane_job(x, o)], stream)
    t = plan.timing()
    assert t["fused_launches"] == 150 and t["finalize_launches"] == 150
    assert t["fused_ms"] > 0 and t["finalize_ms"] > 0
    plan.set_timing(True)  # restarts from zero
    plan.process_batch([plane_job(x, o)], stream)
    assert plan.timing()["fused_launches"] == 1
    plan.set_timing(False)
    plan.process_batch([plane_job(x, o)], stream)
    assert plan.timing()["fused_launches"] == 1
    plan.close()
