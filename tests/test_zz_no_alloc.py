"""dft2_process_batch_noalloc (the entry point that allocates nothing) against
dft2_process_batch, through the C ABI: two calls in a row, the second with a smaller
input, outputs identical (the same kernels over the same work items); one call whose
frames run in several chunks of the fixed scratch array; and the size limits."""

import sys

import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import plane_job
from dfttest2_amd.params import dfttest_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def run_both(plan, clips, stream):
    ref = [torch.empty_like(x) for x in clips]
    out = [torch.full_like(x, 7) for x in clips]
    plan.process_batch([plane_job(x, o) for x, o in zip(clips, ref)], stream)
    plan.process_batch_noalloc([plane_job(x, o) for x, o in zip(clips, out)], stream)
    torch.cuda.synchronize()
    return ref, out


@pytest.mark.parametrize("fmt", ["u8", "f32"])
def test_noalloc_matches_process_batch_on_two_calls(fmt):
    announce("noalloc vs process_batch, two calls, second smaller", fmt)
    p = dfttest_params()
    if fmt == "u8":
        plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, 1, 0)
        mk = lambda n, h, w, s: synth.int_plane(n, h, w, s, 8, 6.0).cuda()  # noqa: E731
    else:
        plan = _lib.Plan(p, _lib.DFT2_SAMPLE_FLOAT, 32, 0, 1, 0)
        mk = lambda n, h, w, s: synth.float_plane(n, h, w, s, 0.05).cuda()  # noqa: E731
    stream = torch.cuda.current_stream().cuda_stream
    first = [mk(6, 540, 960, 81), mk(6, 270, 480, 82), mk(6, 270, 480, 83)]
    ref, out = run_both(plan, first, stream)
    for r, o in zip(ref, out):
        assert torch.equal(r, o)
    second = [mk(3, 100, 120, 84)]
    ref, out = run_both(plan, second, stream)
    assert torch.equal(ref[0], out[0])
    plan.close()


def test_noalloc_runs_frames_in_chunks():
    # 1080p luma needs about 14 MiB of tile regions per frame, so 12 frames take several
    # chunks of the 64 MiB array, each reusing it after the previous one.
    announce("noalloc chunked: 12 frames of 1920x1080 u8, then 2 frames of 64x80")
    p = dfttest_params(tbsize=5)
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, 1, 0)
    stream = torch.cuda.current_stream().cuda_stream
    x = synth.int_plane(12, 1080, 1920, 85, 8, 6.0).cuda()
    ref, out = run_both(plan, [x], stream)
    assert torch.equal(ref[0], out[0])
    y = synth.int_plane(2, 64, 80, 86, 8, 6.0).cuda()
    ref, out = run_both(plan, [y], stream)
    assert torch.equal(ref[0], out[0])
    plan.close()


def test_noalloc_rejects_sizes_beyond_its_array():
    announce("noalloc size limits")
    p = dfttest_params()
    plan = _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8, 0, 1, 0)
    big = torch.zeros((1, 4320, 7680), dtype=torch.uint8, device="cuda")
    out = torch.zeros_like(big)
    with pytest.raises(_lib.Dft2Error, match="MiB of tile scratch"):
        plan.process_batch_noalloc([plane_job(big, out)], None)
    small = torch.zeros((1, 40, 40), dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.Dft2Error, match="at most 64 jobs"):
        plan.process_batch_noalloc([plane_job(small, small)] * 65, None)
    torch.cuda.synchronize()
    assert not out.any()
    plan.close()
