"""The keyed per-frame entry (dft2_process_frame_keyed) on the GPU: outputs equal to the
whole-clip path for sequential, shuffled and repeated frame orders, source planes of two
geometries sharing the plan's pool, replacement of the least recently used planes,
dft2_plan_forget_frames, pinned planes, concurrent calls from several threads over
several stream slots, padded rows (pageable and pinned, 8- and 10-bit), and the generic
block path at radius 4 (also through dft2_process_frame on device planes, whose
frame-pointer table covers 9 slices)."""

import sys
import threading

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.dfttest2 import Backend, VideoClip, filter_clip, get_plan
from dfttest2_amd.params import dfttest_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def sources(clip, n, r):
    """The T source planes of output frame n, clamped at the clip's ends, and their indices."""
    idx = [min(max(n - r + t, 0), clip.shape[0] - 1) for t in range(2 * r + 1)]
    return [np.ascontiguousarray(clip[i]) for i in idx], idx


def keyed_frame(plan, clip, n, r, plane_id, planes=None):
    srcs, idx = sources(clip, n, r)
    if planes is not None:  # caller-held copies (kept alive, e.g. pinned)
        srcs = [planes[i] for i in idx]
    h, w = clip.shape[1:]
    out = np.zeros((h, w), clip.dtype)
    bps = clip.dtype.itemsize
    plan.process_frame_keyed([s.ctypes.data if isinstance(s, np.ndarray) else s.data_ptr() for s in srcs],
                             [4 * i + plane_id for i in idx], w * bps, out.ctypes.data, w * bps, w, h)
    return out


@pytest.mark.parametrize("tbsize", [3, 5])
def test_keyed_frames_match_whole_clip(tbsize):
    announce("keyed frames", tbsize)
    p = dfttest_params(tbsize=tbsize)
    r = (tbsize - 1) // 2
    luma = synth.int_plane(12, 72, 96, seed=81, bits=8, sd_lsb=6.0).numpy()
    chroma = synth.int_plane(12, 36, 48, seed=82, bits=8, sd_lsb=6.0).numpy()
    ref = filter_clip(VideoClip([luma, chroma], "int", 8), p).planes
    plan = get_plan(p, "int", 8, Backend.MI355X(num_streams=1))
    plan.forget_frames()
    # sequential (each call uploads one new plane of each geometry), then shuffled with
    # repeats (hits, misses and replacement: 24 keys against a pool of 2 T planes)
    order = list(range(12)) + [int(v) for v in np.random.default_rng(5).integers(0, 12, 30)]
    for n in order:
        np.testing.assert_array_equal(keyed_frame(plan, luma, n, r, 0), ref[0][n], err_msg=f"luma frame {n}")
        np.testing.assert_array_equal(keyed_frame(plan, chroma, n, r, 1), ref[1][n], err_msg=f"chroma frame {n}")


def test_forget_frames_uploads_new_content_under_old_keys():
    announce("keyed frames: forget")
    p = dfttest_params()
    a = synth.int_plane(6, 64, 80, seed=83, bits=8, sd_lsb=6.0).numpy()
    b = synth.int_plane(6, 64, 80, seed=84, bits=8, sd_lsb=6.0).numpy()
    ref_a = filter_clip(VideoClip([a], "int", 8), p).planes[0]
    ref_b = filter_clip(VideoClip([b], "int", 8), p).planes[0]
    plan = get_plan(p, "int", 8, Backend.MI355X(num_streams=1))
    plan.forget_frames()
    for n in range(6):
        np.testing.assert_array_equal(keyed_frame(plan, a, n, 1, 0), ref_a[n])
    plan.forget_frames()  # the same keys now name clip b
    for n in range(6):
        np.testing.assert_array_equal(keyed_frame(plan, b, n, 1, 0), ref_b[n])


def test_keyed_frames_from_pinned_planes():
    announce("keyed frames: pinned planes")
    p = dfttest_params(tbsize=5)
    x = synth.int_plane(8, 60, 76, seed=85, bits=10, sd_lsb=12.0).numpy()
    ref = filter_clip(VideoClip([x], "int", 10), p).planes[0]
    pinned = [torch.from_numpy(x[i]).pin_memory() for i in range(8)]
    plan = get_plan(p, "int", 10, Backend.MI355X(num_streams=1))
    plan.forget_frames()
    h, w = x.shape[1:]
    for n in (0, 1, 2, 7, 3, 3, 6):
        idx = [min(max(n - 2 + t, 0), 7) for t in range(5)]
        out = torch.zeros((h, w), dtype=torch.int16).pin_memory()
        plan.process_frame_keyed([pinned[i].data_ptr() for i in idx], [4 * i for i in idx], 2 * w, out.data_ptr(),
                                 2 * w, w, h)
        np.testing.assert_array_equal(out.numpy().view(np.uint16), ref[n], err_msg=f"frame {n}")


@pytest.mark.parametrize("memory", ["pageable", "pinned"])
def test_keyed_frames_with_row_pitch(memory):
    # Planes whose rows are padded (pitch 96 samples for 70-sample rows, as VapourSynth's
    # frames are), source and output alike, 8-bit and 16-bit containers.
    announce("keyed frames: padded rows", memory)
    for bits in (8, 10):
        p = dfttest_params()
        x = synth.int_plane(5, 52, 70, seed=88 + bits, bits=bits, sd_lsb=6.0).numpy()
        ref = filter_clip(VideoClip([x], "int", bits), p).planes[0]
        plan = get_plan(p, "int", bits, Backend.MI355X(num_streams=1))
        plan.forget_frames()
        n_f, h, w = x.shape
        pitch = 96
        dt = x.dtype
        if memory == "pinned":
            tdt = torch.uint8 if dt == np.uint8 else torch.int16
            bufs = [torch.zeros((h, pitch), dtype=tdt).pin_memory() for _ in range(n_f)]
            views = [b.numpy().view(dt) for b in bufs]
            out_t = torch.zeros((h, pitch), dtype=tdt).pin_memory()
            out = out_t.numpy().view(dt)
        else:
            views = [np.zeros((h, pitch), dt) for _ in range(n_f)]
            out = np.zeros((h, pitch), dt)
        for i in range(n_f):
            views[i][:, :w] = x[i]
            views[i][:, w:] = dt.type(7)  # padding the filter must not read
        bps = dt.itemsize
        for n in range(n_f):
            idx = [min(max(n - 1 + t, 0), n_f - 1) for t in range(3)]
            out[:] = 0
            plan.process_frame_keyed([views[i].ctypes.data for i in idx], [4 * i for i in idx], pitch * bps,
                                     out.ctypes.data, pitch * bps, w, h)
            np.testing.assert_array_equal(out[:, :w], ref[n], err_msg=f"{memory} {bits}-bit frame {n}")
            assert not out[:, w:].any(), "samples written past the row"


def test_keyed_frames_from_threads():
    # Six threads over three stream slots: a plane one call uploads is used by others,
    # some while its upload is still in flight.
    announce("keyed frames: six threads, three slots")
    p = dfttest_params()
    x = synth.int_plane(30, 90, 120, seed=86, bits=8, sd_lsb=6.0).numpy()
    ref = filter_clip(VideoClip([x], "int", 8), p).planes[0]
    plan = get_plan(p, "int", 8, Backend.MI355X(num_streams=3))
    plan.forget_frames()
    outs = {}
    errors = []

    def worker(k):
        try:
            for rep in range(2):
                for n in range(k, 30, 6):
                    outs[(rep, n)] = keyed_frame(plan, x, n, 1, 0)
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (rep, n), out in outs.items():
        np.testing.assert_array_equal(out, ref[n], err_msg=f"pass {rep} frame {n}")


def test_generic_radius4_keyed_and_device_frames():
    announce("generic path, tbsize 9: keyed and device frames")
    p = dfttest_params(sbsize=8, sosize=4, tbsize=9)
    x = synth.int_plane(11, 40, 52, seed=87, bits=8, sd_lsb=6.0).numpy()
    ref = filter_clip(VideoClip([x], "int", 8), p).planes[0]
    plan = get_plan(p, "int", 8, Backend.MI355X())
    plan.forget_frames()
    h, w = x.shape[1:]
    for n in (0, 4, 5, 10, 6):
        np.testing.assert_array_equal(keyed_frame(plan, x, n, 4, 0), ref[n], err_msg=f"keyed frame {n}")
        srcs, _ = sources(x, n, 4)
        dev = [torch.from_numpy(s).cuda() for s in srcs]
        dout = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        plan.process_frame([d.data_ptr() for d in dev], w, dout.data_ptr(), w, w, h, memory=0)
        np.testing.assert_array_equal(dout.cpu().numpy(), ref[n], err_msg=f"device frame {n}")


def test_keyed_entry_argument_checks():
    announce("keyed frames: argument checks")
    p = dfttest_params()
    plan = get_plan(p, "int", 8, Backend.MI355X())
    buf = np.zeros(64 * 64, np.uint8)
    out = np.zeros(64 * 64, np.uint8)
    with pytest.raises(_lib.Dft2Error, match="row stride"):
        plan.process_frame_keyed([buf.ctypes.data] * 3, [1, 2, 3], 16, out.ctypes.data, 64, 24, 40)
    with pytest.raises(_lib.Dft2Error, match="positive"):
        plan.process_frame_keyed([buf.ctypes.data] * 3, [1, 2, 3], 64, out.ctypes.data, 64, 0, 40)
    assert not out.any()
