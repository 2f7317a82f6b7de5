"""Differentiable path on the GPU: forward of dfttest_diff against the oracle, and the
backward kernel's gradients (input planes and sigma) against the FP64 torch autograd
graph of the same pipeline (oracle/dfttest_torch.py, itself pinned against finite
differences).  Bar: relative error <= 1e-4 (norm-wise).  The backward kernels recompute
the block spectra in FP32, as the forward computes them, so each gain and its derivative
is the forward's; their linear adjoint contractions (split BF16) leave about 1e-5."""

import sys

import numpy as np
import pytest

from dfttest2_amd import synth
from dfttest2_amd.autograd import dfttest_diff, diff_params
from oracle import dfttest_np as onp
from oracle import dfttest_torch as otorch

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def announce(*what):
    print("\n[case]", *what, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-30))


@pytest.mark.parametrize("tbsize", [1, 3, 5])
def test_diff_forward_matches_oracle(tbsize):
    announce("diff forward", tbsize)
    p = diff_params(tbsize=tbsize)
    x = synth.float_plane(5, 44, 52, 7)
    rng = np.random.default_rng(5)
    sigma = np.abs(rng.normal(8.0, 4.0, (p.t_len, 16, 9))).astype(np.float32)  # asymmetric
    y = dfttest_diff(x.cuda(), torch.from_numpy(sigma).cuda(), p)
    p.sigma = (sigma.astype(np.float64) * p.extra["wscale"]).reshape(-1)
    ref = onp.process_plane(x.numpy(), p, 32, True)
    assert float(np.max(np.abs(y.cpu().numpy() - ref))) <= 1e-5


@pytest.mark.parametrize("tbsize,shape,sosize", [(3, (2, 5, 40, 44), 12), (1, (1, 3, 36, 36), 12), (5, (1, 6, 34, 40), 12),
                                                 (3, (1, 4, 38, 42), 14), (3, (1, 4, 44, 36), 8),
                                                 (3, (1, 19, 100, 40), 12), (3, (2, 9, 44, 36), 8),
                                                 (3, (1, 10, 52, 60), 14), (5, (1, 13, 40, 44), 12),
                                                 (7, (1, 11, 36, 36), 12), (3, (1, 1, 36, 40), 12),
                                                 (3, (2, 2, 36, 40), 12), (5, (1, 2, 36, 40), 12)])
def test_gradients_match_fp64_autograd(tbsize, shape, sosize):
    # block steps 4 (default), 2 (eight blocks cover a pixel per axis) and 8; tbsize >= 3
    # takes the cross-frame backward (8 output frames per work item): clips of up to 8
    # frames are one group, 19 frames are groups of 8, 8 and 3, 9 frames 8 and 1; the
    # 100-row plane has two tile rows; clips of 1 and 2 frames clamp most slices
    announce("diff backward", tbsize, shape, sosize)
    p = diff_params(tbsize=tbsize, sosize=sosize)
    b, n, h, w = shape
    clean = torch.stack([synth.clean_plane(n, h, w, 30 + i) for i in range(b)])
    noisy = clean + torch.stack([synth.noise((n, h, w), 0.05, 40 + i) for i in range(b)])
    rng = np.random.default_rng(9)
    sigma0 = torch.from_numpy(np.abs(rng.normal(8.0, 3.0, (p.t_len, 16, 9))).astype(np.float32))

    xg = noisy.cuda().requires_grad_(True)
    sg = sigma0.cuda().requires_grad_(True)
    loss = torch.mean((dfttest_diff(xg, sg, p) - clean.cuda()) ** 2)
    loss.backward()

    xr = noisy.double().requires_grad_(True)
    sr = sigma0.double().requires_grad_(True)
    wscale = p.extra["wscale"]
    yr = torch.stack([otorch.filter_plane(xr[i], sr * wscale, p) for i in range(b)])
    loss_r = torch.mean((yr - clean.double()) ** 2)
    loss_r.backward()

    assert float(loss.detach()) == pytest.approx(float(loss_r.detach()), rel=1e-5)
    ex, es = rel_err(xg.grad, xr.grad), rel_err(sg.grad, sr.grad)
    print(f"vs FP64: input {ex:.2e}, sigma {es:.2e}", flush=True)
    assert ex <= 1e-4
    assert es <= 1e-4


def test_backward_requires_wiener_float_plan():
    announce("backward argument checks")
    from dfttest2_amd import _lib
    from dfttest2_amd.dfttest2 import Backend, get_plan
    from dfttest2_amd.params import dfttest_params

    plan = get_plan(dfttest_params(ftype=1), "float", 32, Backend.MI355X())
    x = torch.zeros((3, 40, 40), device="cuda")
    s = torch.zeros((3, 16, 9), device="cuda")
    job = _lib.Dft2GradJob(src=x.data_ptr(), grad_out=x.data_ptr(), grad_in=x.data_ptr(), src_frame_pitch=6400,
                           src_stride=160, go_frame_pitch=6400, go_stride=160, gi_frame_pitch=6400, gi_stride=160,
                           width=40, height=40, n_frames=3, frame_begin=0, frame_end=3)
    with pytest.raises(_lib.Dft2Error, match="filter_type 0"):
        plan.backward_batch([job], s.data_ptr(), 1.0, s.data_ptr())


def _grads(x0, sig0, p, clean):
    xg = x0.clone().requires_grad_(True)
    sg = sig0.clone().requires_grad_(True)
    loss = torch.mean((dfttest_diff(xg, sg, p) - clean) ** 2)
    loss.backward()
    torch.cuda.synchronize()
    return xg.grad.clone(), sg.grad.clone()


@pytest.mark.parametrize("tbsize,tile_h", [(3, "16"), (3, "4"), (5, "8")])
def test_cross_frame_backward_matches_one_frame_form(tbsize, tile_h, monkeypatch):
    # The cross-frame backward (sums each input frame's spectrum gradients over the output
    # frames that read it, then one adjoint spatial transform) and the one-frame kernel
    # (one adjoint transform per slice and output frame) on the same 2 x 21 frames of
    # 96x120, both against the FP64 autograd graph and against each other.  Both recompute
    # the spectra in FP32, as the forward does, so they take the same side of the Wiener
    # gain's kink (psd = sigma, where d gain / d psd jumps) everywhere; with split-BF16
    # recomputes they differed by 3e-4 where a psd lay within rounding of sigma.  Tile
    # heights 16, 4 and 8 blocks.
    announce("cross-frame vs one-frame backward", tbsize, tile_h)
    p = diff_params(tbsize=tbsize)
    b, n, h, w = 2, 21, 96, 120
    clean = torch.stack([synth.clean_plane(n, h, w, 70 + i) for i in range(b)])
    noisy = clean + torch.stack([synth.noise((n, h, w), 0.05, 80 + i) for i in range(b)])
    rng = np.random.default_rng(11)
    sig0 = torch.from_numpy(np.abs(rng.normal(8.0, 3.0, (p.t_len, 16, 9))).astype(np.float32))
    xr = noisy.double().requires_grad_(True)
    sr = sig0.double().requires_grad_(True)
    wscale = p.extra["wscale"]
    yr = torch.stack([otorch.filter_plane(xr[i], sr * wscale, p) for i in range(b)])
    torch.mean((yr - clean.double()) ** 2).backward()
    dev = torch.device("cuda", 0)
    monkeypatch.setenv("DFT2_BWD_REUSE", "1")
    gx1, gs1 = _grads(noisy.to(dev), sig0.to(dev), p, clean.to(dev))
    monkeypatch.setenv("DFT2_BWD_REUSE", "8")
    monkeypatch.setenv("DFT2_BWD_TILE_H", tile_h)
    gx8, gs8 = _grads(noisy.to(dev), sig0.to(dev), p, clean.to(dev))
    e1 = (rel_err(gx1, xr.grad), rel_err(gs1, sr.grad))
    e8 = (rel_err(gx8, xr.grad), rel_err(gs8, sr.grad))
    print(f"vs FP64: one-frame input {e1[0]:.2e} sigma {e1[1]:.2e}; cross-frame input {e8[0]:.2e} "
          f"sigma {e8[1]:.2e}; between the forms {rel_err(gx8, gx1):.2e}", flush=True)
    assert max(e1) <= 1e-4 and max(e8) <= 1e-4
    assert rel_err(gx8, gx1) <= 1e-4 and rel_err(gs8, gs1) <= 1e-4


@pytest.mark.parametrize("tbsize,reuse", [(3, "8"), (3, "1"), (5, "8")])
def test_backward_frame_range_matches_fp64(tbsize, reuse, monkeypatch):
    # dft2_backward_batch on output frames [4, 15) of a 20-frame clip (the C ABI's frame
    # range): the gradient is that of a loss on those output frames only.  The output
    # gradient comes from the FP64 forward, so the comparison isolates the backward; both
    # backward kernels (cross-frame: 8-frame groups from frame 4, 4 + 8 = 12 and a group of
    # 3; one-frame) against the FP64 autograd graph of the same restricted loss.
    announce("backward frame range", tbsize, reuse)
    from dfttest2_amd.autograd import _grad_job
    from dfttest2_amd.dfttest2 import Backend, get_plan

    p = diff_params(tbsize=tbsize)
    n, h, w, f0, f1 = 20, 44, 52, 4, 15
    clean = synth.clean_plane(n, h, w, 91)
    noisy = clean + synth.noise((n, h, w), 0.05, 92)
    rng = np.random.default_rng(13)
    sig0 = torch.from_numpy(np.abs(rng.normal(8.0, 3.0, (p.t_len, 16, 9))).astype(np.float32))
    wscale = p.extra["wscale"]
    xr = noisy.double().requires_grad_(True)
    sr = sig0.double().requires_grad_(True)
    yr = otorch.filter_plane(xr, sr * wscale, p)
    loss = torch.sum((yr[f0:f1] - clean.double()[f0:f1]) ** 2)
    gy = torch.autograd.grad(loss, yr, retain_graph=True)[0]
    loss.backward()

    dev = torch.device("cuda", 0)
    x = noisy.to(dev)
    gyd = gy.float().to(dev)
    gyd[:f0] = float("nan")  # outside the range: never read
    gyd[f1:] = float("nan")
    gx = torch.zeros_like(x)
    sig = sig0.to(dev)
    gs = torch.zeros_like(sig)
    monkeypatch.setenv("DFT2_BWD_REUSE", reuse)
    plan = get_plan(p, "float", 32, Backend.MI355X())
    job = _grad_job(x, gyd, gx)
    job.frame_begin, job.frame_end = f0, f1
    plan.backward_batch([job], sig.data_ptr(), float(wscale), gs.data_ptr())
    torch.cuda.synchronize()
    ex, es = rel_err(gx, xr.grad), rel_err(gs, sr.grad)
    print(f"vs FP64: input {ex:.2e}, sigma {es:.2e}", flush=True)
    assert torch.isfinite(gx).all() and torch.isfinite(gs).all()
    assert ex <= 1e-4 and es <= 1e-4
