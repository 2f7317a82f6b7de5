"""Pins the FP64 torch restatement used for gradient checks (oracle/dfttest_torch.py):
its forward equals the NumPy oracle, and its autograd gradients equal FP64 central
finite differences along random directions (input planes and sigma)."""

import numpy as np
import pytest
import torch

from dfttest2_amd.params import dfttest_params
from oracle import dfttest_np as onp
from oracle import dfttest_torch as otorch


def _setup(tbsize=3, seed=0, shape=(4, 22, 26), **kw):
    p = dfttest_params(ftype=0, tbsize=tbsize, **kw)
    rng = np.random.default_rng(seed)
    x = rng.random(shape)
    sigma = np.abs(rng.normal(8.0, 3.0, (p.t_len, 16, 9)))
    return p, x, sigma


@pytest.mark.parametrize("tbsize,kw", [(3, {}), (1, {}), (5, dict(swin=1)), (3, dict(sosize=8, zmean=False))])
def test_forward_matches_numpy_oracle(tbsize, kw):
    p, x, sigma = _setup(tbsize, **kw)
    wscale = p.extra["wscale"]
    p.sigma = (sigma * wscale).reshape(-1)
    ref = onp.process_plane(x.astype(np.float32), p, 32, True)
    # the numpy oracle rounds parameters to float32; feed the same sigma values
    s_eff = torch.as_tensor(np.asarray(p.sigma, dtype=np.float32).astype(np.float64).reshape(p.t_len, 16, 9))
    out = otorch.filter_plane(torch.as_tensor(x.astype(np.float32).astype(np.float64)), s_eff, p)
    np.testing.assert_allclose(out.numpy(), ref.astype(np.float64), atol=2e-6)


@pytest.mark.parametrize("tbsize", [1, 3])
def test_autograd_matches_finite_differences(tbsize):
    p, x, sigma = _setup(tbsize, seed=3)
    wscale = p.extra["wscale"]
    rng = np.random.default_rng(11)
    target = torch.as_tensor(rng.random(x.shape))
    xt = torch.as_tensor(x, dtype=torch.float64).requires_grad_(True)
    st = torch.as_tensor(sigma, dtype=torch.float64).requires_grad_(True)

    def loss(xv, sv):
        return torch.mean((otorch.filter_plane(xv, sv * wscale, p) - target) ** 2)

    L = loss(xt, st)
    gx, gs = torch.autograd.grad(L, (xt, st))
    for k in range(3):
        vx = torch.as_tensor(rng.normal(size=x.shape))
        vs = torch.as_tensor(rng.normal(size=sigma.shape))
        h = 1e-5
        with torch.no_grad():
            fd_x = (loss(xt + h * vx, st) - loss(xt - h * vx, st)) / (2 * h)
            fd_s = (loss(xt, st + h * vs) - loss(xt, st - h * vs)) / (2 * h)
        an_x = float((gx * vx).sum())
        an_s = float((gs * vs).sum())
        assert an_x == pytest.approx(float(fd_x), rel=1e-4, abs=1e-10)  # kinks of max(., 0)
        assert an_s == pytest.approx(float(fd_s), rel=1e-4, abs=1e-10)
