"""Writes tests/golden/*.npz from the reference's own block core (oracle/_ref).

    make -C oracle && python tests/golden/make_golden.py

Each fixture holds plugin-level parameters (window, sigma, sigma2, pmin, pmax,
filter_type, radius, block_step, zero_mean, window_freq), the input plane and the
output of the reference GCC backend's fused() under the restated frame glue
(oracle/ref_harness.cpp), with the centre-slice synthesis window of the GPU backends.
Run only where /root/reference is present; the fixtures themselves are plain data.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "vs-dfttest2_amd"))

from dfttest2_amd.params import dfttest_params  # noqa: E402
from dfttest2_amd import synth  # noqa: E402
from oracle import ref_lib  # noqa: E402

CASES = [
    # name, format, (N, H, W), seed, DFTTest kwargs
    ("u8_defaults", "u8", (5, 48, 64), 11, {}),
    ("u10_ft1_t5_curve", "u10", (7, 40, 48), 12, dict(ftype=1, tbsize=5, slocation=[0.0, 4, 0.5, 8, 1.0, 16])),
    ("f32_defaults", "f32", (5, 36, 40), 13, {}),
    ("u8_t1", "u8", (4, 29, 37), 14, dict(tbsize=1)),
    ("u8_t7_hamming_hanning", "u8", (9, 33, 45), 15, dict(tbsize=7, swin=1, twin=0)),
    ("u16_ft3_step8", "u16", (5, 40, 40), 16, dict(ftype=3, sosize=8, pmin=2.0, pmax=400.0)),
    ("f32_ft4_blackman", "f32", (5, 50, 33), 17, dict(ftype=4, swin=2, pmin=1.0)),
    ("u8_ft6_step2", "u8", (5, 34, 50), 18, dict(f0beta=0.5, sosize=14)),
    ("f32_ft5", "f32", (5, 32, 32), 19, dict(f0beta=0.7, sigma=4.0)),
    ("u8_ft2_nozmean", "u8", (5, 32, 40), 20, dict(ftype=2, sigma=0.5, zmean=False)),
    ("f32_ssystem1_sst", "f32", (5, 36, 36), 21, dict(ssystem=1, sst=[0.0, 2.0, 1.0, 20.0], swin=4)),
    ("u8_sigma_large", "u8", (5, 40, 40), 22, dict(sigma=200.0, tbsize=5, swin=3)),
]


def make_input(fmt, shape, seed):
    n, h, w = shape
    if fmt == "u8":
        return synth.int_plane(n, h, w, seed, bits=8, sd_lsb=6.0).numpy(), 8, False
    if fmt == "u10":
        return synth.int_plane(n, h, w, seed, bits=10, sd_lsb=12.0).numpy(), 10, False
    if fmt == "u16":
        return synth.int_plane(n, h, w, seed, bits=16, sd_lsb=500.0).numpy(), 16, False
    return synth.float_plane(n, h, w, seed, sd=0.05).numpy(), 32, True


def main():
    if not ref_lib.available():
        raise SystemExit("oracle/_ref is not built (make -C oracle)")
    for name, fmt, shape, seed, kw in CASES:
        p = dfttest_params(**kw)
        x, bits, is_float = make_input(fmt, shape, seed)
        y = ref_lib.filter_plane(x, p, bits, is_float)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"),
            input=x, output=y, bits=bits, is_float=int(is_float),
            radius=p.radius, block_step=p.block_step, filter_type=p.filter_type, zero_mean=int(p.zero_mean),
            window=p.window, sigma=p.sigma, sigma2=p.sigma2, pmin=p.pmin, pmax=p.pmax,
            window_freq=p.window_freq if p.window_freq is not None else np.zeros(0),
        )
        print(name, x.shape, x.dtype)

    # An asymmetric sigma array (plugin level only): conjugate partners at kx = 0, 8 differ.
    rng = np.random.Generator(np.random.PCG64(23))
    p = dfttest_params()
    p.sigma = (np.abs(rng.normal(0.5, 2.0, p.t_len * 16 * 9)) * 20.0) * p.extra["wscale"]
    x, bits, is_float = make_input("u8", (5, 40, 48), 23)
    y = ref_lib.filter_plane(x, p, bits, is_float)
    np.savez_compressed(
        os.path.join(HERE, "u8_sigma_asymmetric.npz"),
        input=x, output=y, bits=bits, is_float=0, radius=p.radius, block_step=p.block_step,
        filter_type=p.filter_type, zero_mean=int(p.zero_mean), window=p.window, sigma=p.sigma,
        sigma2=p.sigma2, pmin=p.pmin, pmax=p.pmax, window_freq=p.window_freq,
    )
    print("u8_sigma_asymmetric")


if __name__ == "__main__":
    main()
