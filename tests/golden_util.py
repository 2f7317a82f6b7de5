"""Loads the golden fixtures of tests/golden/ as (FilterParams, input, output, bits, is_float)."""

import glob
import os

import numpy as np

from dfttest2_amd.params import FilterParams

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz"))
                  if not os.path.basename(p).startswith("windows"))


def load_fixture(name):
    z = np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False)
    wf = z["window_freq"]
    p = FilterParams(
        radius=int(z["radius"]), block_size=16, block_step=int(z["block_step"]),
        filter_type=int(z["filter_type"]), zero_mean=bool(int(z["zero_mean"])), window=z["window"],
        sigma=z["sigma"], sigma2=float(z["sigma2"]), pmin=float(z["pmin"]), pmax=float(z["pmax"]),
        window_freq=wf if wf.size else None,
    )
    return p, z["input"], z["output"], int(z["bits"]), bool(int(z["is_float"]))


def max_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) if a.size else 0.0
