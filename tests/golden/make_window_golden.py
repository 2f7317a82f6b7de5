"""Writes tests/golden/windows.npz from the reference's own python/helpers.py.

    python tests/golden/make_window_golden.py

Only for the window shapes that no closed form or library routine pins here: the
Kaiser-Bessel window with the reference's truncated I0 series (mode 4), 7-term
Blackman-Harris (5), flat top (6), Bartlett-Hann without abs() (9) and Nuttall (10).
The reference module is loaded by file path in this container only; the fixture keeps
its outputs as plain numbers.
"""

import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_HELPERS = "/root/reference/python/helpers.py"
MODES = (4, 5, 6, 9, 10)
SIZES = (1, 3, 5, 7, 16)
BETAS = (2.5, 1.0, 6.0)


def main():
    spec = importlib.util.spec_from_file_location("ref_helpers", REF_HELPERS)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rows = []
    for mode in MODES:
        for size in SIZES:
            for beta in BETAS if mode == 4 else (2.5,):
                for i in range(size):
                    loc = i + 0.5
                    rows.append((mode, size, beta, loc, mod.get_window_value(loc, size, mode, beta)))
    arr = np.array(rows, dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "windows.npz"), rows=arr)
    print(arr.shape)


if __name__ == "__main__":
    main()
