"""Integer facts the kernels' overlap-add gathers rely on (CPU).

The gathers find the blocks covering a pixel column x of a pass as ceil((x - 15) / S) ..
floor(x / S) without a division: x / S is (x * m) >> 16 with m = ceil(65536 / S), computed
once on the host (LaunchArgs::step_magic) and applied with a 24-bit multiply (div_small in
vs-dfttest2_amd/csrc/dft2_device.h).  Exhaustive check over every block step the plugin
accepts (1..16) and every column a pass can hold.
"""

import numpy as np
import pytest


@pytest.mark.parametrize("step", range(1, 17))
def test_multiply_shift_division_is_exact(step):
    m = (65536 + step - 1) // step
    # a pass holds 8 blocks: columns below step * 7 + 16 <= 128; 0..255 is the stated range
    x = np.arange(256, dtype=np.int64)
    assert np.all(x * m < (1 << 24))  # the 24-bit multiply's operand range
    np.testing.assert_array_equal((x * m) >> 16, x // step)


@pytest.mark.parametrize("step", range(1, 17))
def test_covering_block_range(step):
    # lo = ceil((x - 15) / S) clamped at 0, hi = floor(x / S): exactly the blocks k with
    # k*S <= x < k*S + 16, brute-forced over a pass of 8 blocks
    m = (65536 + step - 1) // step
    for x in range(step * 7 + 16):
        lo = (max(x + step - 16, 0) * m) >> 16
        hi = min((x * m) >> 16, 7)
        covering = [k for k in range(8) if k * step <= x < k * step + 16]
        assert covering == list(range(lo, hi + 1)), (step, x)
