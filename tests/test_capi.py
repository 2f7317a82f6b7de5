"""C ABI checks that need no GPU: the library loads, exports every declared entry point,
validates arguments before touching a device, and its RDFT matches numpy."""

import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from dfttest2_amd import _lib
from dfttest2_amd.params import dfttest_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dfttest2_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*DFT2_API\s+(?:int|const char\*)\s+(dft2_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)
    for name in names:
        assert hasattr(lib, name), name


def test_library_exports_only_the_c_abi():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.split() and line.split()[-2] in "TW"}
    ours = {n for n in names if "dft2" in n}
    assert ours == set(declared_functions())


def test_version_string():
    assert "dfttest2_mi355x" in _lib.version()


@pytest.mark.parametrize("shape", [(16, 16), (3, 16, 16), (7, 16, 16), (5,), (2, 3, 4)])
def test_capi_rdft_matches_numpy(shape):
    rng = np.random.default_rng(2)
    data = rng.normal(size=shape)
    spec = np.fft.rfftn(data)
    out = _lib.rdft(data, shape).reshape(spec.shape + (2,))
    np.testing.assert_allclose(out[..., 0], spec.real, atol=1e-9)
    np.testing.assert_allclose(out[..., 1], spec.imag, atol=1e-9)


def test_capi_rdft_rejects_bad_shape():
    with pytest.raises(_lib.Dft2Error, match="reshape"):
        _lib.rdft(np.zeros(10), (3, 4))


def _params(**over):
    p = dfttest_params()
    for k, v in over.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize(
    "over,msg",
    [
        (dict(radius=8), "radius"),
        (dict(block_size=65), "block_size"),
        (dict(block_size=8), "window must have 192"),  # a 16-point window for 8-point blocks
        (dict(block_size=8, block_step=9), "block_step"),
        (dict(block_step=0), "block_step"),
        (dict(filter_type=7), "filter_type"),
        (dict(window=np.ones(100)), "window"),
        (dict(sigma=np.ones(5)), "sigma"),
        (dict(window_freq=None), "window_freq"),
    ],
)
def test_plan_create_validates_arguments(over, msg):
    with pytest.raises(_lib.Dft2Error, match=msg):
        _lib.Plan(_params(**over), _lib.DFT2_SAMPLE_INT, 8)


def test_generic_path_limits():
    # Block sizes other than 16 (and radius > 3) take the generic block path, whose LDS per
    # block bounds the size: 64-point blocks at tbsize 15 do not fit.
    p = dfttest_params(sbsize=64, sosize=48, tbsize=15)
    with pytest.raises(_lib.Dft2Error, match="LDS"):
        _lib.Plan(p, _lib.DFT2_SAMPLE_INT, 8)
    with pytest.raises(_lib.Dft2Error, match="split BF16"):
        _lib.Plan(dfttest_params(sbsize=8, sosize=4), _lib.DFT2_SAMPLE_INT, 8, precision=_lib.DFT2_PRECISION_SPLIT_BF16)


def test_plan_create_rejects_formats():
    with pytest.raises(_lib.Dft2Error, match="integer"):
        _lib.Plan(_params(), _lib.DFT2_SAMPLE_INT, 17)
    with pytest.raises(_lib.Dft2Error, match="float"):
        _lib.Plan(_params(), _lib.DFT2_SAMPLE_FLOAT, 16)
    with pytest.raises(_lib.Dft2Error, match="num_streams"):
        _lib.Plan(_params(), _lib.DFT2_SAMPLE_INT, 8, num_streams=0)


def test_stream_arguments():
    # None selects the plan's own stream (a blocking call); a torch stream handle is passed
    # through, with 0 (torch's default stream) as DFT2_DEFAULT_STREAM, which the library
    # turns into HIP's stream 0 without waiting.
    assert _lib._stream_arg(None).value is None
    assert _lib._stream_arg(0).value == _lib.DFT2_DEFAULT_STREAM == 1
    text = open(HEADER).read()
    assert re.search(r"#define DFT2_DEFAULT_STREAM \(\(void\*\)1\)", text)
    assert _lib._stream_arg(0x7F00_1234).value == 0x7F00_1234
