"""The C ABI from a compiled C client (tests/abi_client.c): the public header compiles as
strict C99, the client links against the shared library alone, its GPU-free entry points
behave (dft2_rdft against a direct FP64 DFT, argument errors), and on the GPU its
dft2_process_frame and dft2_process_frame_keyed calls on host planes give the bits that
the Python binding's calls give, within 1 LSB of the FP64 oracle."""

import os
import subprocess
import sys

import numpy as np
import pytest

from dfttest2_amd import _lib, synth
from dfttest2_amd.params import dfttest_params
from oracle import dfttest_np as onp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "_bin", "abi_client")


def client(tmp_path):
    """The client built by __graft_entry__.build(); compiled into tmp_path where it is
    missing (CPU tests only: GPU runs use the prebuilt one)."""
    if os.path.exists(CLIENT):
        return CLIENT
    sys.path.insert(0, ROOT)
    import __graft_entry__ as entry

    return entry.build_abi_client(str(tmp_path / "abi_client"))


def test_header_is_strict_c99(tmp_path):
    src = tmp_path / "h.c"
    src.write_text('#include "dfttest2_mi355x.h"\nint main(void) { return dft2_version() == 0; }\n')
    out = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fsyntax-only",
                          "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr


def test_client_cpu_entry_points(tmp_path):
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("native library not built")
    out = subprocess.run([client(tmp_path), "cpu"], capture_output=True, text=True, timeout=60)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert "cpu checks passed" in out.stdout


def write_params(path, p, sample_type, bits, width, height, n_frames):
    window = np.ascontiguousarray(p.window, dtype=np.float64)
    sigma = np.ascontiguousarray(np.atleast_1d(p.sigma), dtype=np.float64)
    wf = np.zeros(0) if p.window_freq is None else np.ascontiguousarray(p.window_freq, dtype=np.float64)
    hdr = np.array([p.radius, p.block_size, p.block_step, p.filter_type, int(p.zero_mean), window.size, sigma.size,
                    wf.size, sample_type, bits, width, height, n_frames], dtype=np.int32)
    scal = np.array([p.sigma2, p.pmin, p.pmax], dtype=np.float64)
    with open(path, "wb") as f:
        for a in (hdr, scal, window, sigma, wf):
            f.write(a.tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,tbsize", [("u8", 3), ("u10", 5), ("f32", 3)])
def test_client_frames_match_python_binding(fmt, tbsize, tmp_path):
    print(f"\n[case] C client frames {fmt} tbsize {tbsize}", flush=True)
    torch = pytest.importorskip("torch")
    p = dfttest_params(tbsize=tbsize)
    n, h, w = 6, 44, 60
    if fmt == "f32":
        x = synth.float_plane(n, h, w, 71, 0.05).numpy()
        st, bits = _lib.DFT2_SAMPLE_FLOAT, 32
    else:
        bits = 8 if fmt == "u8" else 10
        x = synth.int_plane(n, h, w, 71, bits, 6.0).numpy()
        st = _lib.DFT2_SAMPLE_INT
    write_params(tmp_path / "params.bin", p, st, bits, w, h, n)
    (tmp_path / "frames.bin").write_bytes(np.ascontiguousarray(x).tobytes())
    out = subprocess.run([CLIENT, "frames", str(tmp_path / "params.bin"), str(tmp_path / "frames.bin"),
                          str(tmp_path / "plain.bin"), str(tmp_path / "keyed.bin")],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    plain = np.fromfile(tmp_path / "plain.bin", dtype=x.dtype).reshape(x.shape)
    keyed = np.fromfile(tmp_path / "keyed.bin", dtype=x.dtype).reshape(x.shape)
    # the same calls through the Python binding
    plan = _lib.Plan(p, st, bits)
    ref_py = np.empty_like(x)
    bps = x.dtype.itemsize
    for i in range(n):
        srcs = [np.ascontiguousarray(x[min(max(i - p.radius + t, 0), n - 1)]) for t in range(p.t_len)]
        o = np.empty((h, w), x.dtype)
        plan.process_frame([s.ctypes.data for s in srcs], w * bps, o.ctypes.data, w * bps, w, h, memory=1)
        ref_py[i] = o
    plan.close()
    torch.cuda.synchronize()
    assert np.array_equal(plain, ref_py)
    assert np.array_equal(keyed, ref_py)
    ref = onp.process_plane(x, p, bits, fmt == "f32")
    err = float(np.max(np.abs(plain.astype(np.float64) - ref.astype(np.float64))))
    assert err <= (1e-5 if fmt == "f32" else 1.0)
