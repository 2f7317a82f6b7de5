"""The VapourSynth plugin source (vs-dfttest2_amd/vsplugin) on the CPU: its registered
functions and argument strings are the reference HIPRTC plugin's
(hiprtc_source/source.cpp:1025-1057) plus the optional precision, and the front-end diff
for the reference's Python files is well formed.  The plugin itself is built only where
VapourSynth headers exist (none in this image)."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "vs-dfttest2_amd", "vsplugin", "dfttest2_mi355x_vs.cpp")
DIFF = os.path.join(ROOT, "vs-dfttest2_amd", "vsplugin", "reference_frontend.diff")

REFERENCE_DFTTEST_ARGS = ("clip:vnode;window:float[];sigma:float[];sigma2:float;pmin:float;pmax:float;"
                          "filter_type:int;radius:int:opt;block_size:int:opt;block_step:int:opt;zero_mean:int:opt;"
                          "window_freq:float[]:opt;planes:int[]:opt;device_id:int:opt;num_streams:int:opt;")


def registrations():
    text = open(SRC).read()
    out = {}
    for m in re.finditer(r'registerFunction\(\s*"(\w+)",((?:\s*"[^"]*")+),((?:\s*"[^"]*")+),', text):
        join = lambda s: "".join(re.findall(r'"([^"]*)"', s))  # noqa: E731
        out[m.group(1)] = (join(m.group(2)), join(m.group(3)))
    return out


def test_plugin_registers_the_reference_functions():
    reg = registrations()
    assert set(reg) == {"DFTTest", "RDFT", "Version"}
    assert reg["DFTTest"] == (REFERENCE_DFTTEST_ARGS + "precision:int:opt;", "clip:vnode;")
    assert reg["RDFT"] == ("data:float[];shape:int[];", "window_freq:float[];")
    assert reg["Version"] == ("", "any")
    text = open(SRC).read()
    assert '"dfttest2_mi355x"' in text  # the namespace the front-end diff selects
    for call in ("dft2_plan_create", "dft2_process_frame_keyed", "dft2_plan_destroy", "dft2_rdft", "dft2_version"):
        assert call in text


def test_frontend_diff_adds_the_backend():
    diff = open(DIFF).read()
    assert "+++ b/python/backends.py" in diff and "+++ b/python/dfttest2.py" in diff
    added = "\n".join(l[1:] for l in diff.splitlines() if l.startswith("+") and not l.startswith("+++"))
    assert "class MI355X:" in added and '("dfttest2_mi355x", Backend.MI355X)' in added
    assert "plugin = core.dfttest2_mi355x" in added and '"precision": backend_inst.precision' in added


@pytest.mark.skipif(not os.path.isdir("/root/reference/python") or shutil.which("patch") is None,
                    reason="the reference's Python files are not here")
def test_frontend_diff_applies_to_the_reference(tmp_path):
    os.makedirs(tmp_path / "python")
    for f in ("backends.py", "dfttest2.py"):
        shutil.copy(os.path.join("/root/reference/python", f), tmp_path / "python" / f)
    r = subprocess.run(["patch", "-p1", "-i", DIFF], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "class MI355X" in (tmp_path / "python" / "backends.py").read_text()


def test_getframe_holds_the_largest_radius():
    # GetFrame gathers T = 2 * radius + 1 source frames into stack arrays; dft2_plan_create
    # accepts radius up to the generic path's maximum, so the arrays cover that T.
    text = open(SRC).read()
    internal = open(os.path.join(ROOT, "vs-dfttest2_amd", "csrc", "dft2_internal.h")).read()
    host = open(os.path.join(ROOT, "vs-dfttest2_amd", "csrc", "dft2_host.h")).read()
    m = re.search(r"kMaxGenericRadius\s*=\s*(\d+)", internal + host)
    assert m
    max_t = 2 * int(m.group(1)) + 1
    slices = int(re.search(r"constexpr int kMaxSlices\s*=\s*(\d+)", text).group(1))
    assert slices >= max_t
    body = text[text.index("VS_CC GetFrame"):text.index("VS_CC Free")]
    for decl in ("src[kMaxSlices]", "src_n[kMaxSlices]", "planes[kMaxSlices]", "keys[kMaxSlices]"):
        assert decl in body, decl
    assert not re.search(r"\w+\[\d+\]\s*=\s*\{\}", body)  # no other fixed-size per-slice arrays
