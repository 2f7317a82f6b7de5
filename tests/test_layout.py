"""The product never reaches the oracle: no module of the dfttest2_amd package imports from
oracle/ or names the reference build under oracle/_ref, no native source includes a file
from there, and the native library links nothing from it.  The oracle is test and bench
infrastructure only (tests/, the bench's CPU leg, the smoke check)."""

import ast
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "vs-dfttest2_amd" / "dfttest2_amd"
CSRC = ROOT / "vs-dfttest2_amd" / "csrc"


def test_package_imports_nothing_from_the_oracle():
    files = list(PKG.rglob("*.py"))
    assert files
    for f in files:
        text = f.read_text()
        for node in ast.walk(ast.parse(text)):
            if isinstance(node, ast.Import):
                names = [a.name for a in node.names]
            elif isinstance(node, ast.ImportFrom):
                names = [node.module or ""]
            else:
                continue
            assert not any(n.split(".")[0] == "oracle" for n in names), f"{f} imports the oracle"
        assert "libdft2ref" not in text and "oracle/_ref" not in text, f"{f} names the reference build"


def test_native_sources_include_nothing_from_the_oracle():
    for f in list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")) + list(CSRC.glob("*.h")):
        for line in f.read_text().splitlines():
            if line.lstrip().startswith("#include"):
                assert "oracle" not in line, f"{f}: {line}"


def test_native_library_links_nothing_from_the_oracle():
    so = PKG / "_native" / "libdfttest2_mi355x.so"
    if not so.exists():
        pytest.skip("native library not built")
    needed = subprocess.run(["readelf", "-d", str(so)], capture_output=True, text=True, check=True).stdout
    assert "dft2ref" not in needed and "oracle" not in needed
