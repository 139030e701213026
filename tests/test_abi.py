"""CPU checks of the C-ABI library: it loads and exports every declared symbol.

No compute call is made here (no GPU in this environment).
"""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "set_intersection.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(si_\w+)\s*\(", text, flags=re.M)))


def test_header_parses():
    names = declared_functions()
    assert "si_open" in names and "si_pair_count" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    from set_intersection_amd import _abi
    assert os.path.exists(_abi.LIB_PATH), "libsi.so not built"
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from set_intersection_amd import _abi
    assert set(declared_functions()) == set(_abi.SIGNATURES)


def test_library_built_for_gfx950():
    from set_intersection_amd import _abi
    out = subprocess.run(["strings", _abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_open_without_gpu_fails_cleanly():
    """si_open must return an error status (not crash) when no GPU is visible."""
    code = ("import set_intersection_amd as si, sys\n"
            "try:\n    si.Context(0)\nexcept si.SIError as e:\n    print('ERR', e.status); sys.exit(0)\n"
            "print('OPENED')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    assert ("ERR" in r.stdout) or ("OPENED" in r.stdout)


def test_missing_library_raises(monkeypatch, tmp_path):
    from set_intersection_amd import _abi
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_abi, "_lib", None)
    with pytest.raises(_abi.LibraryMissing):
        _abi.lib()
