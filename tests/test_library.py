"""The C-ABI library loads on a host without a GPU and exports every entry
point declared in include/gpudelaunay.h (no compute calls here)."""
import ctypes
import os
import re

import pytest

from gpudelaunaytriangulation_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gpudelaunay.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gdt_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libgpudelaunay.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_status_strings():
    L = _lib.load()
    assert _lib.status_string(0) == "ok"
    assert "stride" in _lib.status_string(_lib.GDT_OVERFLOW)
    assert L.gdt_status_string(12345) == b"unknown status"


def test_cpp_dropin_header_compiles(tmp_path):
    """include/gpudelaunay.hpp (the C++ classes over the C ABI) compiles and
    links against the library with a plain host compiler."""
    import subprocess
    from gpudelaunaytriangulation_amd import build as B
    for src in (B.DROPIN_SRC, B.GPUARRAY_SRC):
        out = str(tmp_path / os.path.basename(src)[:-4])
        r = subprocess.run(B.cxx_dropin_cmd(src, out), capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
