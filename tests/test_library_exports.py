"""The C-ABI library loads without a GPU and exports every entry point that
include/gpuspgemm.h declares; functions that need no device work without one."""
import ctypes
import os
import re
import subprocess

import pytest

import gpuspgemm_amd as g
from gpuspgemm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gpuspgemm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spg_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("spg_multiply", "spg_csr_free", "spg_get_nip", "spg_malloc", "spg_free", "spg_memcpy",
                     "spg_status_str", "spg_last_error", "spg_multiply_batched"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = g.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r"\b(spg_\w+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported
    # the Python binding covers every declared function
    assert set(declared_functions()) == set(_lib._SIG)


def test_status_strings_and_variants_need_no_device():
    lib = g.load_library()
    for code, name in _lib.STATUS.items():
        assert lib.spg_status_str(code)
    assert lib.spg_status_str(99) == b"unknown status"
    assert "capacity=8192" in g.get_variants(1) and "max" in g.get_variants(0)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        _lib.load()


def test_shared_object_targets_gfx950():
    """The library built here carries a gfx950 code object."""
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
