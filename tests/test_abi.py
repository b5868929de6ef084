"""The C-ABI library loads and exports every entry point include/spmd.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from starpolymersmd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "spmd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(spmd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "spmd_mpc_step" in syms and "spmd_md_propagate" in syms and "spmd_create" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    assert set(declared_symbols()) <= set(_lib._SIGS), set(declared_symbols()) - set(_lib._SIGS)


def test_error_strings_without_device():
    lib = _lib.load()
    assert lib.spmd_error_string(0) == b"ok"
    assert b"cell index" in lib.spmd_error_string(4)
    assert b"gfx950" in lib.spmd_version()


def test_library_has_gfx950_code_object():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.SpmdError):
        _lib.load()
