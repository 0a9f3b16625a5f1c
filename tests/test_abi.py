"""The C-ABI library loads and exports every entry point include/pff.h
declares, and the ctypes table matches the header (no GPU calls)."""

import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    from pyfastflow_amd import _lib

    lib = _lib.load()
    declared = _lib.declared_symbols(os.path.join(ROOT, "include", "pff.h"))
    assert len(declared) >= 25
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []


def test_ctypes_table_matches_header():
    from pyfastflow_amd import _lib

    declared = set(_lib.declared_symbols(os.path.join(ROOT, "include", "pff.h")))
    assert declared == set(_lib.SIGNATURES)


def test_version_string():
    from pyfastflow_amd import _lib

    assert b"gfx950" in _lib.load().pff_version()


def test_library_has_gfx950_code_object():
    path = os.path.join(ROOT, "pyfastflow_amd", "libpff_hip.so")
    data = open(path, "rb").read()
    assert b"gfx950" in data


def test_invalid_grid_is_rejected_without_gpu():
    """Argument checks run before any device work: a NULL context is refused."""
    import ctypes

    from pyfastflow_amd import _lib

    lib = _lib.load()
    g = _lib.grid(1, 5, 1.0, 0)
    rc = lib.pff_compute_receivers(None, ctypes.byref(g), None, None, None)
    assert rc == -1
    assert b"context" in lib.pff_last_error()
