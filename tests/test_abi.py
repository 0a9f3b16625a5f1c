"""The C-ABI library loads and exports every entry point include/pff.h
declares, and the ctypes table matches the header (no GPU calls)."""

import os
import shutil
import subprocess

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


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to build the C caller")
def test_plain_c_program_links_and_calls(tmp_path):
    """A C caller with no Python and no torch: it includes the header, links
    libpff_hip.so, and gets the version string and the argument check's
    error code and message (no GPU call is made)."""
    src = tmp_path / "caller.c"
    src.write_text(
        '#include <stdio.h>\n#include "pff.h"\n'
        "int main(void) {\n"
        '  printf("%s\\n", pff_version());\n'
        "  pff_grid g = {16, 16, 1.0f, 0, 0};\n"
        "  int rc = pff_compute_receivers(0, &g, 0, 0, 0);\n"
        '  printf("%d %s\\n", rc, pff_last_error());\n'
        "  return rc == PFF_EINVAL ? 0 : 1;\n"
        "}\n")
    exe = tmp_path / "caller"
    libdir = os.path.join(ROOT, "pyfastflow_amd")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-l:libpff_hip.so", "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib"],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    assert "gfx950" in lines[0]
    assert lines[1].startswith("-1 ") and "context" in lines[1]
