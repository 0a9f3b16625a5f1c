"""The C-ABI library loads and exports every entry point include/pff.h
declares, and the ctypes table matches the header (no GPU calls)."""

import os
import shutil
import subprocess
import sys

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


_NULL_CTX_SCRIPT = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
from pyfastflow_amd import _lib
lib = _lib.load()
for name, args in _lib.SIGNATURES.items():
    if name in ("pff_version", "pff_last_error"):
        continue
    vals = []
    for a in args:
        if a in (ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64):
            vals.append(a(1))
        elif a in (ctypes.c_float, ctypes.c_double):
            vals.append(a(1.0))
        else:
            vals.append(None)
    rc = getattr(lib, name)(*vals)
    print(name, rc, lib.pff_last_error().decode(), flush=True)
"""


def test_every_entry_point_refuses_null_arguments_before_device_work():
    """Each entry point, called with a NULL context (or, for the two without
    one, NULL out/grid pointers) and NULL arrays, returns PFF_EINVAL with a
    message and touches neither the arrays nor the GPU; destroying a NULL
    context is a no-op.  Run in a child process so that a crash would fail
    the test instead of the test run."""
    out = subprocess.run([sys.executable, "-c", _NULL_CTX_SCRIPT, ROOT], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    got = {}
    for line in out.stdout.splitlines():
        name, rc, msg = line.split(" ", 2)
        got[name] = (int(rc), msg)
    from pyfastflow_amd import _lib

    assert set(got) == set(_lib.SIGNATURES) - {"pff_version", "pff_last_error"}
    assert got.pop("pff_ctx_destroy")[0] == 0  # a NULL context: nothing to release
    assert got.pop("pff_ctx_create") == (-1, "out is NULL")
    assert got.pop("pff_accumulate_constant_Q_noalloc") == (-1, "grid is NULL")
    for name, (rc, msg) in got.items():
        assert (rc, msg) == (-1, "context is NULL"), name


def _noalloc(g, rcv=1, Q=1, conv=1):
    import ctypes

    from pyfastflow_amd import _lib

    lib = _lib.load()
    p = lambda v: ctypes.c_void_p(v * 0x1000 if v else 0)  # never dereferenced: the checks fail first
    rc = lib.pff_accumulate_constant_Q_noalloc(None, ctypes.byref(g), p(rcv), 1.0, 1, p(Q), p(conv))
    return rc, lib.pff_last_error().decode()


@pytest.mark.parametrize("nx,ny,dx,mode,msg", [
    (1, 5, 1.0, 0, "nx >= 2"),
    (5, 1, 1.0, 0, "ny >= 2"),
    (65536, 32768, 1.0, 0, "2^31"),
    (8, 8, 1.0, 4, "bound_mode"),
    (8, 8, 1.0, -1, "bound_mode"),
    (8, 8, 1.0, 3, "boundaries"),
    (8, 8, 0.0, 0, "dx"),
    (8, 8, -1.0, 0, "dx"),
    (8, 8, float("nan"), 0, "dx"),
])
def test_grid_contract_is_checked(nx, ny, dx, mode, msg):
    """The header's grid contract (nx, ny >= 2, nx*ny < 2^31, mode 0..3, a
    boundaries array in custom mode, dx > 0) is checked before any device
    work; the entry point without a context shows it on the CPU."""
    from pyfastflow_amd import _lib

    rc, err = _noalloc(_lib.grid(nx, ny, dx, mode))
    assert rc == -1 and msg in err, err


def test_noalloc_checks_pointers_and_its_size_limit_before_device_work():
    from pyfastflow_amd import _lib

    g = _lib.grid(64, 64, 1.0, 0)
    assert _noalloc(g, rcv=0) == (-1, "rcv is NULL")
    assert _noalloc(g, Q=0) == (-1, "Q is NULL")
    assert _noalloc(g, conv=0) == (-1, "converged is NULL")
    rc, err = _noalloc(_lib.grid(2048, 2048, 1.0, 0))
    assert rc == -1 and "fixed scratch" in err, err
