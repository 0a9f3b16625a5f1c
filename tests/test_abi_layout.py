"""The ctypes binding agrees with include/pff.h, read as a C compiler reads it
(no GPU calls).

- every struct the header defines has a ctypes mirror with the same fields in
  the same order, and gcc, compiling the header itself, gives the same size
  and the same offset for every field;
- every entry point's ctypes argument list has the header's parameter count,
  and each argument has the header's kind (pointer, int32, int64, uint64,
  float, double), pointers to the header's structs pointing at their mirrors.
A mismatch here would pass the wrong bytes at the boundary with no error.
"""

import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pff.h")

# C typedef -> ctypes mirror in pyfastflow_amd._lib
MIRRORS = {
    "pff_grid": "PffGrid",
    "pff_hydro": "PffHydro",
    "pff_lake_status": "PffLakeStatus",
    "pff_gf_status": "PffGfStatus",
    "pff_conv_src": "PffConvSrc",
    "pff_tile": "PffTile",
    "pff_batch_status": "PffBatchStatus",
}


def _header_text() -> str:
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return re.sub(r"//[^\n]*", " ", txt)


def _header_structs() -> dict[str, list[str]]:
    """typedef name -> field names in declaration order."""
    out = {}
    for body, name in re.findall(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", _header_text(), flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            rest = re.sub(r"^(?:const\s+)?\w+\s*\**\s*", "", decl)
            for d in rest.split(","):
                d = re.sub(r"\[.*?\]", "", d).strip().lstrip("*").strip()
                fields.append(d)
        out[name] = fields
    return out


def _header_functions() -> dict[str, list[str]]:
    """entry point -> parameter declarations."""
    out = {}
    for name, params in re.findall(r"(?:int|const\s+char\s*\*)\s+(pff_\w+)\s*\(([^)]*)\)\s*;", _header_text()):
        ps = [p.strip() for p in params.split(",")]
        out[name] = [] if ps == ["void"] else [re.sub(r"\s+", " ", p) for p in ps]
    return out


def test_every_header_struct_has_a_mirror_with_the_same_fields():
    from pyfastflow_amd import _lib

    structs = _header_structs()
    assert set(structs) == set(MIRRORS), "a header struct without a ctypes mirror (or a stale mirror)"
    for cname, pyname in MIRRORS.items():
        mirror = getattr(_lib, pyname)
        assert [f[0] for f in mirror._fields_] == structs[cname], cname


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to read the header as C")
def test_struct_sizes_and_offsets_match_the_compiled_header(tmp_path):
    from pyfastflow_amd import _lib

    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pff.h"', "int main(void) {"]
    expect = []
    for cname, pyname in MIRRORS.items():
        mirror = getattr(_lib, pyname)
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        expect.append(f"{cname} sizeof {ctypes.sizeof(mirror)}")
        for fname, _ in mirror._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
            expect.append(f"{cname}.{fname} {getattr(mirror, fname).offset}")
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True, capture_output=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert [g for g in got if g] == expect


def _kind_of_c(param: str) -> tuple[str, str | None]:
    """('ptr', pointee struct or None) or (scalar kind, None)."""
    if "*" in param:
        base = param.replace("const", "").split("*")[0].strip()
        stars = param.count("*")
        if base in MIRRORS and stars == 1:
            return "ptr", base
        if base == "char" and stars == 1:
            return "str", None
        return "ptr", None
    base = param.replace("const", "").split()[0]
    return {"int32_t": "i32", "int64_t": "i64", "uint64_t": "u64", "float": "f32", "double": "f64"}[base], None


def _kind_of_ctypes(t) -> tuple[str, object]:
    scalars = {ctypes.c_int32: "i32", ctypes.c_int64: "i64", ctypes.c_uint64: "u64", ctypes.c_float: "f32",
               ctypes.c_double: "f64"}
    if t in scalars:
        return scalars[t], None
    if t is ctypes.c_char_p:
        return "str", None
    if t is ctypes.c_void_p:
        return "ptr", None
    if issubclass(t, ctypes._Pointer):
        return "ptr", t._type_
    raise AssertionError(f"unexpected ctypes argument type {t}")


def test_every_entry_point_binding_matches_the_header_parameters():
    from pyfastflow_amd import _lib

    funcs = _header_functions()
    assert set(funcs) == set(_lib.SIGNATURES)
    for name, params in funcs.items():
        args = _lib.SIGNATURES[name]
        assert len(args) == len(params), f"{name}: {len(args)} ctypes arguments, {len(params)} in the header"
        for k, (p, a) in enumerate(zip(params, args)):
            ck, cstruct = _kind_of_c(p)
            pk, target = _kind_of_ctypes(a)
            if ck == "str":  # a C string may also be bound as an untyped pointer
                assert pk in ("str", "ptr"), f"{name} argument {k} ({p})"
                continue
            assert ck == pk, f"{name} argument {k} ({p}): ctypes {a}"
            if cstruct is not None:
                assert target is getattr(_lib, MIRRORS[cstruct]), f"{name} argument {k} ({p}): points at {target}"


def test_return_types():
    from pyfastflow_amd import _lib

    lib = _lib.load()
    text = _header_text()
    for name in _lib.SIGNATURES:
        returns_string = re.search(r"const\s+char\s*\*\s*" + name + r"\s*\(", text) is not None
        restype = getattr(lib, name).restype
        assert (restype is ctypes.c_char_p) == returns_string, name
        if not returns_string:
            assert restype is ctypes.c_int32, name
