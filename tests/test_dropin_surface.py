"""The drop-in exposes the reference's Python-callable surface.

tests/golden/reference_api.json lists what pyfastflow.flow and
pyfastflow.graphflood define (tests/golden/gen_reference_api.py reads the
reference's source text; nothing of it is imported): the @ti.kernel
functions and plain functions a caller can reach from Python, the classes and
their public methods, with their parameter names.  Every one of them must
exist under the same name in pyfastflow_amd.flow / pyfastflow_amd.graphflood,
except those listed below with the reason, and take the reference's
parameters by the same names in the same order (it may add optional ones
after them)."""

import inspect
import json
import os

import pytest

import pyfastflow_amd as pf

HERE = os.path.dirname(os.path.abspath(__file__))
API = json.load(open(os.path.join(HERE, "golden", "reference_api.json")))

NOT_EXPOSED = {
    # the three phases of reroute_carve (lakeflow.py:405-438), which is
    # exposed whole: the phases share its tag buffers and change flag
    "init_reroute_carve", "iteration_reroute_carve", "finalise_reroute_carve",
    # the boundary dispatch (neighbourer_flat.py:477-593) that
    # compile_neighbourer turns into Taichi device functions; it runs inside
    # every HIP kernel instead
    "top", "left", "right", "bottom", "validate_link", "neighbour", "can_leave_domain",
}


def _names(pkg):
    v = API[pkg]
    return sorted({e[0] for e in v["kernels"]} | {e[0] for e in v["functions"]} | set(v["classes"]))


def test_inventory_is_complete():
    assert len(_names("flow")) >= 30 and len(_names("graphflood")) >= 3
    assert "FlowRouter" in API["flow"]["classes"] and "Flooder" in API["graphflood"]["classes"]


@pytest.mark.parametrize("pkg", ["flow", "graphflood"])
def test_every_callable_name_exists(pkg):
    mod = getattr(pf, pkg)
    missing = [n for n in _names(pkg) if n not in NOT_EXPOSED and not hasattr(mod, n)]
    assert missing == []


@pytest.mark.parametrize("pkg", ["flow", "graphflood"])
def test_every_public_method_exists(pkg):
    mod = getattr(pf, pkg)
    for cls, methods in API[pkg]["classes"].items():
        ours = getattr(mod, cls)
        assert [m for m in methods if not hasattr(ours, m)] == [], cls


def test_exceptions_are_reference_names():
    every = set(_names("flow")) | set(_names("graphflood"))
    assert NOT_EXPOSED <= every


def _params(fn):
    return list(inspect.signature(fn).parameters)


@pytest.mark.parametrize("pkg", ["flow", "graphflood"])
def test_parameters_match_by_name_and_order(pkg):
    mod = getattr(pf, pkg)
    v = API[pkg]
    for name, where, args in v["kernels"] + v["functions"]:
        if name in NOT_EXPOSED:
            continue
        assert _params(getattr(mod, name))[:len(args)] == args, f"{name} ({where})"
    for cls, methods in v["classes"].items():
        for m, args in methods.items():
            assert _params(getattr(getattr(mod, cls), m))[:len(args)] == args, f"{cls}.{m}"


def test_constants_module_names():
    """pyfastflow.constants' globals and functions (constants.py:20-103)."""
    from pyfastflow_amd import constants

    assert len(API["constants"]) >= 10
    assert [n for n in API["constants"] if not hasattr(constants, n)] == []


TI_DTYPES = {"ti.f32": "torch.float32", "ti.i32": "torch.int32", "ti.i64": "torch.int64",
             "ti.u1": "torch.uint8", "ti.u8": "torch.uint8"}  # ti.u1 fields take one byte per element


def test_router_and_flooder_fields_match_the_reference_layout(monkeypatch):
    """FlowRouter and Flooder built on host storage (the constructors only
    allocate) carry every field the reference's constructors create, with the
    same element type and extent (1-D of N, 4N for the donor lists, 0-d for
    the change flag), so fields are interchangeable with the reference's."""
    for k in ("NX", "NY", "DX", "BOUND_MODE", "PREC", "MANNING", "EDGESW"):  # set by the constructors
        monkeypatch.setattr(pf.constants, k, getattr(pf.constants, k))
    nx, ny = 8, 6
    router = pf.flow.FlowRouter(nx, ny, 1.0, device="cpu")
    flooder = pf.graphflood.Flooder(router)
    for obj, cls in ((router, "FlowRouter"), (flooder, "Flooder")):
        pkg = "flow" if cls == "FlowRouter" else "graphflood"
        for name, (dtype, per_cell) in API[pkg]["fields"][cls].items():
            f = getattr(obj, name)
            assert str(f.dtype) == TI_DTYPES[dtype], f"{cls}.{name}"
            assert f.shape == (() if per_cell == 0 else (per_cell * nx * ny,)), f"{cls}.{name}"
