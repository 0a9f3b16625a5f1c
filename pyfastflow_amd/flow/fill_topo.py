"""Epsilon-gradient filling (pyfastflow/flow/fill_topo.py)."""

from __future__ import annotations

import torch

from .. import _lib
from ._core import check, ctx_of, grid_from_constants, n_cells

__all__ = ["topofill", "fill_z_add_delta"]


def topofill(flow_field, epsilon=1e-6, custom_z=None):
    """Fill the router's z so every cell has a strictly descending path to its
    root (fill_topo.py:48-78).  With custom_z the filled surface goes there
    and router.z is left unchanged.  The reference prints each increment
    (fill_topo.py:71); that debug print is dropped."""
    n = flow_field.nx * flow_field.ny
    z = check("z", flow_field.z, n, torch.float32)
    out = check("custom_z", custom_z, n, torch.float32) if custom_z is not None else z
    r = check("receivers", flow_field.receivers, n, torch.int32)
    r_ = check("receivers_", flow_field.receivers_, n, torch.int32)
    r__ = check("receivers__", flow_field.receivers__, n, torch.int32)
    _lib.call("pff_topofill", ctx_of(z), flow_field._grid(), _lib.ptr(z), _lib.ptr(out), _lib.ptr(r), _lib.ptr(r_),
              _lib.ptr(r__), float(epsilon))


def fill_z_add_delta(zh, h, z_, receivers, receivers_, receivers__, epsilon=1e-4):
    """Fill zh (first increment epsilon/2, fill_topo.py:121-122) and add the
    raise to h: h += filled - zh; zh = filled (fill_topo.py:99-125)."""
    n = n_cells()
    a = [check("zh", zh, n, torch.float32), check("h", h, n, torch.float32), check("z_", z_, n, torch.float32),
         check("receivers", receivers, n, torch.int32), check("receivers_", receivers_, n, torch.int32),
         check("receivers__", receivers__, n, torch.int32)]
    _lib.call("pff_fill_z_add_delta", ctx_of(a[0]), grid_from_constants(), *[_lib.ptr(x) for x in a], float(epsilon))
