"""Grid topology (pyfastflow/flow/neighbourer_flat.py).

The reference's neighbour / can_leave_domain / which_edge are Taichi
functions inlined into every kernel; here they live in the HIP kernels
(csrc/pff_common.hpp).  The two debug-map kernels are exposed.
"""

from __future__ import annotations

import torch

from .. import _lib
from ._core import check, ctx_of, grid_from_constants, n_cells

__all__ = ["fill_edges", "flow_out_nodes", "compile_neighbourer"]


def fill_edges(edges):
    """Edge class 0..8 of every cell (neighbourer_flat.py:112-115)."""
    t = check("edges", edges, n_cells(), torch.uint8)
    g = grid_from_constants()
    _lib.call("pff_fill_edges", ctx_of(t), g, _lib.ptr(t))


def flow_out_nodes(outnodes):
    """1 where flow can leave the domain (neighbourer_flat.py:608-611)."""
    t = check("outnodes", outnodes, n_cells(), torch.uint8)
    g = grid_from_constants()
    _lib.call("pff_flow_out_nodes", ctx_of(t), g, _lib.ptr(t))


def compile_neighbourer():
    """Kept for callers of the reference's initialisation
    (neighbourer_flat.py:595-605, environment.py:36), which binds the
    boundary-mode dispatch into Taichi functions once per process.  Here the
    mode is an argument of every kernel call, so there is nothing to bind."""
    return None
