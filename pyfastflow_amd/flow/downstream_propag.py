"""Donors and rake-and-compress accumulation (pyfastflow/flow/downstream_propag.py)."""

from __future__ import annotations

import torch

from .. import _lib
from ._core import check, ctx_of, grid_from_constants, n_cells

__all__ = ["rcv2donor", "rake_compress_accum", "fuse", "upstream_count"]


def rcv2donor(rcv, dnr, ndnr):
    """Invert receivers into 4 donor slots per node (downstream_propag.py:77-96).
    ndnr must be zero on entry (flowfields.py:275)."""
    n = n_cells()
    tr = check("rcv", rcv, n, torch.int32)
    td = check("dnr", dnr, 4 * n, torch.int32)
    tn = check("ndnr", ndnr, n, torch.int32)
    _lib.call("pff_rcv2donor", ctx_of(tr), grid_from_constants(), _lib.ptr(tr), _lib.ptr(td), _lib.ptr(tn))


def rake_compress_accum(dnr, ndnr, p, src, dnr_, ndnr_, p_, iteration: int):
    """One rake-and-compress round (downstream_propag.py:99-179)."""
    n = n_cells()
    args = [check("dnr", dnr, 4 * n, torch.int32), check("ndnr", ndnr, n, torch.int32),
            check("p", p, n, torch.float32), check("src", src, n, torch.int32),
            check("dnr_", dnr_, 4 * n, torch.int32), check("ndnr_", ndnr_, n, torch.int32),
            check("p_", p_, n, torch.float32)]
    _lib.call("pff_rake_compress_accum", ctx_of(args[0]), grid_from_constants(), *[_lib.ptr(a) for a in args],
              int(iteration))


def fuse(A, src, B, iteration: int):
    """A[i] = B[i] where the ping-pong state selects B (downstream_propag.py:59-74)."""
    n = n_cells()
    ta = check("A", A, n, torch.float32)
    ts = check("src", src, n, torch.int32)
    tb = check("B", B, n, torch.float32)
    _lib.call("pff_fuse", ctx_of(ta), grid_from_constants(), _lib.ptr(ta), _lib.ptr(ts), _lib.ptr(tb), int(iteration))


def upstream_count(rcv, count):
    """count[i] = number of cells whose receiver chain passes through i,
    itself included (uint32 stored in an int32 field).  This is the integer
    core of FlowRouter.accumulate_constant_Q (flowfields.py:254-292): the
    converged discharge is count * Q0."""
    n = n_cells()
    tr = check("rcv", rcv, n, torch.int32)
    tc = check("count", count, n, torch.int32)
    _lib.call("pff_upstream_count", ctx_of(tr), grid_from_constants(), _lib.ptr(tr), _lib.ptr(tc))
