"""Depression handling (pyfastflow/flow/lakeflow.py): basin labels, saddles and
outlets, carving or jumping, and the reroute_flow driver."""

from __future__ import annotations

import ctypes
import math

import torch

from .. import _lib
from ..field import as_tensor
from ._core import check, ctx_of, grid_from_constants, n_cells

__all__ = [
    "depression_counter", "basin_id_init", "propagate_basin", "basin_identification", "saddlesort",
    "reroute_jump", "reroute_carve", "count_N_valid", "reroute_flow", "INVALID_KEY",
]

# pack_float_index(1e8, 42) (lakeflow.py:127)
INVALID_KEY = 0x4CBEBC200000002A


def depression_counter(rec) -> int:
    """Number of interior cells draining to themselves (lakeflow.py:24-44)."""
    t = check("rec", rec, n_cells(), torch.int32)
    out = ctypes.c_int64(0)
    _lib.call("pff_depression_counter", ctx_of(t), grid_from_constants(), _lib.ptr(t), ctypes.byref(out))
    return int(out.value)


def basin_id_init(bid):
    """bid = i+1, or 0 where flow can leave (lakeflow.py:46-61)."""
    t = check("bid", bid, n_cells(), torch.int32)
    _lib.call("pff_basin_id_init", ctx_of(t), grid_from_constants(), _lib.ptr(t))


def propagate_basin(bid, rec_):
    """One pointer-jumping pass bid[i] = bid[rec_[i]]; rec_[i] = rec_[rec_[i]]
    (lakeflow.py:63-80), with every read taken before the pass."""
    n = n_cells()
    tb = check("bid", bid, n, torch.int32)
    tr = check("rec_", rec_, n, torch.int32)
    _lib.call("pff_propagate_basin", ctx_of(tb), grid_from_constants(), _lib.ptr(tb), _lib.ptr(tr))


def basin_identification(bid, rec, rec_, N: int | None = None):
    """bid = label of the root each cell drains to (lakeflow.py:82-101)."""
    n = n_cells()
    if N is not None and int(N) != n:
        raise ValueError(f"N={N} does not match NX*NY={n}")
    tb = check("bid", bid, n, torch.int32)
    tr = check("rec", rec, n, torch.int32)
    tr_ = check("rec_", rec_, n, torch.int32)
    _lib.call("pff_basin_identification", ctx_of(tb), grid_from_constants(), _lib.ptr(tb), _lib.ptr(tr),
              _lib.ptr(tr_))


def saddlesort(bid, is_border, z_prime, basin_saddle, basin_saddlenode, outlet, z):
    """Borders, minimum saddle, saddle node, outlet and 2-cycle removal per
    basin (lakeflow.py:105-281)."""
    n = n_cells()
    a = [check("bid", bid, n, torch.int32), check("is_border", is_border, n, torch.uint8),
         check("z_prime", z_prime, n, torch.float32), check("basin_saddle", basin_saddle, n, torch.int64),
         check("basin_saddlenode", basin_saddlenode, n, torch.int32), check("outlet", outlet, n, torch.int64),
         check("z", z, n, torch.float32)]
    _lib.call("pff_saddlesort", ctx_of(a[0]), grid_from_constants(), *[_lib.ptr(x) for x in a])


def reroute_jump(rec, outlet):
    """rec[b-1] = outlet node of basin b (lakeflow.py:284-304)."""
    n = n_cells()
    tr = check("rec", rec, n, torch.int32)
    to = check("outlet", outlet, n, torch.int64)
    _lib.call("pff_reroute_jump", ctx_of(tr), grid_from_constants(), _lib.ptr(tr), _lib.ptr(to))


def reroute_carve(rec, rec_, rec__, tag, tag_, saddlenode, outlet, change=None):
    """Carve from each saddle node to its pit and join it to the outlet
    (lakeflow.py:405-438): rec receives the carved copy of rec_."""
    n = n_cells()
    a = [check("rec", rec, n, torch.int32), check("rec_", rec_, n, torch.int32),
         check("rec__", rec__, n, torch.int32), check("tag", tag, n, torch.uint8),
         check("tag_", tag_, n, torch.uint8), check("saddlenode", saddlenode, n, torch.int32),
         check("outlet", outlet, n, torch.int64)]
    _lib.call("pff_reroute_carve", ctx_of(a[0]), grid_from_constants(), *[_lib.ptr(x) for x in a])
    if change is not None:
        as_tensor(change).fill_(0)


def count_N_valid(arr) -> int:
    """Entries that are not the sentinel key (lakeflow.py:445-463)."""
    t = as_tensor(arr)
    return int((t != INVALID_KEY).sum().item())


def reroute_flow(bid, rec, rec_, rec__, z, z_prime, is_border, outlet, basin_saddle, basin_saddlenode, tag, tag_,
                 change=None, carve=True, bid_first=None, status=None):
    """Lakeflow driver (lakeflow.py:466-544): route every closed depression to
    a basin that drains further, by carving (carve=True, result in rec) or by
    jumping pits to outlets (carve=False, result in rec_ only, as in the
    reference).  Optional: bid_first receives the first pass's basin labels,
    status (a PffLakeStatus) the per-pass depression counts."""
    n = n_cells()
    a = [check("bid", bid, n, torch.int32), check("rec", rec, n, torch.int32), check("rec_", rec_, n, torch.int32),
         check("rec__", rec__, n, torch.int32), check("z", z, n, torch.float32),
         check("z_prime", z_prime, n, torch.float32), check("is_border", is_border, n, torch.uint8),
         check("outlet", outlet, n, torch.int64), check("basin_saddle", basin_saddle, n, torch.int64),
         check("basin_saddlenode", basin_saddlenode, n, torch.int32)]
    tt = check("tag", tag, n, torch.uint8) if tag is not None else None
    tt_ = check("tag_", tag_, n, torch.uint8) if tag_ is not None else None
    bf = check("bid_first", bid_first, n, torch.int32) if bid_first is not None else None
    st = status if status is not None else _lib.PffLakeStatus()
    _lib.call("pff_reroute_flow", ctx_of(a[0]), grid_from_constants(), *[_lib.ptr(x) for x in a], _lib.ptr(tt),
              _lib.ptr(tt_), 1 if carve else 0, _lib.ptr(bf), ctypes.byref(st))
    if change is not None:
        as_tensor(change).fill_(0)
    return st


def outer_iterations(ndep: int) -> int:
    """Cap of the outer loop (lakeflow.py:500)."""
    return 0 if ndep <= 0 else math.ceil(math.log2(ndep)) + 1
