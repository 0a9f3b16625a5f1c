"""Shared plumbing of the module-level flow functions: the grid struct built
from the constants globals (as the reference's kernels read cte.NX/NY/DX/
BOUND_MODE, constants.py:27-50) and the per-device library context."""

from __future__ import annotations

from .. import _lib
from .. import constants as cte
from ..field import as_tensor


def grid_from_constants():
    nx, ny, dx, mode, b = cte.grid_args()
    return _lib.grid(nx, ny, dx, mode, b)


def n_cells() -> int:
    return int(cte.NX) * int(cte.NY)


def check(name: str, x, n: int, dtype=None):
    t = as_tensor(x)
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < n:
        raise ValueError(f"{name} has {t.numel()} elements, expected {n}")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")
    return t


def ctx_of(t):
    return _lib.ctx_for(as_tensor(t).device)
