"""Steepest-descent receivers (pyfastflow/flow/receivers.py)."""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ._core import check, ctx_of, grid_from_constants, n_cells

__all__ = ["compute_receivers"]


def compute_receivers(z, receivers, gradient, seed=None):
    """D4 steepest descent: receivers[i] = steepest lower neighbour or i,
    gradient[i] = its slope (receivers.py:15-58).

    With constants.RAND_RCV (receivers.py:46-47) each valid link's slope is
    weighted by a uniform draw in [0, 1) and gradient holds the weighted
    slope.  The draws come from a seeded counter-based generator
    (splitmix64 of seed + (4i + k) * golden ratio), `seed` or else
    constants.next_rand_seed(); Taichi's ti.random() is another generator,
    so a run reproduces the reference's distribution, not its draws."""
    from .. import constants as cte

    n = n_cells()
    tz = check("z", z, n, torch.float32)
    tr = check("receivers", receivers, n, torch.int32)
    tg = check("gradient", gradient, n, torch.float32)
    if cte.RAND_RCV:
        s = cte.next_rand_seed() if seed is None else int(seed)
        _lib.call("pff_compute_receivers_stochastic", ctx_of(tz), grid_from_constants(), _lib.ptr(tz), _lib.ptr(tr),
                  _lib.ptr(tg), ctypes.c_uint64(s & 0xFFFFFFFFFFFFFFFF))
        return
    _lib.call("pff_compute_receivers", ctx_of(tz), grid_from_constants(), _lib.ptr(tz), _lib.ptr(tr), _lib.ptr(tg))
