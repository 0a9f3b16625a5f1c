"""Steepest-descent receivers (pyfastflow/flow/receivers.py)."""

from __future__ import annotations

import torch

from .. import _lib
from ._core import check, ctx_of, grid_from_constants, n_cells

__all__ = ["compute_receivers"]


def compute_receivers(z, receivers, gradient):
    """D4 steepest descent: receivers[i] = steepest lower neighbour or i,
    gradient[i] = its slope (receivers.py:15-58).  Stochastic receivers
    (constants.RAND_RCV) are not supported and raise."""
    from .. import constants as cte

    if cte.RAND_RCV:
        raise NotImplementedError("stochastic receivers (RAND_RCV) are not implemented")
    n = n_cells()
    tz = check("z", z, n, torch.float32)
    tr = check("receivers", receivers, n, torch.int32)
    tg = check("gradient", gradient, n, torch.float32)
    _lib.call("pff_compute_receivers", ctx_of(tz), grid_from_constants(), _lib.ptr(tz), _lib.ptr(tr), _lib.ptr(tg))
