"""GraphFlood kernels (pyfastflow/graphflood/gf_hydrodynamics.py)."""

from __future__ import annotations

import torch

from .. import _lib
from .. import constants as cte
from ..flow._core import check, ctx_of, grid_from_constants, n_cells

__all__ = ["diffuse_Q_constant_prec", "graphflood_core_cte_mannings", "hydro_from_constants"]


def hydro_from_constants():
    return _lib.hydro(cte.PREC, cte.MANNING, cte.EDGESW, cte.DT_HYDRO)


def diffuse_Q_constant_prec(zh, Q, temp):
    """One multiple-flow-direction relaxation of Q with constant precipitation
    (gf_hydrodynamics.py:23-75): temp = P*dx^2 + slope-weighted shares of Q
    from upslope neighbours, Q = temp.  Each cell gathers its inflow, so the
    result does not depend on atomic order."""
    n = n_cells()
    a = [check("zh", zh, n, torch.float32), check("Q", Q, n, torch.float32), check("temp", temp, n, torch.float32)]
    _lib.call("pff_diffuse_Q_constant_prec", ctx_of(a[0]), grid_from_constants(), hydro_from_constants(),
              *[_lib.ptr(x) for x in a])


def graphflood_core_cte_mannings(h, zh, dh, rec, Q):
    """Manning depth update (gf_hydrodynamics.py:78-129), including the
    reference's double application of dh to h."""
    n = n_cells()
    a = [check("h", h, n, torch.float32), check("zh", zh, n, torch.float32), check("dh", dh, n, torch.float32),
         check("rec", rec, n, torch.int32), check("Q", Q, n, torch.float32)]
    _lib.call("pff_graphflood_core_cte_mannings", ctx_of(a[0]), grid_from_constants(), hydro_from_constants(),
              *[_lib.ptr(x) for x in a])
