"""Flooder (pyfastflow/graphflood/gf_fields.py): water depth state and the
GraphFlood iteration on a FlowRouter."""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from .. import constants as cte
from ..field import Field, f32

__all__ = ["Flooder"]


class Flooder:
    """GraphFlood on `router` (gf_fields.py:28-108).

    precipitation_rates in m/s, manning's n, edge_slope for cells draining out
    of the domain; dt is constants.DT_HYDRO (5e-3 s) unless given.
    """

    def __init__(self, router, precipitation_rates=10e-3 / 3600, manning=0.033, edge_slope=1e-2, dt=None):
        self.nx, self.ny, self.dx = router.nx, router.ny, router.dx
        self.rshp = router.rshp
        self.router = router
        N = self.nx * self.ny
        self.h = Field(f32, (N,), device=router.device)
        self.dh = Field(f32, (N,), device=router.device)
        self.precipitation_rates = float(precipitation_rates)
        self.manning = float(manning)
        self.edge_slope = float(edge_slope)
        self.dt = float(cte.DT_HYDRO if dt is None else dt)
        # globals for the module-level kernels (gf_fields.py:106-108)
        cte.PREC = self.precipitation_rates
        cte.MANNING = self.manning
        cte.EDGESW = self.edge_slope
        self.iterations = 0

    def _hydro(self):
        return _lib.hydro(self.precipitation_rates, self.manning, self.edge_slope, self.dt)

    def step(self, n_diffuse=5):
        """One GraphFlood iteration (gf_fields.py:126-141)."""
        r = self.router
        ctx, g, hy = r._ctx(), r._grid(), self._hydro()
        r.compute_receivers()
        r.reroute_flow()
        _lib.call("pff_fill_z_add_delta", ctx, g, _lib.ptr(r.z), _lib.ptr(self.h), _lib.ptr(r.z_),
                  _lib.ptr(r.receivers), _lib.ptr(r.receivers_), _lib.ptr(r.receivers__), 1e-3)
        for _ in range(n_diffuse):
            _lib.call("pff_diffuse_Q_constant_prec", ctx, g, hy, _lib.ptr(r.z), _lib.ptr(r.Q), _lib.ptr(r.Q_))
        _lib.call("pff_graphflood_core_cte_mannings", ctx, g, hy, _lib.ptr(self.h), _lib.ptr(r.z), _lib.ptr(self.dh),
                  _lib.ptr(r.receivers), _lib.ptr(r.Q))
        self.iterations += 1

    def run_graphflood(self, N=10):
        """N GraphFlood iterations (gf_fields.py:111-141)."""
        for _ in range(int(N)):
            self.step()

    def run_to_steady(self, tol=1e-6, check_every=100, max_iter=1_000_000):
        """Iterate until max|h(t) - h(t - check_every)| < tol, checked every
        `check_every` iterations.  Returns (iterations, last max|dh|,
        converged)."""
        prev = self.h.data.clone()
        it = 0
        res = float("inf")
        while it < max_iter:
            for _ in range(check_every):
                self.step()
            it += check_every
            res = float((self.h.data - prev).abs().max().item())
            if res < tol:
                return it, res, True
            prev.copy_(self.h.data)
        return it, res, False

    def set_h(self, val):
        """Set h (gf_fields.py:143-150); zh is not touched, as in the reference."""
        self.h.from_numpy(np.asarray(val, dtype=np.float32).ravel())

    def get_h(self):
        return self.h.to_numpy().reshape(self.rshp)
