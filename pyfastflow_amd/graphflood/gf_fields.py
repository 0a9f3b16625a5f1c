"""Flooder (pyfastflow/graphflood/gf_fields.py): water depth state and the
GraphFlood iteration on a FlowRouter."""

from __future__ import annotations

import ctypes

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

    update selects the depth update of graphflood_core_cte_mannings:
    "reference" (default) adds dh to h twice and to zh once, exactly as
    gf_hydrodynamics.py:121-129; "single" adds dh once to both and clamps h
    at 0.  With "reference", filled lakes lose depth every iteration until h
    is negative and h^(5/3) is NaN, so runs to steady state use "single".
    """

    def __init__(self, router, precipitation_rates=10e-3 / 3600, manning=0.033, edge_slope=1e-2, dt=None,
                 update="reference"):
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
        if update not in ("reference", "single"):
            raise ValueError("update must be 'reference' or 'single'")
        self.update = update
        # globals for the module-level kernels (gf_fields.py:106-108)
        cte.PREC = self.precipitation_rates
        cte.MANNING = self.manning
        cte.EDGESW = self.edge_slope
        self.iterations = 0

    def _hydro(self):
        return _lib.hydro(self.precipitation_rates, self.manning, self.edge_slope, self.dt,
                          0 if self.update == "reference" else 1)

    def step_modules(self, n_diffuse=5):
        """One GraphFlood iteration through the module-level calls, as the
        reference sequences them (gf_fields.py:126-141)."""
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

    def step(self, n_diffuse=5):
        """One GraphFlood iteration (gf_fields.py:126-141)."""
        self.run_graphflood(1, n_diffuse=n_diffuse)

    def run_graphflood(self, N=10, n_diffuse=5):
        """N GraphFlood iterations (gf_fields.py:111-141), looped inside the
        HIP library (on a router with stochastic receivers, through the
        module-level calls, so that each iteration draws its receivers)."""
        r = self.router
        if r.stochastic_receivers:
            for _ in range(int(N)):
                self.step_modules(n_diffuse)
            return None
        st = _lib.PffGfStatus()
        _lib.call("pff_graphflood_run", r._ctx(), r._grid(), self._hydro(), _lib.ptr(r.z), _lib.ptr(self.h),
                  _lib.ptr(r.Q), _lib.ptr(r.receivers), int(N), int(n_diffuse), 1e-3, ctypes.byref(st))
        self.iterations += int(N)
        return st

    def run_to_steady(self, tol=1e-6, check_every=100, max_iter=1_000_000, n_diffuse=5, frac=0.0):
        """Iterate until steady, checked every `check_every` iterations on the
        GPU on d = |h(t) - h(t - check_every)|: frac=0 asks max d < tol; with
        0 < frac < 1 at most that fraction of cells may have d > tol.  Stops
        at max_iter.  Returns the PffGfStatus (iterations, residual (max d),
        above_frac, converged, NaN cells, and the first check's values)."""
        r = self.router
        st = _lib.PffGfStatus()
        _lib.call("pff_graphflood_to_steady", r._ctx(), r._grid(), self._hydro(), _lib.ptr(r.z), _lib.ptr(self.h),
                  _lib.ptr(r.Q), _lib.ptr(r.receivers), float(tol), float(frac), int(check_every), int(max_iter),
                  int(n_diffuse), 1e-3, ctypes.byref(st))
        self.iterations += int(st.iterations)
        return st

    def set_h(self, val):
        """Set h (gf_fields.py:143-150); zh is not touched, as in the reference."""
        self.h.from_numpy(np.asarray(val, dtype=np.float32).ravel())

    def get_h(self):
        return self.h.to_numpy().reshape(self.rshp)

    # ---- checkpoint ------------------------------------------------------------
    # No reference counterpart: the reference moves state by hand with
    # set_z/get_Z and set_h/get_h (flowfields.py:202-214,310-319,
    # gf_fields.py:143-159).  An iteration reads the surface zh (the router's
    # z), h and the discharge Q of the previous one (the diffusion starts from
    # it, gf_hydrodynamics.py:23-75); receivers, lakes and the fill are
    # recomputed from zh, so these three arrays are the whole state.
    _PARAMS = ("precipitation_rates", "manning", "edge_slope", "dt")

    def save_state(self, path):
        """Write zh, h, Q, the iteration count, the grid and the parameters to
        an .npz file (numpy arrays only, no pickled objects)."""
        r = self.router
        extra = {k: np.float64(getattr(self, k)) for k in self._PARAMS}
        np.savez(path, zh=r.z.to_numpy(), h=self.h.to_numpy(), Q=r.Q.to_numpy(), iterations=np.int64(self.iterations),
                 nx=np.int64(self.nx), ny=np.int64(self.ny), dx=np.float64(self.dx), update=np.str_(self.update),
                 **extra)

    def load_state(self, path):
        """Restore what save_state wrote on a Flooder of the same grid (nx, ny,
        dx): zh into the router's z, h, Q and the iteration count, so that the
        next iteration continues the saved run.  This Flooder keeps its own
        parameters; the saved ones are returned (a dict).  With stochastic
        receivers the draws continue from this process's seed counter."""
        r = self.router
        with np.load(path, allow_pickle=False) as f:
            shape = (int(f["nx"]), int(f["ny"]), float(f["dx"]))
            if shape != (self.nx, self.ny, self.dx):
                raise ValueError(f"state of a {shape[0]} x {shape[1]} grid with dx = {shape[2]} cannot be loaded "
                                 f"on a {self.nx} x {self.ny} grid with dx = {self.dx}")
            r.z.from_numpy(f["zh"])
            self.h.from_numpy(f["h"])
            r.Q.from_numpy(f["Q"])
            self.iterations = int(f["iterations"])
            saved = {k: float(f[k]) for k in self._PARAMS}
            saved["update"] = str(f["update"])
        return saved
