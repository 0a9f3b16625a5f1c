"""FlowRouter (pyfastflow/flow/flowfields.py): fields, boundary setup and the
receivers -> lakeflow -> accumulation pipeline on the HIP library."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from .. import constants as cte
from ..field import Field, f32, i32, i64, u1

__all__ = ["FlowRouter", "BOUNDARY_MODES"]

BOUNDARY_MODES = {"normal": 0, "periodic_EW": 1, "periodic_NS": 2, "custom": 3}


class FlowRouter:
    """Routing on an (ny, nx) grid of spacing dx (flowfields.py:37-182).

    boundary_mode: 'normal', 'periodic_EW', 'periodic_NS' or 'custom'
    (unknown strings fall back to 'normal', flowfields.py:169-175);
    boundaries: (ny, nx) uint8 codes 0/1/3/7/9 for 'custom'.

    Fields are Field objects over HBM tensors, with the reference's names.
    Unlike the reference, several routers may coexist: each keeps its own grid
    and passes it to every kernel call.  The constructor still sets the
    constants globals for the module-level functions.
    """

    def __init__(self, nx, ny, dx, boundary_mode="normal", boundaries=None, lakeflow=True,
                 stochastic_receivers=False, device=None, seed=None):
        self.nx = int(nx)
        self.ny = int(ny)
        self.dx = float(dx)
        self.rshp = (self.ny, self.nx)
        self.boundary_mode = boundary_mode
        self.boundaries = boundaries
        self.lakeflow = lakeflow
        self.stochastic_receivers = stochastic_receivers
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.mode = BOUNDARY_MODES.get(boundary_mode, 0)
        N = self.nx * self.ny
        dev = self.device

        def F(dtype, n=N):
            return Field(dtype, (n,), device=dev)

        # core fields (flowfields.py:66-115)
        self.z = F(f32)
        self.Q = F(f32)
        self.Q_ = F(f32)
        self.gradient = F(f32)
        self.receivers = F(i32)
        self.src = F(i32)
        self.ndonors = F(i32)
        self.ndonors_ = F(i32)
        self.donors = F(i32, 4 * N)
        self.donors_ = F(i32, 4 * N)
        # lake fields (flowfields.py:119-159)
        if self.lakeflow:
            self.z_ = F(f32)
            self.z__ = F(f32)
            self.receivers_ = F(i32)
            self.receivers__ = F(i32)
            self.bid = F(i32)
            self.basin_saddlenode = F(i32)
            self.outlet = F(i64)
            self.basin_saddle = F(i64)
            self.is_border = F(u1)
            self.tag = F(u1)
            self.tag_ = F(u1)
            self.change = Field(u1, (), device=dev)
        self._boundaries_dev = None
        if self.mode == 3:
            if boundaries is None:
                raise ValueError("boundary_mode='custom' needs a boundaries array")
            b = np.ascontiguousarray(np.asarray(boundaries).ravel(), dtype=np.uint8)
            if b.size != N:
                raise ValueError(f"boundaries has {b.size} entries, expected {N}")
            self._boundaries_dev = torch.from_numpy(b).to(dev)
        self.last_lake_status = None
        # globals for the module-level functions (flowfields.py:163-179)
        cte.NX, cte.NY, cte.DX = self.nx, self.ny, self.dx
        cte.RAND_RCV = bool(stochastic_receivers)
        if seed is not None:
            cte.RAND_SEED = int(seed)
        cte.BOUND_MODE = self.mode
        cte.boundaries = self._boundaries_dev
        cte.INITIALISED = True

    # ---- plumbing ----------------------------------------------------------
    def _grid(self):
        return _lib.grid(self.nx, self.ny, self.dx, self.mode, self._boundaries_dev)

    def _ctx(self):
        return _lib.ctx_for(self.device)

    def free(self):
        """Drop the fields (flowfields.py:186-200)."""
        for k in list(vars(self)):
            if isinstance(getattr(self, k), Field):
                setattr(self, k, None)

    # ---- I/O (flowfields.py:202-214, 298-319) --------------------------------
    def set_z(self, z):
        self.z.from_numpy(np.asarray(z).ravel().astype(np.float32))

    def get_Q(self):
        return self.Q.to_numpy().reshape(self.rshp)

    def get_Z(self):
        return self.z.to_numpy().reshape(self.rshp)

    # ---- pipeline ------------------------------------------------------------
    def compute_receivers(self):
        """D4 steepest-descent receivers and gradient (flowfields.py:216-226);
        with stochastic_receivers, slopes weighted by seeded uniform draws
        (receivers.py:46-47, see flow.compute_receivers), a fresh seed per
        call from constants.next_rand_seed()."""
        if self.stochastic_receivers:
            _lib.call("pff_compute_receivers_stochastic", self._ctx(), self._grid(), _lib.ptr(self.z),
                      _lib.ptr(self.receivers), _lib.ptr(self.gradient), ctypes.c_uint64(cte.next_rand_seed()))
            return
        _lib.call("pff_compute_receivers", self._ctx(), self._grid(), _lib.ptr(self.z), _lib.ptr(self.receivers),
                  _lib.ptr(self.gradient))

    def reroute_flow(self, carve=True, bid_first=None, log_outlets=False):
        """Route closed depressions through their spill saddles
        (flowfields.py:228-252, lakeflow.py:466-544).

        log_outlets=True also records, for every trip, each basin's outlet
        key and saddle node after the saddle passes (the state the carve
        uses); they are left in self.last_lake_outlets as a list of
        (basin ids, outlet keys, saddle nodes) numpy arrays, one per trip."""
        if not self.lakeflow:
            raise RuntimeError("Flow field not compiled for lakeflow")
        st = _lib.PffLakeStatus()
        bf = bid_first.data if isinstance(bid_first, Field) else bid_first
        log = None
        if log_outlets:
            N = self.nx * self.ny
            log = (torch.empty(N, dtype=torch.int32, device=self.device),
                   torch.empty(N, dtype=torch.int64, device=self.device),
                   torch.empty(N, dtype=torch.int32, device=self.device),
                   (ctypes.c_int64 * 64)())
            _lib.call("pff_ctx_lake_log", self._ctx(), _lib.ptr(log[0]), _lib.ptr(log[1]), _lib.ptr(log[2]), N,
                      log[3], 64)
        _lib.call("pff_reroute_flow", self._ctx(), self._grid(), _lib.ptr(self.bid), _lib.ptr(self.receivers),
                  _lib.ptr(self.receivers_), _lib.ptr(self.receivers__), _lib.ptr(self.z), _lib.ptr(self.z_),
                  _lib.ptr(self.is_border), _lib.ptr(self.outlet), _lib.ptr(self.basin_saddle),
                  _lib.ptr(self.basin_saddlenode), _lib.ptr(self.tag), _lib.ptr(self.tag_), 1 if carve else 0,
                  _lib.ptr(bf), ctypes.byref(st))
        self.last_lake_status = st
        if log is not None:
            b, o, sn = (t.cpu().numpy() for t in log[:3])
            trips, prev = [], 0
            for k in range(64):
                c = int(log[3][k])
                if c < 0:
                    break
                c = min(c, b.size)
                trips.append((b[prev:c].copy(), o[prev:c].copy(), sn[prev:c].copy()))
                prev = c
            self.last_lake_outlets = trips
        return st

    def accumulate_constant_Q(self, value, area=True):
        """Q[i] = value*dx^2 (area) or value, summed over i and everything
        upstream (flowfields.py:254-292).  Computed as exact integer upstream
        counts times the f32 unit value; returns False if the receivers hold
        a cycle (then Q is not meaningful)."""
        conv = ctypes.c_int32(1)
        _lib.call("pff_accumulate_constant_Q", self._ctx(), self._grid(), _lib.ptr(self.receivers), float(value),
                  1 if area else 0, _lib.ptr(self.Q), ctypes.byref(conv))
        return bool(conv.value)

    def accumulate_constant_Q_rake_compress(self, value, area=True):
        """The reference's own schedule: rcv2donor, ceil(log2 N)+2 rake-and-
        compress rounds and fuse (flowfields.py:271-292), on the GPU."""
        logn = math.ceil(math.log2(self.nx * self.ny)) + 1
        ctx, g = self._ctx(), self._grid()
        self.ndonors.fill(0)
        self.src.fill(0)
        self.Q.fill(float(np.float32(value * (self.dx * self.dx if area else 1.0))))
        _lib.call("pff_rcv2donor", ctx, g, _lib.ptr(self.receivers), _lib.ptr(self.donors), _lib.ptr(self.ndonors))
        for it in range(logn + 1):
            _lib.call("pff_rake_compress_accum", ctx, g, _lib.ptr(self.donors), _lib.ptr(self.ndonors),
                      _lib.ptr(self.Q), _lib.ptr(self.src), _lib.ptr(self.donors_), _lib.ptr(self.ndonors_),
                      _lib.ptr(self.Q_), it)
        _lib.call("pff_fuse", ctx, g, _lib.ptr(self.Q), _lib.ptr(self.src), _lib.ptr(self.Q_), logn)

    def fill_z(self, epsilon=1e-4):
        """topofill of router.z in place (flowfields.py:295-296)."""
        from .fill_topo import topofill

        topofill(self, epsilon=epsilon, custom_z=None)
