"""Ragged batches of independent grids routed together (no reference
counterpart: the reference routes one grid per process, environment.py:28-29).

``route_batch`` gives, for every DEM of the batch, what this sequence gives
for that DEM alone::

    router = FlowRouter(nx, ny, dx, boundary_mode="normal")
    router.set_z(dem)
    router.compute_receivers()
    router.reroute_flow(carve=True)          # if lakeflow
    router.accumulate_constant_Q(value, area=area)
    router.receivers, router.Q

(receivers.py:15-58, lakeflow.py:466-544, flowfields.py:254-292), bit for bit,
in one launch sequence over the whole batch.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib

__all__ = ["route_batch", "BatchResult"]


@dataclass
class BatchResult:
    receivers: list = field(default_factory=list)  # int32 (ny*nx,) device tensors
    Q: list = field(default_factory=list)          # float32 (ny*nx,) device tensors
    shapes: list = field(default_factory=list)     # (ny, nx) per tile
    status: dict = field(default_factory=dict)

    def get_Q(self, k: int) -> np.ndarray:
        return self.Q[k].cpu().numpy().reshape(self.shapes[k])

    def get_receivers(self, k: int) -> np.ndarray:
        return self.receivers[k].cpu().numpy().reshape(self.shapes[k])


def _as_dem(d, device) -> torch.Tensor:
    if isinstance(d, torch.Tensor):
        t = d.to(device=device, dtype=torch.float32)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(d, dtype=np.float32))).to(device)
    if t.dim() != 2:
        raise ValueError("each DEM must be a 2-D (ny, nx) array")
    return t.contiguous()


def route_batch(dems, dx: float = 1.0, lakeflow: bool = True, value: float = 1e-5, area: bool = True,
                carve: bool = True, device=None, out: BatchResult | None = None) -> BatchResult:
    """Receivers (after lakeflow) and discharge of each DEM.

    dems: sequence of 2-D (ny, nx) arrays (numpy or torch), 'normal'
    boundaries, all with the same cell size dx.  carve=False selects the
    jump mode, whose result the reference leaves in receivers_ (the returned
    receivers are then the unrerouted ones); it runs tile by tile.  ``out``
    may be the result of an earlier call on DEMs of the same shapes, to reuse
    its output tensors.
    """
    if not torch.cuda.is_available():
        raise RuntimeError("pyfastflow_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    zs = [_as_dem(d, dev) for d in dems]
    shapes = [tuple(z.shape) for z in zs]
    if out is None or out.shapes != shapes:
        out = BatchResult(
            receivers=[torch.empty(ny * nx, dtype=torch.int32, device=dev) for ny, nx in shapes],
            Q=[torch.empty(ny * nx, dtype=torch.float32, device=dev) for ny, nx in shapes],
            shapes=shapes,
        )
    tiles = (_lib.PffTile * max(len(zs), 1))()
    for k, (z, (ny, nx)) in enumerate(zip(zs, shapes)):
        tiles[k] = _lib.PffTile(nx, ny, _lib.ptr(z), _lib.ptr(out.receivers[k]), _lib.ptr(out.Q[k]))
    st = _lib.PffBatchStatus()
    mode = 0 if not lakeflow else (1 if carve else 2)
    _lib.call("pff_batch_route", _lib.ctx_for(dev), len(zs), tiles, float(dx), mode, float(value),
              int(bool(area)), ctypes.byref(st))
    out.status = {
        "cells": int(st.cells),
        "canvas_cells": int(st.canvas_cells),
        "canvas": (int(st.canvas_ny), int(st.canvas_nx)),
        "ndep0": int(st.ndep0),
        "trips": int(st.n_loop),
        "per_tile": bool(st.per_tile),
        "converged": bool(st.converged),
        "capped_tiles": int(st.capped_tiles),
    }
    del zs  # the inputs were consumed by the synchronising call
    return out
