"""Grid and hydraulic parameters (pyfastflow/constants.py:20-103).

The reference compiles these module globals into its kernels at first JIT;
here they are read at each call and passed to the HIP library as a
``pff_grid`` / ``pff_hydro`` struct, so changing them takes effect
immediately and several grids can coexist in one process.  Module-level
kernel functions (``pyfastflow_amd.flow.compute_receivers`` and friends) use
these globals as the reference does; ``FlowRouter`` sets them in its
constructor (flowfields.py:163-179).
"""

from __future__ import annotations

import numpy as np
import torch

INITIALISED = False

# grid (constants.py:27-50)
DX = 1.0
NX = 512
NY = 512
RAND_RCV = False
# stochastic receivers draw from a seeded counter-based generator (not
# Taichi's ti.random): call k of compute_receivers uses next_rand_seed()
RAND_SEED = 0x5EED
_rand_calls = 0


def next_rand_seed() -> int:
    """Seed of the next stochastic receivers call: splitmix64 of
    RAND_SEED + call count, so every call draws fresh numbers and a run is
    reproducible from RAND_SEED."""
    global _rand_calls
    x = (RAND_SEED + _rand_calls * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    _rand_calls += 1
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)
# 0 normal, 1 periodic_EW, 2 periodic_NS, 3 custom (codes 0/1/3/7/9)
BOUND_MODE = 0

boundaries = None  # device uint8 tensor of custom codes (mode 3)

# hydraulics (constants.py:93-103)
PREC = 10 * 1e-3 / 3600
MANNING = 0.033
EDGESW = 1e-2
DT_HYDRO = 5e-3


def init_custom_boundaries(tboundaries: np.ndarray):
    """Install per-node boundary codes and switch to mode 3 (constants.py:56-83)."""
    global boundaries, BOUND_MODE
    arr = np.ascontiguousarray(np.asarray(tboundaries).ravel(), dtype=np.uint8)
    if arr.size != NX * NY:
        raise ValueError(f"boundaries has {arr.size} entries, expected NX*NY = {NX * NY}")
    boundaries = torch.from_numpy(arr).to(torch.device("cuda", torch.cuda.current_device()))
    BOUND_MODE = 3


def grid_args():
    """(nx, ny, dx, mode, boundaries) from the globals."""
    return NX, NY, DX, BOUND_MODE, boundaries if BOUND_MODE == 3 else None
