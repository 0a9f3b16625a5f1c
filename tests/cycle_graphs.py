"""Receiver graphs with cycles for the lakeflow tests: a DEM's steepest-
descent receivers with loops written over them (a long loop across many
64x64 tiles, a loop of 128 cells, 4-cycles inside a tile and across a tile
corner, a 2-cycle across a tile edge, and a 4-cycle through grid-edge
cells).  Cells whose receivers point into a loop drain into a cycle."""

import numpy as np

from pyfastflow_amd import synth


def _loop(r0, c0, h, w):
    """Cells of the perimeter of the rectangle of h x w steps with corner
    (r0, c0), walked clockwise; 2(h + w) cells."""
    cells = []
    cells += [(r0, c0 + k) for k in range(w)]
    cells += [(r0 + k, c0 + w) for k in range(h)]
    cells += [(r0 + h, c0 + w - k) for k in range(w)]
    cells += [(r0 + h - k, c0) for k in range(h)]
    return cells


def cycle_case(ny=200, nx=300, seed=5, dx=10.0):
    """(z, rcv, loops): z a DEM, rcv its receivers with the loops written in."""
    from oracle import pff_oracle as ora

    z = synth.dem(ny, nx, 2.0, 60.0, seed, n_pits=40, dx=dx)
    g = ora.Grid(nx, ny, dx, 0)
    rcv, _ = ora.compute_receivers(g, z.ravel())
    rcv = rcv.copy()
    loops = [
        _loop(20, 20, 130, 230),   # 720 cells over many tiles
        _loop(160, 30, 32, 32),    # 128 = 2^7 cells, across a tile row
        _loop(70, 70, 1, 1),       # 4-cycle inside a tile
        _loop(63, 127, 1, 1),      # 4-cycle across a tile corner
        _loop(0, 0, 1, 1),         # 4-cycle through grid-edge cells
    ]
    loops.append([(10, 127), (10, 128)])  # 2-cycle across a tile edge
    for cells in loops:
        idx = [r * nx + c for r, c in cells]
        for a, b in zip(idx, idx[1:] + idx[:1]):
            rcv[a] = b
    return z.ravel().astype(np.float32), rcv.astype(np.int32), loops


def power(f, k):
    """f composed with itself 2^k times, by squaring."""
    p = f.copy()
    for _ in range(k):
        p = p[p]
    return p


def jumped(f, npass):
    """f^(2^npass - 1) = f o f^2 o f^4 o ... o f^(2^(npass-1))."""
    out = np.arange(f.size, dtype=np.int64)
    p = f.astype(np.int64)
    for _ in range(npass):
        out = p[out]
        p = p[p]
    return out


def forward_closure(f, seeds):
    """Every cell reachable from a seed along f (seeds included)."""
    tag = np.zeros(f.size, bool)
    for s in seeds:
        x = int(s)
        while not tag[x]:
            tag[x] = True
            x = int(f[x])
    return tag
