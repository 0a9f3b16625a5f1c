"""Writes tests/golden/routing_cases.npz: small DEMs (the arrays, not only
their seeds) with the oracle's outputs for the routing path.

    python tests/golden/gen_fixtures.py

Each case stores the DEM, the receivers and gradient, the receivers after
reroute_flow(carve=True), the first trip's basin labels, the depression count
of every trip, every trip's lake outlets (basin id, packed outlet key, saddle
node, for the basins of that trip that have a saddle), the upstream cell counts, and one
fill_z_add_delta pass from h = 0 (the last two are empty where the carve
leaves a receiver cycle).  The oracle that produced them is pinned by
the closed-form and brute-force tests in tests/test_oracle_*.py; the fixtures
freeze its answers (and the DEM synthesiser) against later changes.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

MODES = {"normal": 0, "periodic_EW": 1, "periodic_NS": 2}
INVALID_KEY = 0x4CBEBC200000002A

# name -> (ny, nx, mode, beta, relief, seed, pits, dx); "sentinel"/"capped"
# entries use the small-DEM family of oracle_util.sentinel_dem
CASES = {
    "normal_48x64": (48, 64, "normal", 2.0, 200.0, 11, 5, 10.0),
    "periodic_ew_40x40": (40, 40, "periodic_EW", 2.0, 200.0, 12, 4, 10.0),
    "periodic_ns_32x50": (32, 50, "periodic_NS", 2.2, 150.0, 13, 3, 10.0),
    "tilted_plain_33x17": (33, 17, "normal", 2.0, 50.0, 14, 0, 1.0),
    "sentinel_68": "sentinel",
    "capped_157": "capped",
    # a 32x32 window (rows 270-301, cols 923-954) of a GraphFlood water
    # surface z + h after 15 iterations on the 2048x2048 GraphFlood workload
    # (synth.graphflood_plain): the lakeflow carve leaves a receiver cycle
    # there (only 2-cycles are removed), so every later step runs on a graph
    # that is not a forest; stored as tests/golden/cycle_dem_32x32.npy
    "cycle_32x32": "file",
}


def case_dem(name):
    from oracle_util import sentinel_dem
    from pyfastflow_amd import synth

    spec = CASES[name]
    if spec == "file":
        return np.load(os.path.join(HERE, "cycle_dem_32x32.npy")).astype(np.float32), "normal", 10.0
    if isinstance(spec, str):
        seed = int(name.split("_")[1])
        z = sentinel_dem(seed)
        return z, "normal", 10.0
    ny, nx, mode, beta, relief, seed, pits, dx = spec
    tilt = 0.01 if pits == 0 else 0.0
    z = synth.dem(ny, nx, beta, relief, seed, n_pits=pits, tilt=tilt, dx=dx, periodic_ew=(mode == "periodic_EW"))
    return z, mode, dx


def oracle_outputs(z, mode, dx):
    from oracle import pff_oracle as O

    ny, nx = z.shape
    g = O.Grid(nx, ny, dx, MODES[mode])
    rcv, grad = O.compute_receivers(g, z)
    lk = O.reroute_flow(g, z, rcv, carve=True, log_iters=64)
    trips_b, trips_o, trips_s, trip_len = [], [], [], []
    for t in range(lk["n_iter"]):
        ol, sn = lk["outlet_log"][t], lk["saddlenode_log"][t]
        # basins with a saddle or an outlet; a basin without a saddle keeps the
        # invalid entries every slot starts the trip with and is left out
        b = np.flatnonzero((ol != INVALID_KEY) | (sn != -1)).astype(np.int32)
        trips_b.append(b)
        trips_o.append(ol[b])
        trips_s.append(sn[b])
        trip_len.append(b.size)
    try:
        cnt = O.upstream_count(g, lk["rec"])
        zh, h = O.fill_z_add_delta(g, z, np.zeros(nx * ny, np.float32), lk["rec"], 1e-3)
    except ValueError:  # the carve left a receiver cycle: no upstream counts or fill
        cnt = np.zeros(0, np.uint32)
        zh = h = np.zeros(0, np.float32)
    return dict(rcv=rcv, gradient=grad, rec=lk["rec"], bid_first=lk["bid_first"],
                ndep=np.asarray(lk["ndep"], np.int64), trip_len=np.asarray(trip_len, np.int64),
                trip_basin=np.concatenate(trips_b) if trips_b else np.zeros(0, np.int32),
                trip_outlet=np.concatenate(trips_o) if trips_o else np.zeros(0, np.int64),
                trip_snode=np.concatenate(trips_s) if trips_s else np.zeros(0, np.int32),
                upstream=cnt.astype(np.uint32), fill_zh=zh, fill_h=h)


def main():
    out = {}
    for name in CASES:
        z, mode, dx = case_dem(name)
        res = oracle_outputs(z, mode, dx)
        out[f"{name}/z"] = z
        out[f"{name}/meta"] = np.asarray([z.shape[0], z.shape[1], MODES[mode]], np.int64)
        out[f"{name}/dx"] = np.asarray([dx], np.float64)
        for k, v in res.items():
            out[f"{name}/{k}"] = v
    path = os.path.join(HERE, "routing_cases.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(CASES)} cases, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
