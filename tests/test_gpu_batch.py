"""GPU parity of the ragged batch (pff_batch_route): every tile of a batch gets
what the single-grid path, and the oracle, give for that tile alone.

Integer outputs (receivers) bit-exact; Q bit-exact against the single-grid
GPU run and within 1e-5 relative of f32(q0) * exact upstream count.
"""

import os

import numpy as np
import pytest
import torch

from oracle_util import CAPPED_SEEDS, SENTINEL_SEEDS, sentinel_dem
from pyfastflow_amd import synth

pytestmark = pytest.mark.gpu

DX = 10.0


def _single(pf, z, lakeflow=True, carve=True, value=1e-5):
    ny, nx = z.shape
    r = pf.flow.FlowRouter(nx, ny, DX, boundary_mode="normal")
    r.set_z(z)
    r.compute_receivers()
    if lakeflow:
        r.reroute_flow(carve=carve)
    r.accumulate_constant_Q(value, area=True)
    return r.receivers.to_numpy().copy(), r.Q.to_numpy().copy()


def _oracle(ora, z, lakeflow=True):
    ny, nx = z.shape
    g = ora.Grid(nx, ny, DX, 0)
    rcv, _ = ora.compute_receivers(g, z)
    if lakeflow:
        rcv = ora.reroute_flow(g, z, rcv, carve=True)["rec"]
    return g, rcv


def _check_q(ora, g, rcv, Q, value=1e-5):
    cnt = ora.upstream_count(g, rcv).astype(np.float64)
    q0 = float(np.float32(value * DX * DX))
    ref = cnt * q0
    rel = np.abs(Q.astype(np.float64) - ref) / ref
    assert rel.max() <= 1e-5, rel.max()


def _mixed_tiles():
    tiles = [sentinel_dem(s, DX) for s in SENTINEL_SEEDS + CAPPED_SEEDS]
    rng = np.random.default_rng(7)
    for k in range(20):
        ny, nx = int(rng.integers(2, 80)), int(rng.integers(2, 80))
        tiles.append(synth.dem(ny, nx, 2.0, 80.0, 1000 + k, n_pits=int(rng.integers(0, 6)), dx=DX))
    tiles += [np.zeros((2, 2), np.float32), np.ones((3, 5), np.float32), synth.dem(2, 9, 2.0, 5.0, 3, dx=DX)]
    return tiles


def test_batch_vs_oracle_per_tile(pf, ora):
    tiles = _mixed_tiles()
    print(f"case batch mixed {len(tiles)} tiles", flush=True)
    res = pf.flow.route_batch(tiles, dx=DX, lakeflow=True, value=1e-5)
    assert not res.status["per_tile"]
    assert res.status["cells"] == sum(t.size for t in tiles)
    capped = 0
    for k, z in enumerate(tiles):
        g, rcv = _oracle(ora, z)
        capped += ora.depression_counter(g, rcv) > 0
        got = res.receivers[k].cpu().numpy()
        np.testing.assert_array_equal(got, rcv, err_msg=f"tile {k} {z.shape}")
        _check_q(ora, g, rcv, res.Q[k].cpu().numpy())
    # tiles that stop at their own trip cap (the canvas freezes them there)
    assert capped >= len(CAPPED_SEEDS)
    assert res.status["capped_tiles"] == capped


def test_ragged_256_tiles_vs_oracle(pf, ora):
    """The batched-routing workload itself: 256 tiles with sides log-uniform
    in [128, 1024] (seed 404), routed in one launch sequence; every tile's
    receivers equal the oracle's reroute_flow(carve) bit for bit, and its Q
    equals f32(q0) times the oracle's exact upstream count (error stated per
    tile: the largest relative error of all tiles against the FP64 sums)."""
    tiles = [t["z"] for t in synth.ragged_batch(256)]
    ncell = sum(t.size for t in tiles)
    print(f"case batch ragged-256: {ncell} cells, largest side {max(t.shape[0] for t in tiles)}", flush=True)
    res = pf.flow.route_batch(tiles, dx=DX, lakeflow=True, value=1e-5)
    torch.cuda.synchronize()
    assert not res.status["per_tile"]
    assert res.status["cells"] == ncell
    q0 = np.float32(1e-5 * DX * DX)
    worst = (0.0, -1)
    for k, z in enumerate(tiles):
        g, rcv = _oracle(ora, z)
        np.testing.assert_array_equal(res.receivers[k].cpu().numpy(), rcv, err_msg=f"tile {k} {z.shape}")
        cnt = ora.upstream_count(g, rcv)
        Q = res.Q[k].cpu().numpy()
        np.testing.assert_array_equal(Q, (q0 * cnt.astype(np.float32)).astype(np.float32),
                                      err_msg=f"tile {k} {z.shape}")
        rel = float((np.abs(Q.astype(np.float64) - cnt * float(q0)) / (cnt * float(q0))).max())
        worst = max(worst, (rel, k))
        if k % 32 == 31:
            print(f"  {k + 1} tiles checked", flush=True)
    print(f"  Q max relative error vs FP64: {worst[0]:.3e} (tile {worst[1]})", flush=True)
    assert worst[0] <= 1e-5


def test_batch_vs_single_grid_runs(pf):
    tiles = [t["z"] for t in synth.ragged_batch(n_tiles=24, seed=404, lo=64, hi=512)]
    print(f"case batch ragged {len(tiles)} tiles, {sum(t.size for t in tiles)} cells", flush=True)
    res = pf.flow.route_batch(tiles, dx=DX, lakeflow=True, value=1e-5)
    assert not res.status["per_tile"]
    assert res.status["ndep0"] > 0
    for k, z in enumerate(tiles):
        rcv, Q = _single(pf, z)
        np.testing.assert_array_equal(res.receivers[k].cpu().numpy(), rcv, err_msg=f"tile {k} {z.shape}")
        np.testing.assert_array_equal(res.Q[k].cpu().numpy().view(np.uint32), Q.view(np.uint32))


def test_batch_without_lakeflow(pf, ora):
    tiles = _mixed_tiles()[:10]
    res = pf.flow.route_batch(tiles, dx=DX, lakeflow=False, value=1.0)
    for k, z in enumerate(tiles):
        g, rcv = _oracle(ora, z, lakeflow=False)
        np.testing.assert_array_equal(res.receivers[k].cpu().numpy(), rcv)
        _check_q(ora, g, rcv, res.Q[k].cpu().numpy(), value=1.0)


def test_batch_jump_mode_runs_tile_by_tile(pf):
    """Jump mode leaves the reference's receivers unrerouted (its result is
    in receivers_, lakeflow.py:527-544); the batch then runs per tile."""
    tiles = _mixed_tiles()[:8]
    res = pf.flow.route_batch(tiles, dx=DX, lakeflow=True, carve=False, value=1e-5)
    assert res.status["per_tile"]
    for k, z in enumerate(tiles):
        rcv, Q = _single(pf, z, lakeflow=True, carve=False)
        np.testing.assert_array_equal(res.receivers[k].cpu().numpy(), rcv)
        np.testing.assert_array_equal(res.Q[k].cpu().numpy().view(np.uint32), Q.view(np.uint32))


def test_batch_reuse_and_empty(pf):
    tiles = _mixed_tiles()[4:9]
    a = pf.flow.route_batch(tiles, dx=DX)
    first = [q.clone() for q in a.Q]
    b = pf.flow.route_batch(tiles, dx=DX, out=a)
    assert b is a
    for x, y in zip(first, b.Q):
        assert torch.equal(x, y)
    e = pf.flow.route_batch([], dx=DX)
    assert e.status["cells"] == 0 and e.receivers == []


def test_batch_with_a_receiver_cycle_runs_tile_by_tile(pf, ora):
    """A tile whose lakeflow carve leaves a receiver cycle (only 2-cycles are
    removed; tests/golden/cycle_dem_32x32.npy) sends the batch down the
    tile-by-tile path; every tile's receivers still match the oracle and its
    Q the single-grid run."""
    cyc = np.load(os.path.join(os.path.dirname(__file__), "golden", "cycle_dem_32x32.npy")).astype(np.float32)
    tiles = _mixed_tiles()[:6] + [cyc]
    print(f"case batch with a cycle tile, {len(tiles)} tiles", flush=True)
    res = pf.flow.route_batch(tiles, dx=DX, lakeflow=True, value=1e-5)
    assert res.status["per_tile"]
    for k, z in enumerate(tiles):
        g, rcv = _oracle(ora, z)
        np.testing.assert_array_equal(res.receivers[k].cpu().numpy(), rcv, err_msg=f"tile {k} {z.shape}")
        _, q1 = _single(pf, z)
        np.testing.assert_array_equal(res.Q[k].cpu().numpy().view(np.uint32), q1.view(np.uint32))
