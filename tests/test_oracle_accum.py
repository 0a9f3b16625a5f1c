"""Accumulation of the oracle: the reference's rake-and-compress schedule
(downstream_propag.py, flowfields.py:254-292) and the exact forms, against
brute-force path walks and the outlet-flux invariant."""

import numpy as np
import pytest

from oracle_util import upstream_by_walking
from pyfastflow_amd import synth


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("ny,nx,seed", [(5, 7, 1), (40, 30, 2), (64, 64, 3)])
def test_counts_match_path_walks(ora, ny, nx, seed, mode):
    z = synth.dem(ny, nx, 2.0, 300.0, seed, n_pits=1)
    g = ora.Grid(nx, ny, 1.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    cnt = ora.upstream_count(g, rcv)
    np.testing.assert_array_equal(cnt, upstream_by_walking(rcv))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_reference_schedule_converges_to_counts(ora, mode):
    ny, nx = 96, 80
    z = synth.dem(ny, nx, 2.0, 300.0, 11, n_pits=3)
    g = ora.Grid(nx, ny, 1.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    Q, done = ora.accumulate_constant_Q(g, rcv, 1.0, area=True)
    assert done
    np.testing.assert_array_equal(Q, upstream_by_walking(rcv).astype(np.float32))


def test_outlet_flux_is_total_input(ora):
    """Sum of Q over roots = N * Q0 (the reference's own boundary-flux check,
    archive/_save_fasftflow/fastflow/_archived/main.py:58-67)."""
    ny, nx = 128, 128
    z = synth.dem(ny, nx, 2.0, 500.0, 5, n_pits=4, dx=10.0)
    g = ora.Grid(nx, ny, 10.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    r = ora.reroute_flow(g, z, rcv)
    Q = ora.accumulate_f64(g, r["rec"], np.full(g.N, 1e-3))
    roots = r["rec"] == np.arange(g.N)
    assert abs(Q[roots].sum() - g.N * 1e-3) < 1e-9 * g.N


def test_snake_needs_log_rounds(ora):
    """A single chain of 2*L cells: log2(2L) rake-and-compress rounds finish
    it, fewer leave nodes unfinished (the truncation the schedule's fixed cap
    would produce on deeper trees)."""
    L = 64
    ny, nx = 2, L
    rcv = np.empty(2 * L, np.int32)
    rcv[:L] = np.arange(1, L + 1)  # top row east, last top cell down
    rcv[L - 1] = 2 * L - 1
    rcv[L:] = np.arange(L - 1, 2 * L - 1)  # bottom row west
    rcv[L] = L  # root at bottom-left
    g = ora.Grid(nx, ny, 1.0, 0)
    w = np.ones(2 * L, np.float32)
    Qfull, done = ora.accumulate_rake_compress(g, rcv, w, 8)
    assert done
    np.testing.assert_array_equal(Qfull, upstream_by_walking(rcv).astype(np.float32))
    _, done_short = ora.accumulate_rake_compress(g, rcv, w, 3)
    assert not done_short


def test_f64_weights(ora):
    ny, nx = 50, 40
    z = synth.dem(ny, nx, 2.0, 300.0, 9)
    g = ora.Grid(nx, ny, 1.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    w = np.random.default_rng(0).uniform(0, 1, g.N)
    Q = ora.accumulate_f64(g, rcv, w)
    # brute force: add each cell's weight along its path
    ref = np.zeros(g.N)
    for i in range(g.N):
        j = i
        ref[j] += w[i]
        while rcv[j] != j:
            j = rcv[j]
            ref[j] += w[i]
    np.testing.assert_allclose(Q, ref, rtol=1e-12)
