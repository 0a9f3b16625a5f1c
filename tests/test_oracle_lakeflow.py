"""Lakeflow of the oracle (lakeflow.py) against closed forms (labels =
init(root), depression count), the priority-flood spill elevation, hand-built
DEMs and invariants of the rerouted receivers."""

import numpy as np
import pytest

from oracle_util import neighbour_table, outlets, roots_by_walking, spill_elevation
from pyfastflow_amd import synth

INVALID_KEY = 0x4CBEBC200000002A

INVALID = 0x4CBEBC200000002A


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("ny,nx,seed", [(5, 7, 1), (48, 40, 2), (64, 64, 3)])
def test_labels_are_init_of_root(ora, ny, nx, seed, mode):
    z = synth.dem(ny, nx, 2.0, 300.0, seed, n_pits=2)
    g = ora.Grid(nx, ny, 1.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    bid = ora.basin_identification(g, rcv)
    root = roots_by_walking(rcv)
    out = outlets(ny, nx, mode)
    want = np.where(out[root], 0, root + 1)
    np.testing.assert_array_equal(bid, want)
    assert ora.depression_counter(g, rcv) == int(np.sum((rcv == np.arange(g.N)) & ~out))


def _bowl(ny=9, nx=9, tilt=0.1):
    """A tilted plane draining south with one closed bowl in the middle."""
    r, c = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    z = 10.0 + tilt * (ny - 1 - r) + 0.01 * c
    z[3:6, 3:6] += 2.0          # rim
    z[4, 4] = 9.0               # pit below the rim
    return z.astype(np.float32)


def test_single_bowl_saddle_is_priority_flood_spill(ora):
    z = _bowl()
    ny, nx = z.shape
    g = ora.Grid(nx, ny, 1.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    pit = 4 * nx + 4
    assert rcv[pit] == pit
    bid = ora.basin_identification(g, rcv)
    s = ora.saddlesort(g, bid, z)
    b = pit + 1
    key = int(s["basin_saddle"][b])
    f, _ = ora.unpack_array(np.array([key], np.int64))
    assert float(f[0]) == pytest.approx(spill_elevation(z, ny, nx, 0, pit), abs=0)
    # the outlet is a neighbour of the saddle node in another basin, lowest first
    node = int(s["saddlenode"][b])
    nb = neighbour_table(ny, nx, 0)[:, node]
    of, oi = ora.unpack_array(np.array([s["outlet"][b]], np.int64))
    assert oi[0] in nb and bid[oi[0]] != b
    cand = [z.ravel()[j] for j in nb if j >= 0 and bid[j] != b]
    assert of[0] == min(cand)


def test_single_bowl_carve_reverses_the_pit_path(ora):
    z = _bowl()
    ny, nx = z.shape
    g = ora.Grid(nx, ny, 1.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    pit = 4 * nx + 4
    bid = ora.basin_identification(g, rcv)
    s = ora.saddlesort(g, bid, z)
    node = int(s["saddlenode"][pit + 1])
    # original path saddle node -> pit
    path = [node]
    while rcv[path[-1]] != path[-1]:
        path.append(rcv[path[-1]])
    assert path[-1] == pit
    r = ora.reroute_flow(g, z, rcv)
    new = r["rec"]
    # reversed: each path node now drains to its predecessor; the saddle node to the outlet
    for a, b in zip(path[:-1], path[1:]):
        assert new[b] == a
    _, oi = ora.unpack_array(np.array([s["outlet"][pit + 1]], np.int64))
    assert new[node] == oi[0]
    assert ora.depression_counter(g, new) == 0


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_rerouted_receivers_drain_out(ora, mode, seed):
    """After carving: every receiver is a D4 link or self, no interior pit is
    left, the graph is acyclic and every cell reaches a cell that can leave."""
    ny, nx = 96, 96
    z = synth.dem(ny, nx, 2.0, 300.0, seed, n_pits=6, dx=10.0, periodic_ew=(mode == 1))
    g = ora.Grid(nx, ny, 10.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    r = ora.reroute_flow(g, z, rcv, log_iters=16)
    new = r["rec"]
    nb = neighbour_table(ny, nx, mode)
    i = np.arange(g.N)
    assert np.all((new == i) | np.any(nb == new[None, :], axis=0))
    assert ora.depression_counter(g, new) == 0
    root = roots_by_walking(new)  # raises on a cycle
    assert np.all(outlets(ny, nx, mode)[root])
    # per trip: no 2-cycles remain in the basin graph
    bidf = r["bid_first"]
    ol = r["outlet_log"][0]
    valid = np.nonzero(ol != INVALID)[0]
    _, onode = ora.unpack_array(ol[valid])
    tgt = bidf[onode]
    back = {}
    for b, t in zip(valid, tgt):
        back[int(b)] = int(t)
    for b, t in back.items():
        if t != 0 and t in back:
            assert back[t] != b


def test_trip_count_respects_reference_cap(ora):
    ny, nx = 128, 128
    z = synth.dem(ny, nx, 2.0, 300.0, 5, n_pits=10, dx=10.0)
    g = ora.Grid(nx, ny, 10.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    r = ora.reroute_flow(g, z, rcv)
    cap = int(np.ceil(np.log2(r["ndep0"]))) + 1
    assert r["n_loop"] <= cap
    # each trip at least halves the number of pits (2-cycles lose one side)
    nd = r["ndep"]
    for a, b in zip(nd[:-1], nd[1:]):
        assert b <= a // 2 + (a % 2)


def test_jump_mode_leaves_receivers_unchanged(ora):
    """reroute_flow(carve=False) writes only the working copy (lakeflow.py:527-544)."""
    z = _bowl()
    ny, nx = z.shape
    g = ora.Grid(nx, ny, 1.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    r = ora.reroute_flow(g, z, rcv, carve=False)
    np.testing.assert_array_equal(r["rec"], rcv)
    pit = 4 * nx + 4
    assert r["rec_"][pit] != pit


def test_sentinel_rule_fires_on_known_dems(ora):
    """The cycle pass reads outlets that the same pass removed earlier (serial
    order, lakeflow.py:255-281); a removed outlet is the sentinel and decodes
    to node 42.  These DEMs exercise that read, so GPU parity on them pins
    the rule (and, in a batch, that node 42 is the tile's own cell 42)."""
    from oracle_util import SENTINEL_SEEDS, sentinel_dem

    for s in SENTINEL_SEEDS:
        z = sentinel_dem(s)
        ny, nx = z.shape
        g = ora.Grid(nx, ny, 10.0, 0)
        rcv, _ = ora.compute_receivers(g, z)
        ora.sentinel_hits(True)
        lk = ora.reroute_flow(g, z, rcv, carve=True)
        assert ora.sentinel_hits(True) >= 1, s
        # the rerouted receivers drain every cell out of the grid
        roots = roots_by_walking(lk["rec"])
        assert outlets(ny, nx, 0)[roots].all()


def test_capped_dems_keep_a_depression(ora):
    """A closed basin whose every edge cell drains into it, touching the only
    open cells through edge cells, has no border cell (can_leave cells are
    skipped, lakeflow.py:132-157) and so no saddle: the trip loop runs its
    cap ceil(log2 Ndep0)+1 (:500) and ends with that depression."""
    from oracle_util import CAPPED_SEEDS, sentinel_dem

    for s in CAPPED_SEEDS:
        z = sentinel_dem(s)
        ny, nx = z.shape
        g = ora.Grid(nx, ny, 10.0, 0)
        rcv, _ = ora.compute_receivers(g, z)
        lk = ora.reroute_flow(g, z, rcv, carve=True)
        cap = int(np.ceil(np.log2(lk["ndep0"]))) + 1
        assert lk["n_loop"] == cap, s
        assert ora.depression_counter(g, lk["rec"]) >= 1
        roots = roots_by_walking(lk["rec"])
        left = np.unique(roots[~outlets(ny, nx, 0)[roots]])
        assert left.size >= 1
        # that basin has no interior cell next to a cell of another basin
        bid = ora.basin_identification(g, lk["rec"], 64)
        nb = neighbour_table(ny, nx, 0)
        out = outlets(ny, nx, 0)
        for p in left:
            cells = np.flatnonzero(bid == p + 1)
            for c in cells:
                if out[c]:
                    continue
                for k in range(4):
                    j = nb[k, c]
                    assert j < 0 or bid[j] == p + 1


def test_labels_on_cycles_are_the_jumped_pointer(ora):
    """ceil(log2 N)+1 synchronous passes bid <- bid[rec], rec <- rec[rec]
    give bid0[f^(2^P - 1)(i)] on any receiver graph (by induction on the
    passes); on a forest that is the root's label.  Checked on graphs with
    loops of 720, 128 (a power of two, which the doubling folds onto
    itself), 4 and 2 cells."""
    from cycle_graphs import cycle_case, jumped

    z, rcv, _ = cycle_case()
    g = ora.Grid(300, 200, 10.0, 0)
    npass = int(np.ceil(np.log2(g.N))) + 1
    bid0 = ora.basin_id_init(g)
    want = bid0[jumped(rcv, npass)]
    np.testing.assert_array_equal(ora.basin_identification(g, rcv), want)


def test_carve_on_cycles_tags_the_forward_closure(ora):
    """The carve's tag loop runs until nothing changes, so it tags every cell
    reachable from a saddle node, loops included; the reversal then points
    each tagged cell at its largest tagged donor and the saddle nodes at
    their outlets."""
    from cycle_graphs import cycle_case, forward_closure

    z, rcv, loops = cycle_case()
    g = ora.Grid(300, 200, 10.0, 0)
    bid = ora.basin_identification(g, rcv)
    ss = ora.saddlesort(g, bid, z)
    sn, ol = ss["saddlenode"], ss["outlet"]
    seeds = sn[sn >= 0]
    tag = forward_closure(rcv, seeds)
    loop_cells = {r * 300 + c for cells in loops for r, c in cells}
    assert any(tag[c] for c in loop_cells), "no saddle path reaches a loop"
    rec, otag, _ = ora.reroute_carve(g, rcv, sn, ol)
    np.testing.assert_array_equal(otag.astype(bool), tag)
    want = rcv.astype(np.int64).copy()
    i = np.flatnonzero(tag & (rcv != np.arange(g.N)))
    i = i[tag[rcv[i]]]
    best = np.full(g.N, -1, np.int64)
    np.maximum.at(best, rcv[i], i)
    hit = best >= 0
    want[hit] = best[hit]
    b = np.flatnonzero(ol != INVALID_KEY)
    want[sn[b]] = (ol[b] & 0xFFFFFFFF).astype(np.int64)
    np.testing.assert_array_equal(rec, want)
