"""GPU parity of the flow path (receivers, accumulation, lakeflow) against the
oracle, through the C ABI (pyfastflow_amd -> libpff_hip.so).

Integer outputs must be bit-exact; Q must equal f32(q0) * exact upstream
count, which is within 1e-5 relative of the FP64 oracle.
"""

import numpy as np
import pytest
import torch

from pyfastflow_amd import synth

pytestmark = pytest.mark.gpu

MODES = ["normal", "periodic_EW", "periodic_NS"]

# (ny, nx, mode, n_pits, seed)
SMALL = [
    (2, 2, "normal", 0, 1),
    (5, 7, "normal", 0, 2),
    (7, 5, "periodic_EW", 0, 3),
    (9, 11, "periodic_NS", 0, 4),
    (48, 64, "normal", 3, 5),
    (64, 64, "periodic_EW", 4, 6),
    (65, 33, "periodic_NS", 2, 7),
    (128, 96, "normal", 10, 8),
    (256, 256, "periodic_EW", 40, 9),
]


def _dem(ny, nx, seed, pits, mode):
    return synth.dem(ny, nx, 2.0, 200.0, seed, n_pits=pits, dx=10.0, periodic_ew=(mode == "periodic_EW"))


def _router(pf, nx, ny, dx, mode, z, boundaries=None):
    r = pf.flow.FlowRouter(nx, ny, dx, boundary_mode=mode, boundaries=boundaries)
    r.set_z(z)
    return r


def _ograd(ora, nx, ny, dx, mode, boundaries=None):
    return ora.Grid(nx, ny, dx, mode, boundaries)


@pytest.mark.parametrize("ny,nx,mode,pits,seed", SMALL)
def test_receivers_bit_exact(pf, ora, ny, nx, mode, pits, seed):
    print(f"case receivers {ny}x{nx} {mode}", flush=True)
    z = _dem(ny, nx, seed, pits, mode)
    r = _router(pf, nx, ny, 10.0, mode, z)
    r.compute_receivers()
    g = _ograd(ora, nx, ny, 10.0, mode)
    rcv, grad = ora.compute_receivers(g, z)
    np.testing.assert_array_equal(r.receivers.to_numpy(), rcv)
    np.testing.assert_array_equal(r.gradient.to_numpy().view(np.uint32), grad.view(np.uint32))


def test_receivers_ties_and_flats(pf, ora):
    """Flat areas drain to themselves; equal drops pick the first direction."""
    z = np.zeros((6, 8), np.float32)
    z[2:4, 3:5] = 1.0
    r = _router(pf, 8, 6, 1.0, "normal", z)
    r.compute_receivers()
    rcv, _ = ora.compute_receivers(_ograd(ora, 8, 6, 1.0, "normal"), z)
    np.testing.assert_array_equal(r.receivers.to_numpy(), rcv)


def test_receivers_custom_boundaries(pf, ora):
    ny, nx = 32, 40
    z = _dem(ny, nx, 11, 2, "normal")
    b = np.ones((ny, nx), np.uint8)
    b[0, :] = 3
    b[-1, :] = 3
    b[:, 0] = 9
    b[:, -1] = 9
    b[10:14, 10:14] = 0
    b[20, 5] = 7
    r = _router(pf, nx, ny, 10.0, "custom", z, boundaries=b)
    r.compute_receivers()
    rcv, grad = ora.compute_receivers(_ograd(ora, nx, ny, 10.0, "custom", b), z)
    np.testing.assert_array_equal(r.receivers.to_numpy(), rcv)
    np.testing.assert_array_equal(r.gradient.to_numpy(), grad)


@pytest.mark.parametrize("ny,nx,mode,pits,seed", SMALL)
def test_accumulate_exact(pf, ora, ny, nx, mode, pits, seed):
    print(f"case accumulate {ny}x{nx} {mode}", flush=True)
    z = _dem(ny, nx, seed, pits, mode)
    r = _router(pf, nx, ny, 10.0, mode, z)
    r.compute_receivers()
    assert r.accumulate_constant_Q(1e-5, area=True)
    g = _ograd(ora, nx, ny, 10.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    cnt = ora.upstream_count(g, rcv)
    q0 = np.float32(1e-5 * 10.0 * 10.0)
    expect = (q0 * cnt.astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(r.Q.to_numpy(), expect)
    ref64 = ora.accumulate_f64(g, rcv, np.full(g.N, float(q0)))
    rel = np.abs(r.Q.to_numpy() - ref64) / np.abs(ref64)
    assert rel.max() <= 1e-5


def test_accumulate_unit_area_matches_reference_schedule(pf, ora):
    """The 512^2 plain workload: Q (dx=1, value 1) is integer-valued and must
    equal the reference's own rake-and-compress schedule exactly."""
    c = synth.plain_small()
    r = _router(pf, c["nx"], c["ny"], c["dx"], "normal", c["z"])
    r.compute_receivers()
    assert r.accumulate_constant_Q(1.0, area=True)
    g = _ograd(ora, c["nx"], c["ny"], c["dx"], "normal")
    rcv, _ = ora.compute_receivers(g, c["z"])
    Qref, done = ora.accumulate_constant_Q(g, rcv, 1.0, True)
    assert done
    np.testing.assert_array_equal(r.Q.to_numpy(), Qref)
    # the GPU run of the reference schedule agrees as well
    r.accumulate_constant_Q_rake_compress(1.0, area=True)
    np.testing.assert_array_equal(r.Q.to_numpy(), Qref)


@pytest.mark.parametrize("ny,nx,mode,pits,seed", SMALL)
@pytest.mark.parametrize("carve", [True, False])
def test_reroute_flow_bit_exact(pf, ora, ny, nx, mode, pits, seed, carve):
    print(f"case reroute {ny}x{nx} {mode} carve={carve}", flush=True)
    z = _dem(ny, nx, seed, pits, mode)
    r = _router(pf, nx, ny, 10.0, mode, z)
    r.compute_receivers()
    bf = torch.zeros(nx * ny, dtype=torch.int32, device="cuda")
    st = r.reroute_flow(carve=carve, bid_first=bf)
    g = _ograd(ora, nx, ny, 10.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    ref = ora.reroute_flow(g, z, rcv, carve=carve)
    assert st.ndep0 == ref["ndep0"]
    assert st.n_iter == ref["n_iter"]
    np.testing.assert_array_equal(r.receivers.to_numpy(), ref["rec"])
    np.testing.assert_array_equal(r.receivers_.to_numpy(), ref["rec_"])
    if ref["ndep0"] > 0:
        np.testing.assert_array_equal(bf.cpu().numpy(), ref["bid_first"])


def test_saddlesort_module_bit_exact(pf, ora):
    """Module-level saddlesort on first-pass labels: z', borders, saddle keys,
    saddle nodes and outlets per basin."""
    ny, nx, mode = 96, 80, "periodic_EW"
    z = _dem(ny, nx, 21, 8, mode)
    r = _router(pf, nx, ny, 10.0, mode, z)
    r.compute_receivers()
    pf.flow.basin_identification(r.bid, r.receivers, r.receivers_, nx * ny)
    pf.flow.saddlesort(r.bid, r.is_border, r.z_, r.basin_saddle, r.basin_saddlenode, r.outlet, r.z)
    g = _ograd(ora, nx, ny, 10.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    bid = ora.basin_identification(g, rcv)
    np.testing.assert_array_equal(r.bid.to_numpy(), bid)
    s = ora.saddlesort(g, bid, z)
    np.testing.assert_array_equal(r.z_.to_numpy(), s["z_prime"])
    np.testing.assert_array_equal(r.is_border.to_numpy(), s["is_border"])
    np.testing.assert_array_equal(r.basin_saddle.to_numpy(), s["basin_saddle"])
    np.testing.assert_array_equal(r.basin_saddlenode.to_numpy(), s["saddlenode"])
    np.testing.assert_array_equal(r.outlet.to_numpy(), s["outlet"])


def test_depressions_1024_full_pipeline(pf, ora):
    """The depressions workload at 1024^2 (periodic east-west, pits):
    receivers, lakeflow and Q against the oracle."""
    c = synth.depressions(n=1024, n_pits=125)
    nx, ny, dx = c["nx"], c["ny"], c["dx"]
    r = _router(pf, nx, ny, dx, c["boundary_mode"], c["z"])
    r.compute_receivers()
    st = r.reroute_flow()
    assert r.accumulate_constant_Q(c["precipitation"], area=True)
    g = _ograd(ora, nx, ny, dx, c["boundary_mode"])
    rcv, _ = ora.compute_receivers(g, c["z"])
    ref = ora.reroute_flow(g, c["z"], rcv)
    assert st.ndep0 == ref["ndep0"] and st.n_iter == ref["n_iter"]
    np.testing.assert_array_equal(r.receivers.to_numpy(), ref["rec"])
    assert ora.depression_counter(g, ref["rec"]) == 0
    cnt = ora.upstream_count(g, ref["rec"])
    q0 = np.float32(c["precipitation"] * dx * dx)
    np.testing.assert_array_equal(r.Q.to_numpy(), q0 * cnt.astype(np.float32))


def test_pack_unpack_gpu(pf, ora):
    rng = np.random.default_rng(42)
    f = rng.uniform(-1e6, 1e6, 10000).astype(np.float32)
    f[:4] = [0.0, -0.0, np.inf, -np.inf]
    idx = rng.integers(0, 750000, 10000).astype(np.int32)
    tf = torch.from_numpy(f).cuda()
    ti = torch.from_numpy(idx).cuda()
    keys = torch.zeros(10000, dtype=torch.int64, device="cuda")
    pf.flow.pack_full_float_index(keys, tf, ti)
    np.testing.assert_array_equal(keys.cpu().numpy(), ora.pack_array(f, idx))
    f2 = torch.zeros_like(tf)
    i2 = torch.zeros_like(ti)
    pf.flow.unpack_full_float_index(keys, f2, i2)
    np.testing.assert_array_equal(f2.cpu().numpy().view(np.uint32), f.view(np.uint32))
    np.testing.assert_array_equal(i2.cpu().numpy(), idx)


@pytest.mark.parametrize("mode", MODES)
def test_edge_maps(pf, ora, mode):
    ny, nx = 6, 9
    pf.flow.FlowRouter(nx, ny, 1.0, boundary_mode=mode)
    e = torch.zeros(nx * ny, dtype=torch.uint8, device="cuda")
    o = torch.zeros_like(e)
    pf.flow.fill_edges(e)
    pf.flow.flow_out_nodes(o)
    g = _ograd(ora, nx, ny, 1.0, mode)
    np.testing.assert_array_equal(e.cpu().numpy(), [ora.which_edge(g, i) for i in range(g.N)])
    np.testing.assert_array_equal(o.cpu().numpy(), [int(ora.can_leave(g, i)) for i in range(g.N)])


@pytest.mark.parametrize("seed", [68, 132, 149, 447, 113, 157, 382, 1533])
def test_reroute_sentinel_and_capped_dems(pf, ora, seed):
    """DEMs on which the cycle pass reads the sentinel (node 42), or the trip
    loop stops at its cap with a depression left: receivers bit-exact."""
    from oracle_util import sentinel_dem

    print(f"case reroute sentinel/capped seed {seed}", flush=True)
    z = sentinel_dem(seed)
    ny, nx = z.shape
    r = _router(pf, nx, ny, 10.0, "normal", z)
    r.compute_receivers()
    r.reroute_flow(carve=True)
    g = _ograd(ora, nx, ny, 10.0, 0)
    rcv, _ = ora.compute_receivers(g, z)
    lk = ora.reroute_flow(g, z, rcv, carve=True)
    np.testing.assert_array_equal(r.receivers.to_numpy(), lk["rec"])
    st = r.last_lake_status
    assert [int(st.ndep[i]) for i in range(st.n_loop)] == lk["ndep"]


INVALID_KEY = 0x4CBEBC200000002A


@pytest.mark.parametrize("case", [
    (64, 64, "periodic_EW", 4, 6),
    (128, 96, "normal", 10, 8),
    (256, 256, "periodic_EW", 40, 9),
    (200, 300, "periodic_NS", 30, 12),
    "sentinel68", "sentinel447", "capped157",
])
def test_lake_outlets_per_trip(pf, ora, case):
    """Lake outlets bit-exact per trip: after each trip's saddle and cycle
    passes, every basin's packed outlet key (elevation, node) and saddle
    node equal the oracle's (lakeflow.py:105-281, 8.P)."""
    from oracle_util import sentinel_dem

    if isinstance(case, str):
        seed = int("".join(ch for ch in case if ch.isdigit()))
        z = sentinel_dem(seed)
        ny, nx = z.shape
        mode = "normal"
    else:
        ny, nx, mode, pits, seed = case
        z = _dem(ny, nx, seed, pits, mode)
    print(f"case lake outlets {case}", flush=True)
    r = _router(pf, nx, ny, 10.0, mode, z)
    r.compute_receivers()
    r.reroute_flow(carve=True, log_outlets=True)
    g = _ograd(ora, nx, ny, 10.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    lk = ora.reroute_flow(g, z, rcv, carve=True, log_iters=64)
    trips = r.last_lake_outlets
    assert len(trips) == lk["n_iter"]
    for t, (b, o, sn) in enumerate(trips):
        ol, snl = lk["outlet_log"][t], lk["saddlenode_log"][t]
        np.testing.assert_array_equal(o, ol[b], err_msg=f"trip {t}: outlet keys")
        np.testing.assert_array_equal(sn, snl[b], err_msg=f"trip {t}: saddle nodes")
        rest = np.ones(ol.size, bool)
        rest[b] = False
        assert (ol[rest] == INVALID_KEY).all() and (snl[rest] == -1).all(), f"trip {t}: basins missing from the log"


def _snake_dem(n=96):
    """A channel winding through the grid (channel rows 1, 3, 5, ... joined
    at alternating ends, walls between them), descending 0.01 m per cell to a
    pit at its end; a low border cell next to the channel's start.  The pit's
    basin is the whole channel: its carve path runs back up thousands of
    cells, and inside each 64 x 64 tile the walks along it are thousands of
    steps long."""
    z = np.full((n, n), 500.0)
    order = []
    rows = list(range(1, n - 1, 2))
    for k, r in enumerate(rows):
        cols = list(range(1, n - 1)) if k % 2 == 0 else list(range(n - 2, 0, -1))
        order += [(r, c) for c in cols]
        if k + 1 < len(rows):
            order.append((r + 1, cols[-1]))
    L = len(order)
    for s, (r, c) in enumerate(order):
        z[r, c] = 100.0 + 0.01 * (L - s)
    z[order[-1]] = 10.0
    z[0, 2] = 5.0
    return z.astype(np.float32)


@pytest.mark.parametrize("n,mode", [(96, "normal"), (160, "normal"), (130, "periodic_EW")])
def test_reroute_long_in_tile_paths(pf, ora, n, mode):
    """Carve along a winding channel: in-tile walks far longer than the
    walk cap (the tag walks then find local ends by pointer jumping in LDS).
    Receivers, first-trip labels and per-trip outlets bit-exact."""
    print(f"case snake {n} {mode}", flush=True)
    z = _snake_dem(n)
    r = _router(pf, n, n, 10.0, mode, z)
    r.compute_receivers()
    bf = torch.empty(n * n, dtype=torch.int32, device="cuda")
    r.reroute_flow(carve=True, bid_first=bf, log_outlets=True)
    g = _ograd(ora, n, n, 10.0, mode)
    rcv, _ = ora.compute_receivers(g, z)
    lk = ora.reroute_flow(g, z, rcv, carve=True, log_iters=16)
    np.testing.assert_array_equal(bf.cpu().numpy(), lk["bid_first"])
    np.testing.assert_array_equal(r.receivers.to_numpy(), lk["rec"])
    assert ora.depression_counter(g, lk["rec"]) == 0
    for t, (b, o, sn) in enumerate(r.last_lake_outlets):
        np.testing.assert_array_equal(o, lk["outlet_log"][t][b])
        np.testing.assert_array_equal(sn, lk["saddlenode_log"][t][b])
