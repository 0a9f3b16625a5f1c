"""GPU parity of the fill and GraphFlood kernels against the oracle.

The fill is the same f32 arithmetic in the same order as the oracle, so it
must be bit-exact.  Diffusion and the Manning update are checked against the
FP64 restatement from the same f32 inputs, within 1e-5 relative or 1e-4
absolute (m for h, m^3/s for Q).
"""

import numpy as np
import pytest
import torch

from pyfastflow_amd import synth

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-4


def _close(got, ref, what):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    err = np.abs(got - ref)
    ok = (err <= ATOL) | (err <= RTOL * np.abs(ref))
    worst = int(np.argmax(np.where(ok, 0, err)))
    assert ok.all(), f"{what}: {(~ok).sum()} cells out of tolerance, worst cell {worst}: {got[worst]} vs {ref[worst]}"
    return err.max()


def _setup(pf, ny, nx, mode, seed, pits):
    z = synth.dem(ny, nx, 2.2, 100.0, seed, n_pits=pits, dx=10.0, periodic_ew=(mode == "periodic_EW"))
    r = pf.flow.FlowRouter(nx, ny, 10.0, boundary_mode=mode)
    r.set_z(z)
    return r, z


@pytest.mark.parametrize("mode", ["normal", "periodic_EW", "periodic_NS"])
def test_fill_z_add_delta_bit_exact(pf, ora, mode):
    ny, nx = 96, 128
    r, z = _setup(pf, ny, nx, mode, 31, 6)
    r.compute_receivers()
    r.reroute_flow()
    h0 = np.random.default_rng(0).uniform(0, 0.5, nx * ny).astype(np.float32)
    h = pf.field.Field(torch.float32, (nx * ny,))
    h.from_numpy(h0)
    pf.flow.fill_z_add_delta(r.z, h, r.z_, r.receivers, r.receivers_, r.receivers__, epsilon=1e-3)
    g = ora.Grid(nx, ny, 10.0, mode)
    rec = r.receivers.to_numpy()
    zh_ref, h_ref = ora.fill_z_add_delta(g, z, h0, rec, 1e-3)
    np.testing.assert_array_equal(r.z.to_numpy(), zh_ref)
    np.testing.assert_array_equal(h.to_numpy(), h_ref)


def test_topofill_bit_exact(pf, ora):
    ny, nx = 80, 72
    r, z = _setup(pf, ny, nx, "normal", 32, 5)
    r.compute_receivers()
    r.reroute_flow()
    rec = r.receivers.to_numpy()
    r.fill_z(epsilon=1e-4)
    zf = ora.topofill(ora.Grid(nx, ny, 10.0, 0), z, rec, 1e-4)
    np.testing.assert_array_equal(r.z.to_numpy(), zf)


@pytest.mark.parametrize("mode", ["normal", "periodic_EW", "periodic_NS"])
def test_diffuse_vs_f64(pf, ora, mode):
    ny, nx = 64, 96
    r, z = _setup(pf, ny, nx, mode, 33, 4)
    rng = np.random.default_rng(1)
    Q0 = rng.uniform(0, 5.0, nx * ny).astype(np.float32)
    r.Q.from_numpy(Q0)
    pf.constants.PREC = 1e-5
    pf.graphflood.diffuse_Q_constant_prec(r.z, r.Q, r.Q_)
    g = ora.Grid(nx, ny, 10.0, mode)
    hy = ora.hydro(prec=1e-5)
    ref = ora.diffuse(g, hy, z, Q0, f64=True)
    _close(r.Q.to_numpy(), ref, "diffuse")


@pytest.mark.parametrize("update", [0, 1])
def test_manning_vs_f64(pf, ora, update):
    ny, nx = 64, 64
    r, z = _setup(pf, ny, nx, "normal", 34, 3)
    r.compute_receivers()
    rec = r.receivers.to_numpy()
    rng = np.random.default_rng(2)
    h0 = rng.uniform(0, 2.0, nx * ny).astype(np.float32)
    Q0 = rng.uniform(0, 50.0, nx * ny).astype(np.float32)
    h = pf.field.Field(torch.float32, (nx * ny,))
    dh = pf.field.Field(torch.float32, (nx * ny,))
    h.from_numpy(h0)
    r.Q.from_numpy(Q0)
    pf.constants.MANNING, pf.constants.EDGESW, pf.constants.DT_HYDRO = 0.033, 1e-2, 5e-3
    h0[:100] = 1e-4  # shallow cells that a strong outflow empties
    h.from_numpy(h0)
    if update == 0:
        pf.graphflood.graphflood_core_cte_mannings(h, r.z, dh, r.receivers, r.Q)
    else:
        from pyfastflow_amd import _lib

        hy_g = _lib.hydro(1e-5, 0.033, 1e-2, 5e-3, 1)
        _lib.call("pff_graphflood_core_cte_mannings", _lib.ctx_for(), r._grid(), hy_g, _lib.ptr(h), _lib.ptr(r.z),
                  _lib.ptr(dh), _lib.ptr(r.receivers), _lib.ptr(r.Q))
        assert (h.to_numpy() >= 0).all()
    g = ora.Grid(nx, ny, 10.0, 0)
    hy = ora.hydro(manning=0.033, edge_slope=1e-2, dt=5e-3, update=update)
    h_ref, zh_ref = ora.manning(g, hy, h0, z, rec, Q0, f64=True)
    _close(h.to_numpy(), h_ref, "h")
    _close(r.z.to_numpy(), zh_ref, "zh")


@pytest.mark.parametrize("update", ["reference", "single"])
def test_graphflood_iterations_vs_oracle(pf, ora, update):
    """Five Flooder iterations from a cold start against the f32 oracle run
    from the same state each iteration: receivers and the filled surface are
    bit-exact, h and Q within tolerance (the diffusion sums in a different
    order and powf/sqrtf may differ by an ulp).

    The FP64 restatement is compared per kernel (test_diffuse_vs_f64,
    test_manning_vs_f64): across a whole iteration the diffusion weights are
    slopes of an epsilon-filled surface, (zh_i - zh_j)/dx with differences of
    ~1e-3 m on ~100 m elevations, which FP32 and FP64 fills resolve
    differently by construction."""
    ny, nx = 64, 80
    r, z = _setup(pf, ny, nx, "normal", 35, 4)
    fl = pf.graphflood.Flooder(r, precipitation_rates=1e-5, manning=0.033, edge_slope=1e-2, update=update)
    g = ora.Grid(nx, ny, 10.0, 0)
    hy = ora.hydro(prec=1e-5, manning=0.033, edge_slope=1e-2, dt=5e-3, update=0 if update == "reference" else 1)
    zh = z.ravel().copy()
    h = np.zeros(nx * ny, np.float32)
    Q = np.zeros(nx * ny, np.float32)
    for it in range(5):
        fl.step()
        zh_n, h_n, Q_n, rec = ora.graphflood_iteration(g, hy, zh, h, Q)
        np.testing.assert_array_equal(r.receivers.to_numpy(), rec)
        _close(fl.h.to_numpy(), h_n, f"h it{it}")
        _close(r.Q.to_numpy(), Q_n, f"Q it{it}")
        _close(r.z.to_numpy(), zh_n, f"zh it{it}")
        # continue both from the GPU state so the routing stays shared
        zh, h, Q = r.z.to_numpy().copy(), fl.h.to_numpy().copy(), r.Q.to_numpy().copy()


def _err_stats(got, ref):
    """max / p99 absolute error, argmax cell, and the cells outside 1e-5
    relative / 1e-4 absolute, of got against ref (FP64)."""
    e = np.abs(np.asarray(got, np.float64) - ref)
    bad = (e > ATOL) & (e > RTOL * np.abs(ref))
    return dict(max=float(e.max()), p99=float(np.quantile(e, 0.99)), argmax=int(np.argmax(e)),
                out_of_tol=int(bad.sum()))


@pytest.mark.parametrize("update", ["reference", "single"])
def test_graphflood_60_iterations_end_to_end(pf, ora, update):
    """60 iterations of the reference schedule from a cold start on a GraphFlood
    DEM (beta 2.2, 300 m relief, 2 pits, open edges, 96 x 128) on the GPU and
    in the f32 oracle, with no resync: receivers, zh, h and Q bit-exact at
    every 10th iteration (same f32 operations in the same order, no FMA
    contraction, h^(5/3) from exact double operations on both).

    Both runs are also compared with the FP64 restatement (routing on the f32
    rounding of its zh): the error per result (max, p99, argmax cell) is
    printed, and it is the f32 arithmetic's own, equal for GPU and oracle.
    zh and h are stored in f32 at the elevation's magnitude (ulp(300 m) =
    3e-5 m) and every iteration rounds them, so this drift grows past 1e-4 m
    within tens of iterations for the reference's f32 method itself."""
    ny, nx = 96, 128
    z = synth.dem(ny, nx, 2.2, 300.0, 303, n_pits=2, dx=10.0)
    r = pf.flow.FlowRouter(nx, ny, 10.0, boundary_mode="normal")
    r.set_z(z)
    fl = pf.graphflood.Flooder(r, precipitation_rates=1e-5, manning=0.033, edge_slope=1e-2, update=update)
    g = ora.Grid(nx, ny, 10.0, 0)
    hy = ora.hydro(prec=1e-5, manning=0.033, edge_slope=1e-2, dt=5e-3, update=0 if update == "reference" else 1)
    zh, h, Q = z.ravel().copy(), np.zeros(nx * ny, np.float32), np.zeros(nx * ny, np.float32)
    zh64, h64, Q64 = zh.astype(np.float64), np.zeros(nx * ny), np.zeros(nx * ny)
    for blk in range(6):
        fl.run_graphflood(10)
        for _ in range(10):
            zh, h, Q, rec = ora.graphflood_iteration(g, hy, zh, h, Q)
            zh64, h64, Q64, _ = ora.graphflood_iteration(g, hy, zh64, h64, Q64, f64=True)
        it = 10 * (blk + 1)
        print(f"case graphflood e2e {update} iteration {it}", flush=True)
        np.testing.assert_array_equal(r.receivers.to_numpy(), rec)
        np.testing.assert_array_equal(r.z.to_numpy().view(np.uint32), zh.view(np.uint32))
        np.testing.assert_array_equal(fl.h.to_numpy().view(np.uint32), h.view(np.uint32))
        np.testing.assert_array_equal(r.Q.to_numpy().view(np.uint32), Q.view(np.uint32))
        for name, a, b in (("h", h, h64), ("zh", zh, zh64), ("Q", Q, Q64)):
            print(f"  {name} vs FP64: {_err_stats(a, b)}", flush=True)


def test_graphflood_survives_carve_cycles(pf):
    """On the 2048x2048 GraphFlood workload the lakeflow carve of the water
    surface leaves a receiver cycle within the first 20 iterations (only
    2-cycles are removed); the reroute must detect it and take the exact
    path rather than fail."""
    from pyfastflow_amd import synth

    cg = synth.graphflood_plain()
    r = pf.flow.FlowRouter(cg["nx"], cg["ny"], cg["dx"], boundary_mode="normal")
    r.set_z(cg["z"])
    fl = pf.graphflood.Flooder(r, precipitation_rates=cg["precipitation"], manning=cg["manning"], edge_slope=1e-2)
    print("case graphflood 2048^2, 25 iterations", flush=True)
    fl.run_graphflood(25)
    h = fl.get_h()
    assert np.isfinite(h).all()
    assert float(h.max()) > 0.0


@pytest.mark.parametrize("n_diffuse", [1, 3, 5])
def test_fused_hydro_matches_module_calls(pf, n_diffuse):
    """run_graphflood on a 'normal' grid fuses the fill's apply, the
    diffusion passes and Manning into one tiled kernel (with the fill's
    passes stopping early once they change nothing); step_modules runs the
    reference's sequence of module-level kernels.  Both give the same bits
    over 8 iterations on a 300 x 520 grid (tiles cut by the grid edges)."""
    ny, nx = 300, 520
    z = synth.dem(ny, nx, 2.2, 300.0, 48, n_pits=6, dx=10.0)
    out = []
    for fused in (True, False):
        r = pf.flow.FlowRouter(nx, ny, 10.0, boundary_mode="normal")
        r.set_z(z)
        fl = pf.graphflood.Flooder(r, precipitation_rates=1e-5, manning=0.033, edge_slope=1e-2)
        for _ in range(8):
            if fused:
                fl.run_graphflood(1, n_diffuse=n_diffuse)
            else:
                fl.step_modules(n_diffuse=n_diffuse)
        out.append((r.z.to_numpy().copy(), fl.h.to_numpy().copy(), r.Q.to_numpy().copy()))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
