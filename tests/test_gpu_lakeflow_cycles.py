"""GPU parity of reroute_flow on receivers that hold cycles (the carve can
leave them: only 2-cycles of basins are removed).  The trips then follow the
synchronous-pass semantics, which the library computes on the cycle cells
alone and with the tiled contraction elsewhere; the results must be the
oracle's bit for bit: final receivers, first-trip labels and the per-trip
depression counts."""

import numpy as np
import pytest
import torch

from cycle_graphs import cycle_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [5, 6])
def test_reroute_with_receiver_cycles_matches_oracle(pf, ora, seed):
    ny, nx, dx = 200, 300, 10.0
    z, rcv, _ = cycle_case(ny, nx, seed, dx)
    print(f"case cycles seed {seed}", flush=True)
    g = ora.Grid(nx, ny, dx, 0)
    want = ora.reroute_flow(g, z, rcv, carve=True)
    r = pf.flow.FlowRouter(nx, ny, dx, boundary_mode="normal")
    r.set_z(z.reshape(ny, nx))
    r.receivers.from_numpy(rcv)
    bf = torch.empty(nx * ny, dtype=torch.int32, device="cuda")
    r.reroute_flow(carve=True, bid_first=bf)
    np.testing.assert_array_equal(bf.cpu().numpy(), want["bid_first"])
    np.testing.assert_array_equal(r.receivers.to_numpy(), want["rec"])
    st = r.last_lake_status
    assert [int(st.ndep[i]) for i in range(st.n_loop)] == want["ndep"]


def test_reroute_cycles_with_stale_buffers(pf, ora):
    """The same after a larger grid has run on the same context and with the
    router's work arrays holding junk."""
    big = pf.flow.FlowRouter(700, 650, 10.0, boundary_mode="normal")
    from pyfastflow_amd import synth

    big.set_z(synth.dem(650, 700, 2.0, 80.0, 3, n_pits=300, dx=10.0))
    big.compute_receivers()
    big.reroute_flow(carve=True)
    ny, nx, dx = 200, 300, 10.0
    z, rcv, _ = cycle_case(ny, nx, 7, dx)
    g = ora.Grid(nx, ny, dx, 0)
    want = ora.reroute_flow(g, z, rcv, carve=True)
    r = pf.flow.FlowRouter(nx, ny, dx, boundary_mode="normal")
    r.set_z(z.reshape(ny, nx))
    r.receivers.from_numpy(rcv)
    gen = torch.Generator().manual_seed(3)
    for f in (r.bid, r.receivers_, r.receivers__, r.z_, r.is_border, r.outlet, r.basin_saddle, r.basin_saddlenode):
        d = f.data
        if d.dtype == torch.float32:
            d.copy_(torch.rand(d.shape, generator=gen) * 100.0)
        elif d.dtype == torch.uint8:
            d.fill_(1)
        else:
            d.copy_(torch.randint(-1, nx * ny, d.shape, generator=gen).to(d.dtype))
    r.reroute_flow(carve=True)
    np.testing.assert_array_equal(r.receivers.to_numpy(), want["rec"])
