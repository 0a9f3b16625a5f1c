"""The committed routing fixtures (tests/golden/routing_cases.npz): the DEM
synthesiser still makes the stored DEMs, the oracle still gives the stored
outputs, and (GPU) the HIP path gives them bit for bit."""

import numpy as np
import pytest

from golden_cases import load

CASES = load()
INVALID_KEY = 0x4CBEBC200000002A


@pytest.mark.parametrize("name", sorted(CASES))
def test_synthesiser_makes_the_stored_dem(name):
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from gen_fixtures import case_dem

    z, mode, dx = case_dem(name)
    c = CASES[name]
    np.testing.assert_array_equal(z, c["z"])
    assert mode == c["mode_name"] and dx == c["dx"]


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(ora, name):
    c = CASES[name]
    z = c["z"]
    g = ora.Grid(c["nx"], c["ny"], c["dx"], c["mode"])
    rcv, grad = ora.compute_receivers(g, z)
    np.testing.assert_array_equal(rcv, c["rcv"])
    np.testing.assert_array_equal(grad.view(np.uint32), c["gradient"].view(np.uint32))
    lk = ora.reroute_flow(g, z, rcv, carve=True, log_iters=64)
    np.testing.assert_array_equal(lk["rec"], c["rec"])
    np.testing.assert_array_equal(lk["bid_first"], c["bid_first"])
    assert lk["ndep"] == list(c["ndep"])
    if c["upstream"].size == 0:
        # the carve left a receiver cycle (only 2-cycles are removed)
        with pytest.raises(ValueError):
            ora.upstream_count(g, lk["rec"])
        return
    np.testing.assert_array_equal(ora.upstream_count(g, lk["rec"]), c["upstream"])
    zh, h = ora.fill_z_add_delta(g, z, np.zeros(z.size, np.float32), lk["rec"], 1e-3)
    np.testing.assert_array_equal(zh, c["fill_zh"])
    np.testing.assert_array_equal(h, c["fill_h"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_fixture(pf, name):
    import torch

    c = CASES[name]
    print(f"case golden {name}", flush=True)
    r = pf.flow.FlowRouter(c["nx"], c["ny"], c["dx"], boundary_mode=c["mode_name"])
    r.set_z(c["z"])
    r.compute_receivers()
    np.testing.assert_array_equal(r.receivers.to_numpy(), c["rcv"])
    np.testing.assert_array_equal(r.gradient.to_numpy().view(np.uint32), c["gradient"].view(np.uint32))
    bf = torch.empty(c["nx"] * c["ny"], dtype=torch.int32, device="cuda")
    r.reroute_flow(carve=True, bid_first=bf, log_outlets=True)
    np.testing.assert_array_equal(r.receivers.to_numpy(), c["rec"])
    np.testing.assert_array_equal(bf.cpu().numpy(), c["bid_first"])
    st = r.last_lake_status
    assert [int(st.ndep[i]) for i in range(st.n_loop)] == list(c["ndep"])
    assert len(r.last_lake_outlets) == len(c["trips"])
    for t, ((b, o, s), (fb, fo, fs)) in enumerate(zip(r.last_lake_outlets, c["trips"])):
        keep = (o != INVALID_KEY) | (s != -1)
        order = np.argsort(b[keep])
        np.testing.assert_array_equal(b[keep][order], fb, err_msg=f"trip {t} basins")
        np.testing.assert_array_equal(o[keep][order], fo, err_msg=f"trip {t} outlets")
        np.testing.assert_array_equal(s[keep][order], fs, err_msg=f"trip {t} saddle nodes")
    if c["upstream"].size == 0:
        return  # receivers with a cycle: no upstream counts
    cnt = torch.empty(c["nx"] * c["ny"], dtype=torch.int32, device="cuda")
    pf.flow.upstream_count(r.receivers, cnt)
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), c["upstream"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_ignores_stale_work_buffers(pf, name):
    """The lakeflow work arrays (labels, z', border flags, saddles, outlets,
    the receiver copies) may hold anything on entry, as they do when a
    buffer is reused: the result is the fixture's all the same."""
    import torch

    c = CASES[name]
    print(f"case golden {name} with junk work buffers", flush=True)
    r = pf.flow.FlowRouter(c["nx"], c["ny"], c["dx"], boundary_mode=c["mode_name"])
    r.set_z(c["z"])
    r.compute_receivers()
    n = c["nx"] * c["ny"]
    gen = torch.Generator().manual_seed(11)
    for f in (r.bid, r.receivers_, r.receivers__, r.z_, r.is_border, r.outlet, r.basin_saddle, r.basin_saddlenode):
        d = f.data
        if d.dtype == torch.float32:
            d.copy_(torch.rand(d.shape, generator=gen) * 100.0)
        elif d.dtype == torch.uint8:
            d.fill_(1)
        else:
            d.copy_(torch.randint(-1, n, d.shape, generator=gen).to(d.dtype))
    r.reroute_flow(carve=True)
    np.testing.assert_array_equal(r.receivers.to_numpy(), c["rec"])
