"""pff_accumulate_constant_Q_noalloc (fixed device scratch, no allocation)
against pff_accumulate_constant_Q, through the C ABI.

Two calls in a row on the same device, the second on a smaller grid, so that
the second call runs on the first call's leftovers in the shared array.  Q
must be bit-exact with the allocating entry point (both are f32(q0) times the
exact integer upstream count), and the device-side converged flag must be 1
on these forests.
"""

import ctypes

import numpy as np
import pytest
import torch

from pyfastflow_amd import synth

pytestmark = pytest.mark.gpu


def _routed(pf, ny, nx, seed, pits, mode):
    """Receivers after reroute_flow(carve=True) on a synthetic DEM (a forest
    whose every path reaches an outlet)."""
    z = synth.dem(ny, nx, 2.0, 500.0, seed, n_pits=pits, dx=10.0, periodic_ew=(mode == "periodic_EW"))
    r = pf.flow.FlowRouter(nx, ny, 10.0, boundary_mode=mode)
    r.set_z(z)
    r.compute_receivers()
    r.reroute_flow(carve=True)
    return r


def _noalloc(pf, r, value, area, Q, conv):
    lib = pf._lib
    g = lib.grid(r.nx, r.ny, r.dx, r.mode)
    stream = torch.cuda.current_stream().cuda_stream
    lib.call("pff_accumulate_constant_Q_noalloc", ctypes.c_void_p(stream), ctypes.byref(g),
             lib.ptr(r.receivers), ctypes.c_double(value), ctypes.c_int32(area), lib.ptr(Q), lib.ptr(conv))


@pytest.mark.parametrize("value,area", [(1e-5, 1), (2.5, 0)])
def test_noalloc_matches_allocating_entry(pf, value, area):
    cases = [(1000, 1024, 11, 60, "periodic_EW"), (200, 300, 12, 8, "normal")]
    for ny, nx, seed, pits, mode in cases:
        print(f"case noalloc {ny}x{nx} {mode} value={value} area={area}", flush=True)
        r = _routed(pf, ny, nx, seed, pits, mode)
        assert r.accumulate_constant_Q(value, area=bool(area))
        ref = r.Q.to_numpy().copy()
        Q = torch.full((ny * nx,), float("nan"), dtype=torch.float32, device="cuda")
        conv = torch.full((1,), -7, dtype=torch.int32, device="cuda")
        _noalloc(pf, r, value, area, Q, conv)
        torch.cuda.synchronize()
        assert int(conv.item()) == 1
        np.testing.assert_array_equal(Q.cpu().numpy(), ref)


def test_noalloc_rejects_grid_beyond_fixed_scratch(pf):
    """A 2048 x 2048 grid needs more than the fixed 64 MiB: refused before any
    device work, with the size in the message."""
    lib = pf._lib
    g = lib.grid(2048, 2048, 10.0, 0)
    dummy = ctypes.c_void_p(16)
    rc = lib.load().pff_accumulate_constant_Q_noalloc(None, ctypes.byref(g), dummy, 1.0, 1, dummy, dummy)
    assert rc != 0
    assert b"fixed scratch" in lib.load().pff_last_error()
