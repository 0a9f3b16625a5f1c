"""Host-side logic, without a GPU.

- No CPU fallback: every compute path raises without a ROCm device, refuses
  host tensors, and loading the binding without the built library raises.
- The Field shim behaves as the Taichi field methods the reference's callers
  use (fill, copy_from, from_numpy, to_numpy, field[None]).
- The stochastic receivers' per-call seed is splitmix64 (published vector).
- The synthetic workloads are deterministic and have the benchmark's shapes.
- A GraphFlood state saved to .npz loads back bit for bit, on the same grid only.
- bench.py refuses more than one GPU; smoke() refuses to run without one.
"""

import os

import numpy as np
import pytest
import torch

no_gpu = pytest.mark.skipif(torch.cuda.is_available(), reason="checks the behaviour without a GPU")


# ---- no CPU fallback --------------------------------------------------------

@no_gpu
def test_context_needs_a_gpu():
    from pyfastflow_amd import _lib

    with pytest.raises(RuntimeError, match="ROCm GPU"):
        _lib.ctx_for()


@no_gpu
def test_router_and_batch_need_a_gpu():
    import pyfastflow_amd as pf

    with pytest.raises(RuntimeError):
        pf.flow.FlowRouter(16, 16, 1.0)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        pf.flow.route_batch([np.zeros((8, 8), np.float32)])


def test_module_functions_refuse_host_tensors(monkeypatch):
    import pyfastflow_amd as pf
    from pyfastflow_amd import constants as cte

    monkeypatch.setattr(cte, "NX", 8)
    monkeypatch.setattr(cte, "NY", 8)
    monkeypatch.setattr(cte, "BOUND_MODE", 0)
    z = torch.zeros(64, dtype=torch.float32)
    rcv = torch.zeros(64, dtype=torch.int32)
    with pytest.raises(ValueError, match="GPU"):
        pf.flow.compute_receivers(z, rcv, z.clone())


def test_missing_library_raises(monkeypatch, tmp_path):
    from pyfastflow_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libpff_hip.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError, match="no CPU fallback"):
        _lib.load()


def test_custom_boundaries_checked_before_upload(monkeypatch):
    from pyfastflow_amd import constants as cte

    monkeypatch.setattr(cte, "NX", 8)
    monkeypatch.setattr(cte, "NY", 6)
    monkeypatch.setattr(cte, "BOUND_MODE", 0)
    with pytest.raises(ValueError, match="NX\\*NY"):
        cte.init_custom_boundaries(np.ones((6, 7), np.uint8))
    assert cte.BOUND_MODE == 0


# ---- Field shim ---------------------------------------------------------------

def test_field_methods_on_host_storage():
    from pyfastflow_amd.field import Field, f32, i32

    a = Field(i32, (2, 3), device="cpu")
    assert a.shape == (2, 3) and a.dtype == torch.int32 and len(a) == 6
    a.fill(7)
    assert (a.to_numpy() == 7).all()
    a.from_numpy(np.arange(6, dtype=np.int64))  # reshaped and cast like Taichi's from_numpy
    assert a.to_numpy().tolist() == [[0, 1, 2], [3, 4, 5]]
    b = Field(i32, (2, 3), device="cpu")
    b.copy_from(a)
    assert b[1, 2] == 5
    b.copy_from(torch.full((2, 3), -1, dtype=torch.int32))
    assert (b.to_numpy() == -1).all()
    s = Field(f32, (), device="cpu")
    s[None] = 2.5
    assert s[None] == 2.5
    a[0, 0] = 9
    assert a.to_torch()[0, 0].item() == 9


# ---- stochastic receivers' seeds -------------------------------------------------

def test_rand_seed_sequence_is_splitmix64(monkeypatch):
    """Call k uses splitmix64's output function on RAND_SEED + k * golden
    ratio; from seed 0 its second call is the first published splitmix64
    output, 0xE220A8397B1DCDAF (Vigna's reference generator)."""
    from pyfastflow_amd import constants as cte

    monkeypatch.setattr(cte, "RAND_SEED", 0)
    monkeypatch.setattr(cte, "_rand_calls", 0)
    first = cte.next_rand_seed()
    assert first == 0  # the output function maps 0 to 0
    assert cte.next_rand_seed() == 0xE220A8397B1DCDAF
    assert cte.next_rand_seed() == 0x6E789E6AA1B965F4
    monkeypatch.setattr(cte, "_rand_calls", 1)
    assert cte.next_rand_seed() == 0xE220A8397B1DCDAF  # replayable from the call count


# ---- synthetic workloads ------------------------------------------------------------

def test_synthetic_dem_is_deterministic():
    from pyfastflow_amd import synth

    a = synth.depressions(n=128, seed=7, n_pits=12)
    b = synth.depressions(n=128, seed=7, n_pits=12)
    c = synth.depressions(n=128, seed=8, n_pits=12)
    assert a["z"].dtype == np.float32 and a["z"].shape == (128, 128)
    assert np.array_equal(a["z"], b["z"])
    assert not np.array_equal(a["z"], c["z"])
    assert a["boundary_mode"] == "periodic_EW" and a["dx"] == 10.0


def test_pits_lower_the_surface():
    from pyfastflow_amd import synth

    for seed in range(3, 8):
        flat = synth.dem(96, 96, 2.0, 100.0, seed=seed)
        pitted = synth.dem(96, 96, 2.0, 100.0, seed=seed, n_pits=1)
        d = flat.astype(np.float64) - pitted
        assert d.min() >= -1e-4  # a pit only digs
        # depth U[5, 50] m; the deepest cell centre lies within half a cell
        # diagonal of the pit centre, where the Gaussian (sigma = R/2 >= 2.5
        # cells) keeps more than 0.96 of the depth
        assert 0.96 * 5.0 <= d.max() <= 50.0 + 1e-3


def test_ragged_batch_workload_shape():
    """The 256-tile batch: square tiles with sides in [128, 1024], 66 888 835
    cells in all (the figure the benchmark routes per step)."""
    from pyfastflow_amd import synth

    tiles = synth.ragged_batch(256)
    sides = [t["nx"] for t in tiles]
    assert len(tiles) == 256
    assert all(t["nx"] == t["ny"] and t["z"].shape == (t["ny"], t["nx"]) for t in tiles)
    assert 128 <= min(sides) and max(sides) <= 1024
    assert sum(s * s for s in sides) == 66_888_835


# ---- GraphFlood checkpoint ----------------------------------------------------------

class _HostRouter:
    """The router fields a Flooder's state lives in, on the CPU."""

    def __init__(self, nx, ny, dx):
        from pyfastflow_amd.field import Field, f32

        self.nx, self.ny, self.dx = nx, ny, dx
        self.rshp = (ny, nx)
        self.device = torch.device("cpu")
        self.z = Field(f32, (nx * ny,), device="cpu")
        self.Q = Field(f32, (nx * ny,), device="cpu")


def _flooder(monkeypatch, nx, ny, dx, **kw):
    from pyfastflow_amd import constants as cte
    from pyfastflow_amd.graphflood import Flooder

    for k in ("PREC", "MANNING", "EDGESW"):  # the constructor sets these globals
        monkeypatch.setattr(cte, k, getattr(cte, k))
    return Flooder(_HostRouter(nx, ny, dx), **kw)


def test_graphflood_state_round_trip(monkeypatch, tmp_path):
    rng = np.random.Generator(np.random.PCG64(5))
    a = _flooder(monkeypatch, 12, 7, 10.0, precipitation_rates=2e-5, manning=0.05, update="single")
    zh, h, Q = (rng.uniform(0, 100, 84).astype(np.float32) for _ in range(3))
    a.router.z.from_numpy(zh)
    a.set_h(h)
    a.router.Q.from_numpy(Q)
    a.iterations = 1234
    path = tmp_path / "state.npz"
    a.save_state(path)

    b = _flooder(monkeypatch, 12, 7, 10.0)  # other parameters: kept, the saved ones returned
    saved = b.load_state(path)
    assert np.array_equal(b.router.z.to_numpy(), zh)
    assert np.array_equal(b.get_h(), h.reshape(7, 12))
    assert np.array_equal(b.router.Q.to_numpy(), Q)
    assert b.iterations == 1234
    assert saved == {"precipitation_rates": 2e-5, "manning": 0.05, "edge_slope": 1e-2, "dt": a.dt,
                     "update": "single"}
    assert b.manning == 0.033 and b.update == "reference"


def test_graphflood_state_refuses_another_grid(monkeypatch, tmp_path):
    a = _flooder(monkeypatch, 12, 7, 10.0)
    path = tmp_path / "state.npz"
    a.save_state(path)
    for nx, ny, dx in ((7, 12, 10.0), (12, 7, 5.0)):
        with pytest.raises(ValueError, match="cannot be loaded"):
            _flooder(monkeypatch, nx, ny, dx).load_state(path)


# ---- entry points of the repository -------------------------------------------------

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys

    out = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=ROOT, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode != 0
    assert "one GPU" in out.stdout + out.stderr


@no_gpu
def test_smoke_needs_a_gpu():
    import __graft_entry__ as g

    with pytest.raises(AssertionError, match="needs a GPU"):
        g.smoke()
