"""What the trained surrogate gains over the plain solver: GraphFlood from the
network's warm start against the cold start on a DEM it was not trained on.

Uses the committed weights (pyfastflow_amd/surrogate/unet_warmstart.pt, made
by python -m pyfastflow_amd.surrogate.experiment).  Both starts run the same
convergence test (at most half of the cells change by more than 1e-6 m over
100 iterations).
"""

import os

import numpy as np
import pytest
import torch

from pyfastflow_amd import synth

pytestmark = pytest.mark.gpu


def test_warm_start_gains_over_cold(pf):
    from pyfastflow_amd.surrogate import UNet
    from pyfastflow_amd.surrogate.warmstart import WEIGHTS, load_weights, solve

    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights committed")
    model = load_weights(UNet(), WEIGHTS).cuda().eval()
    n = 512
    z = synth.dem(n, n, 2.0, 800.0, 4242, n_pits=int(round(2000 * n * n / 4096.0**2)), dx=10.0)
    print("case warm start 512^2", flush=True)
    out = solve(z, 2e-5, 0.04, model, tol=1e-6, frac=0.5, max_iter=30000, chunk=5000, log=lambda *a: None)
    cold, warm = out["cold"], out["warm"]
    print(f"cold {cold['iterations']} it, warm {warm['iterations']} it; first mean change "
          f"{cold['first_mean_change']:.3e} vs {warm['first_mean_change']:.3e}", flush=True)
    assert cold["converged"] and warm["converged"]
    # the network's start is closer to the steady state ...
    assert warm["first_mean_change"] < cold["first_mean_change"]
    assert out["used"] == "warm"
    # ... and does not slow the solver down
    assert warm["iterations"] <= cold["iterations"]


def test_warm_start_2048_held_out(pf):
    """The warm-start configuration itself: a held-out 2048^2 DEM (config-2
    style synthesis, relief 1000 m, 500 pits, seed 9091), rain 1e-5 m/s,
    n = 0.033, single depth update.  Inference, then the cold and the warm
    solve to the same test (at most half of the cells change by more than
    1e-6 m over 100 iterations).  The warm start needs at most half of the
    cold start's iterations, and the two stop states agree on the median
    cell to 1 mm (they differ in lakes the cold run is still filling)."""
    from pyfastflow_amd.surrogate import UNet
    from pyfastflow_amd.surrogate.warmstart import WEIGHTS, load_weights, solve

    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights committed")
    model = load_weights(UNet(), WEIGHTS).cuda().eval()
    n = 2048
    z = synth.dem(n, n, 2.0, 1000.0, 9091, n_pits=500, dx=10.0)
    print("case warm start 2048^2 held-out", flush=True)
    out = solve(z, 1e-5, 0.033, model, tol=1e-6, frac=0.5, max_iter=40000, chunk=5000, update="single",
                log=lambda *a: print(*a, flush=True))
    cold, warm, agree = out["cold"], out["warm"], out["agreement"]
    print(f"cold {cold['iterations']} it {cold['seconds']:.1f} s, warm {warm['iterations']} it "
          f"{warm['seconds']:.1f} s; agreement {agree}", flush=True)
    assert cold["converged"] and warm["converged"]
    assert cold["nan_cells"] == 0 and warm["nan_cells"] == 0
    assert out["used"] == "warm"
    assert warm["iterations"] <= 0.5 * cold["iterations"]
    assert agree["median_m"] <= 1e-3
