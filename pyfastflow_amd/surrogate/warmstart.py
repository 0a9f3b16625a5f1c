"""Training data, training and the warm-started GraphFlood solve.

Inputs of the network for a DEM with rainfall P and Manning's n: the elevation
(standardised per tile), the log drainage area after lakeflow rerouting (the
accumulation of FlowRouter on the DEM), log10 P and n, as four planes.  The
target is log(h + H_FLOOR) of the GraphFlood water depth the plain solver
reaches from a cold start.

A warm start sets h to the prediction and zh = z + h (Flooder.set_h alone
does not touch zh, gf_fields.py:143-150), with Q from the accumulation of the
rainfall; the cold start is h = 0 with the same Q.  The solver then runs with
the same convergence test from either start, and the warm start is abandoned
for the cold one when its first residual is worse (``solve``).
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from .. import _lib, synth
from ..flow import FlowRouter
from ..graphflood import Flooder
from .unet import UNet

H_FLOOR = 1e-3  # m: log(h + H_FLOOR) keeps dry cells finite


def _planes(router: FlowRouter, rain: float, manning: float) -> torch.Tensor:
    """(1, 4, ny, nx) input planes; leaves router.Q = drainage area (m^2)."""
    router.compute_receivers()
    router.reroute_flow(carve=True)
    router.accumulate_constant_Q(1.0, area=True)
    ny, nx = router.ny, router.nx
    z = router.z.data.view(ny, nx)
    area = router.Q.data.view(ny, nx)
    zn = (z - z.mean()) / (z.std() + 1e-6)
    la = (torch.log(area) - 10.0) / 3.0
    pr = torch.full_like(z, (math.log10(rain) + 4.65) / 0.5)
    mn = torch.full_like(z, (manning - 0.04) / 0.012)
    return torch.stack([zn, la, pr, mn])[None].float()


def target_of(h: torch.Tensor) -> torch.Tensor:
    return torch.log(h.clamp_min(0) + H_FLOOR)


def depth_of(pred: torch.Tensor) -> torch.Tensor:
    return (torch.exp(pred) - H_FLOOR).clamp_min(0)


def _reset_start(router: FlowRouter, fl: Flooder, z0: np.ndarray, h0: torch.Tensor | None, rain: float):
    """zh = z + h0 (or z), h = h0 (or 0), Q = rain * drainage area."""
    router.set_z(z0)
    router.compute_receivers()
    router.reroute_flow(carve=True)
    router.accumulate_constant_Q(rain, area=True)
    if h0 is None:
        fl.h.data.zero_()
    else:
        fl.h.data.copy_(h0.reshape(-1))
        router.z.data.add_(fl.h.data)


def make_dataset(n_domains: int = 8, size: int = 1024, tile: int = 256, iters: int = 20000, seed: int = 505,
                 update: str = "single", log=print):
    """Plain-solver depth fields on config-2 style DEMs (beta 2.0, relief
    U[300, 1000] m, pits in proportion to area) with rainfall U[5e-6, 5e-5]
    m/s and n U[0.02, 0.06], cut into tile x tile samples.  Returns planes
    (S, 4, tile, tile) and targets (S, tile, tile) on the GPU."""
    rng = np.random.Generator(np.random.PCG64(seed))
    X, Y = [], []
    for d in range(n_domains):
        rain = float(rng.uniform(5e-6, 5e-5))
        man = float(rng.uniform(0.02, 0.06))
        relief = float(rng.uniform(300.0, 1000.0))
        pits = int(round(2000 * size * size / 4096.0**2))
        z = synth.dem(size, size, 2.0, relief, int(rng.integers(0, 2**31 - 1)), n_pits=pits, dx=10.0)
        r = FlowRouter(size, size, 10.0, boundary_mode="normal")
        r.set_z(z)
        planes = _planes(r, rain, man)
        fl = Flooder(r, precipitation_rates=rain, manning=man, edge_slope=1e-2, update=update)
        _reset_start(r, fl, z, None, rain)
        t0 = time.time()
        st = fl.run_graphflood(iters)
        torch.cuda.synchronize()
        h = fl.h.data.view(size, size)
        log(f"[surrogate] domain {d}: rain {rain:.2e} n {man:.3f} relief {relief:.0f} m, {iters} iterations in "
            f"{time.time() - t0:.1f} s, h mean {float(h.mean()):.4f} max {float(h.max()):.2f}")
        y = target_of(h)
        for i in range(0, size, tile):
            for j in range(0, size, tile):
                X.append(planes[0, :, i:i + tile, j:j + tile])
                Y.append(y[i:i + tile, j:j + tile])
        del fl, r
    return torch.stack(X).contiguous(), torch.stack(Y).contiguous()


def train(model: UNet, X: torch.Tensor, Y: torch.Tensor, epochs: int = 20, batch: int = 8, lr: float = 2e-3,
          seed: int = 0, log=print):
    """Adam on the mean squared error of log depth, with random flips and
    transpositions of the tiles.  Returns the per-epoch mean losses."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=max(epochs, 1))
    losses = []
    n = X.shape[0]
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        tot, cnt = 0.0, 0
        for k in range(0, n - batch + 1, batch):
            idx = perm[k:k + batch].to(X.device)
            xb, yb = X[idx], Y[idx]
            f = int(torch.randint(0, 8, (1,), generator=g))
            if f & 1:
                xb, yb = xb.flip(-1), yb.flip(-1)
            if f & 2:
                xb, yb = xb.flip(-2), yb.flip(-2)
            if f & 4:
                xb, yb = xb.transpose(-1, -2), yb.transpose(-1, -2)
            pred = model(xb.contiguous())
            loss = ((pred - yb) ** 2).mean()
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            tot += float(loss.detach())
            cnt += 1
        sched.step()
        losses.append(tot / max(cnt, 1))
        log(f"[surrogate] epoch {ep}: loss {losses[-1]:.4f}")
    return losses


@torch.no_grad()
def predict_depth(model: UNet, router: FlowRouter, rain: float, manning: float) -> torch.Tensor:
    planes = _planes(router, rain, manning)
    return depth_of(model(planes))[0]


def solve(z: np.ndarray, rain: float, manning: float, model: UNet | None, tol: float, frac: float = 0.0,
          check_every: int = 100, max_iter: int = 200000, update: str = "single", log=print):
    """GraphFlood to steady (Flooder.run_to_steady with tol and frac) from a
    cold start, and from the model's warm start (if a model is given).  The
    warm start is kept only if its first residual (the fraction of cells
    still moving by more than tol, or the max change when frac is 0) is not
    worse than the cold start's.  Returns a dict with iterations, wall times
    (inference included) and residuals of both."""
    ny, nx = z.shape
    out = {}
    for mode in (["cold", "warm"] if model is not None else ["cold"]):
        r = FlowRouter(nx, ny, 10.0, boundary_mode="normal")
        r.set_z(z)
        fl = Flooder(r, precipitation_rates=rain, manning=manning, edge_slope=1e-2, update=update)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h0 = None
        if mode == "warm":
            h0 = predict_depth(model, r, rain, manning)
        _reset_start(r, fl, z, h0, rain)
        st = fl.run_to_steady(tol=tol, check_every=check_every, max_iter=max_iter, frac=frac)
        torch.cuda.synchronize()
        out[mode] = {"iterations": int(st.iterations), "seconds": time.perf_counter() - t0,
                     "residual": float(st.residual), "first_residual": float(st.first_residual),
                     "above_frac": float(st.above_frac), "first_above_frac": float(st.first_above_frac),
                     "converged": bool(st.converged), "nan_cells": int(st.nan_cells)}
        log(f"[surrogate] {mode}: {out[mode]}")
        del fl, r
    key = "first_above_frac" if frac > 0 else "first_residual"
    if "warm" in out and out["warm"][key] > out["cold"][key]:
        out["used"] = "cold (warm start's first residual was worse)"
    else:
        out["used"] = "warm" if "warm" in out else "cold"
    return out
