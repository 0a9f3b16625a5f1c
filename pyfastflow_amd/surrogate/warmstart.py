"""Training data, training and the warm-started GraphFlood solve.

Inputs of the network for a DEM with rainfall P and Manning's n: the elevation
(standardised per tile), the log drainage area after lakeflow rerouting (the
accumulation of FlowRouter on the DEM), log10 P and n, as four planes.  The
target is log(h + H_FLOOR) of the GraphFlood water depth the plain solver
reaches from a cold start, learnt as a correction to an analytic first guess
(Manning depth on hillslopes, fill depth in lakes).

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


def _planes(router: FlowRouter, rain: float, manning: float, with_base: bool = False):
    """(1, 6, ny, nx) input planes: standardised z, log drainage area after
    lakeflow, log steepest-descent slope, log depth of the filled
    depressions, log10 rainfall and Manning's n.  Leaves router.Q = drainage
    area (m^2).  On a hillslope the steady depth is a power law of area,
    slope, rainfall and n (Manning), and in a filled lake it is the fill
    depth, so these planes carry what the depth depends on.

    with_base also returns the (ny, nx) analytic first guess the network
    corrects: log(max(h_manning, fill depth) + H_FLOOR), with h_manning the
    depth at which Manning outflow over the steepest slope carries the
    rainfall of the drainage area, (P A n / (dx sqrt(S)))^(3/5)."""
    from ..flow.fill_topo import topofill

    ny, nx = router.ny, router.nx
    router.compute_receivers()
    slope = router.gradient.data.view(ny, nx).clone()
    router.reroute_flow(carve=True)
    router.accumulate_constant_Q(1.0, area=True)
    z = router.z.data.view(ny, nx)
    zf = torch.empty_like(router.z.data)
    topofill(router, epsilon=1e-4, custom_z=zf)
    depth = (zf.view(ny, nx) - z).clamp_min(0)
    area = router.Q.data.view(ny, nx)
    zn = (z - z.mean()) / (z.std() + 1e-6)
    la = (torch.log(area) - 10.0) / 3.0
    ls = (torch.log(slope.clamp_min(1e-4)) + 3.0) / 2.0
    ld = (torch.log(depth + H_FLOOR) + 3.0) / 3.0
    pr = torch.full_like(z, (math.log10(rain) + 4.65) / 0.5)
    mn = torch.full_like(z, (manning - 0.04) / 0.012)
    planes = torch.stack([zn, la, ls, ld, pr, mn])[None].float()
    if not with_base:
        return planes
    dx = float(router.dx)
    hm = (rain * area * manning / (dx * torch.sqrt(slope.clamp_min(1e-4)))) ** 0.6
    base = torch.log(torch.maximum(hm, depth) + H_FLOOR)
    return planes, base.float()


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
    (S, 6, tile, tile) and targets (S, tile, tile) on the GPU; a target is
    the log depth minus the analytic first guess of _planes."""
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
        planes, base = _planes(r, rain, man, with_base=True)
        fl = Flooder(r, precipitation_rates=rain, manning=man, edge_slope=1e-2, update=update)
        _reset_start(r, fl, z, None, rain)
        t0 = time.time()
        st = fl.run_graphflood(iters)
        torch.cuda.synchronize()
        h = fl.h.data.view(size, size)
        log(f"[surrogate] domain {d}: rain {rain:.2e} n {man:.3f} relief {relief:.0f} m, {iters} iterations in "
            f"{time.time() - t0:.1f} s, h mean {float(h.mean()):.4f} max {float(h.max()):.2f}")
        y = target_of(h) - base  # the network learns the correction to the analytic guess
        for i in range(0, size, tile):
            for j in range(0, size, tile):
                X.append(planes[0, :, i:i + tile, j:j + tile])
                Y.append(y[i:i + tile, j:j + tile])
        del fl, r
    return torch.stack(X).contiguous(), torch.stack(Y).contiguous()


def make_dataset_tiles(n_tiles: int = 2000, tile: int = 256, per_side: int = 10, seed: int = 505,
                       max_iter: int = 8000, tol: float = 1e-6, frac: float = 0.5, update: str = "single",
                       log=print):
    """Training set of independent tile x tile DEMs (config-2 style: beta
    2.0, relief U[300, 1000] m, pits in proportion to the area), each solved
    by the plain GraphFlood solver to the steady-state test of solve() (at
    most frac of the cells change by more than tol over 100 iterations, or
    max_iter iterations).  per_side^2 tiles are solved at once on one grid in
    custom boundary mode: each tile is framed by cells that drain out (code
    3, a 'normal' tile of its own) and separated from its neighbours by a row
    and column of no-data cells (code 0), so the tiles do not interact.
    Rainfall U[5e-6, 5e-5] m/s and n U[0.02, 0.06] are drawn per grid.  The
    solve starts from the analytic first guess of _planes (the state the
    network corrects), Q from the accumulated rainfall.  Returns planes (S,
    6, tile, tile), targets (S, tile, tile) and per-grid records."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pits = int(round(2000 * tile * tile / 4096.0**2))
    S = per_side * (tile + 1) + 1
    X, Y, recs = [], [], []
    done = 0
    while done < n_tiles:
        k = min(per_side * per_side, n_tiles - done)
        rain = float(rng.uniform(5e-6, 5e-5))
        man = float(rng.uniform(0.02, 0.06))
        zc = np.zeros((S, S), np.float32)
        codes = np.zeros((S, S), np.uint8)
        h0 = np.zeros((S, S), np.float32)
        slots, bases = [], []
        t0 = time.time()
        for t in range(k):
            ty, tx = divmod(t, per_side)
            y0, x0 = 1 + ty * (tile + 1), 1 + tx * (tile + 1)
            z = synth.dem(tile, tile, 2.0, float(rng.uniform(300.0, 1000.0)), int(rng.integers(0, 2**31 - 1)),
                          n_pits=pits, dx=10.0)
            zc[y0:y0 + tile, x0:x0 + tile] = z
            c = codes[y0:y0 + tile, x0:x0 + tile]
            c[:] = 1
            c[0, :] = c[-1, :] = 3
            c[:, 0] = c[:, -1] = 3
            r = FlowRouter(tile, tile, 10.0, boundary_mode="normal")
            r.set_z(z)
            planes, base = _planes(r, rain, man, with_base=True)
            X.append(planes[0])
            bases.append(base)
            h0[y0:y0 + tile, x0:x0 + tile] = depth_of(base).cpu().numpy()
            slots.append((y0, x0))
            del r
        router = FlowRouter(S, S, 10.0, boundary_mode="custom", boundaries=codes)
        fl = Flooder(router, precipitation_rates=rain, manning=man, edge_slope=1e-2, update=update)
        _reset_start(router, fl, zc, torch.from_numpy(h0).cuda(), rain)
        its = 0
        while its < max_iter:
            st = fl.run_to_steady(tol=tol, check_every=100, max_iter=min(2000, max_iter - its), frac=frac)
            its += int(st.iterations)
            if st.converged or st.nan_cells:
                break
        torch.cuda.synchronize()
        h = fl.h.data.view(S, S)
        for (y0, x0), base in zip(slots, bases):
            Y.append(target_of(h[y0:y0 + tile, x0:x0 + tile]) - base)
        rec = {"tiles": k, "rain": rain, "manning": man, "iterations": its, "converged": bool(st.converged),
               "above_frac": float(st.above_frac), "q50": float(st.q50), "seconds": round(time.time() - t0, 1)}
        recs.append(rec)
        log(f"[surrogate] grid {len(recs)}: {rec}")
        done += k
        del fl, router
    return torch.stack(X).contiguous(), torch.stack(Y).contiguous(), recs


def train(model: UNet, X: torch.Tensor, Y: torch.Tensor, epochs: int = 20, batch: int = 8, lr: float = 2e-3,
          seed: int = 0, log=print):
    """Adam on the mean squared error of log depth, with random flips and
    transpositions of the tiles.  Returns the per-epoch mean losses."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    model.target_range.copy_(torch.stack([Y.min(), Y.max()]).to(model.target_range))
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
    """Predicted depth: the analytic first guess corrected by the network,
    the correction clamped to the range it was trained on
    (model.target_range)."""
    planes, base = _planes(router, rain, manning, with_base=True)
    lo, hi = (float(v) for v in model.target_range)
    return depth_of(base + model(planes)[0].clamp(lo, hi))


def solve(z: np.ndarray, rain: float, manning: float, model: UNet | None, tol: float, frac: float = 0.0,
          check_every: int = 100, max_iter: int = 200000, update: str = "single", log=print, chunk: int = 5000):
    """GraphFlood to steady (Flooder.run_to_steady with tol and frac) from a
    cold start, and from the model's warm start (if a model is given).  The
    warm start is kept only if its first residual (the mean change of h over
    the first check block) is not worse than the cold start's.  Returns a
    dict with iterations, wall times (inference included) and residuals of
    both."""
    ny, nx = z.shape
    out = {}
    finals = {}
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
        # in chunks (whole check blocks) so that progress is visible
        its, first, first_frac, first_mean = 0, None, None, None
        while True:
            n = min(chunk, max_iter - its)
            st = fl.run_to_steady(tol=tol, check_every=check_every, max_iter=n, frac=frac)
            its += int(st.iterations)
            if first is None:
                first, first_frac = float(st.first_residual), float(st.first_above_frac)
                first_mean = float(st.first_mean_change)
            log(f"[surrogate] {mode}: {its} iterations, max change {float(st.residual):.3e} m, mean change "
                f"{float(st.mean_change):.3e} m, {100 * float(st.above_frac):.1f}% of cells above tol")
            if st.converged or st.nan_cells or its >= max_iter:
                break
        torch.cuda.synchronize()
        out[mode] = {"iterations": its, "seconds": time.perf_counter() - t0,
                     "residual": float(st.residual), "first_residual": first,
                     "above_frac": float(st.above_frac), "first_above_frac": first_frac,
                     "mean_change": float(st.mean_change), "first_mean_change": first_mean,
                     "q50": float(st.q50), "q90": float(st.q90), "q99": float(st.q99),
                     "converged": bool(st.converged), "nan_cells": int(st.nan_cells)}
        log(f"[surrogate] {mode}: {out[mode]}")
        finals[mode] = fl.h.data.clone()
        del fl, r
    if "warm" in finals:
        # where the two end states agree: |h_warm - h_cold| per cell
        d = (finals["warm"] - finals["cold"]).abs().float()
        q = torch.quantile(d[::7][:1 << 24], torch.tensor([0.5, 0.9, 0.99], device=d.device)).tolist()
        out["agreement"] = {"max_m": float(d.max()), "median_m": q[0], "p90_m": q[1], "p99_m": q[2],
                            "frac_within_1e-3_m": float((d <= 1e-3).float().mean()),
                            "frac_within_1e-2_m": float((d <= 1e-2).float().mean())}
        log(f"[surrogate] warm vs cold end states: {out['agreement']}")
    # the initial residual: mean change of h over the first check block
    key = "first_mean_change"
    if "warm" in out and out["warm"][key] > out["cold"][key]:
        out["used"] = "cold (warm start's first residual was worse)"
    else:
        out["used"] = "warm" if "warm" in out else "cold"
    return out


WEIGHTS = __file__.rsplit("/", 1)[0] + "/unet_warmstart.pt"


def save_weights(model: UNet, path: str = WEIGHTS):
    """bf16 copy of the weights (the kernels use bf16 operands) and the f32
    target range."""
    sd = {k: (v.detach().to(torch.bfloat16).cpu() if v.dtype == torch.float32 and k != "target_range"
              else v.detach().cpu()) for k, v in model.state_dict().items()}
    torch.save(sd, path)


def load_weights(model: UNet, path: str = WEIGHTS) -> UNet:
    sd = torch.load(path, map_location="cpu", weights_only=True)
    model.load_state_dict({k: v.float() for k, v in sd.items()})
    return model
