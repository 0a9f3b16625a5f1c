"""Train the warm-start surrogate and compare warm- and cold-started
GraphFlood on a held-out DEM.

    python -m pyfastflow_amd.surrogate.experiment [--domains 8] [--size 1024]
        [--iters 24000] [--epochs 40] [--eval-size 2048] [--out weights.pt]

Steps: (1) plain-solver depth fields on config-2 style DEMs with random
rainfall and roughness, cut into 256x256 tiles; (2) training of the U-Net on
the GPU with the library's BF16 kernels; (3) GraphFlood to the same
convergence test from h = 0 and from the prediction on a 2048x2048 DEM the
network has not seen.  Prints one JSON line with the results.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from .. import synth
from .unet import UNet
from .warmstart import load_weights, make_dataset, make_dataset_tiles, save_weights, solve, train


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--domains", type=int, default=8)
    p.add_argument("--size", type=int, default=1024)
    p.add_argument("--iters", type=int, default=24000)
    p.add_argument("--epochs", type=int, default=40)
    p.add_argument("--eval-size", type=int, default=2048)
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--frac", type=float, default=0.5)
    p.add_argument("--max-iter", type=int, default=60000)
    p.add_argument("--out", default=None, help="save the trained weights (torch state dict)")
    p.add_argument("--weights", default=None, help="skip training: load these weights")
    p.add_argument("--tiles", type=int, default=0,
                   help="train on this many independent tile x tile DEMs solved to steady state "
                        "(make_dataset_tiles) instead of tiles cut from --domains larger DEMs")
    p.add_argument("--tile-iters", type=int, default=8000, help="iteration cap of each tile grid's solve")
    a = p.parse_args(argv)

    res = {}
    model = UNet().cuda()
    if a.weights:
        load_weights(model, a.weights)
    else:
        t0 = time.time()
        if a.tiles:
            X, Y, recs = make_dataset_tiles(n_tiles=a.tiles, max_iter=a.tile_iters, tol=a.tol, frac=a.frac, log=log)
            res["data"] = {"tiles": int(X.shape[0]), "tile": int(X.shape[-1]), "grids": recs,
                           "seconds": round(time.time() - t0, 1)}
        else:
            X, Y = make_dataset(n_domains=a.domains, size=a.size, iters=a.iters, log=log)
            res["data"] = {"tiles": int(X.shape[0]), "tile": int(X.shape[-1]), "domains": a.domains,
                           "domain_size": a.size, "solver_iterations": a.iters, "seconds": round(time.time() - t0, 1)}
        n_val = max(X.shape[0] // 8, 1)
        t0 = time.time()
        losses = train(model, X[n_val:], Y[n_val:], epochs=a.epochs, log=log)
        torch.cuda.synchronize()
        with torch.no_grad():
            val = float(((model(X[:n_val]) - Y[:n_val]) ** 2).mean())
            base = float(((Y[:n_val] - Y[n_val:].mean()) ** 2).mean())
            guess = float((Y[:n_val] ** 2).mean())  # error of the analytic guess alone
        flops = model.flops_per_pixel() * 3 * X[n_val:].shape[0] * X.shape[-1] ** 2 * a.epochs
        dt = time.time() - t0
        res["train"] = {"epochs": a.epochs, "first_loss": losses[0], "last_loss": losses[-1], "val_mse": val,
                        "val_mse_of_mean_correction": base, "val_mse_of_analytic_guess": guess,
                        "seconds": round(dt, 1),
                        "TFLOP_per_s": round(flops / dt / 1e12, 2)}
        log(f"[surrogate] training {res['train']}")
        if a.out:
            save_weights(model, a.out)
    model.eval()

    # held-out DEM: config-2 style synthesis at eval_size, open edges
    n = a.eval_size
    z = synth.dem(n, n, 2.0, 1000.0, 9091, n_pits=int(round(2000 * n * n / 4096.0**2)), dx=10.0)
    rain, man = 1e-5, 0.033
    with torch.no_grad():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        planes = torch.randn(1, model.in_planes, n, n, device="cuda")
        model(planes)
        torch.cuda.synchronize()
        t_inf = time.perf_counter() - t0
    res["inference"] = {"pixels": n * n, "seconds": round(t_inf, 4),
                        "TFLOP_per_s": round(model.flops_per_pixel() * n * n / t_inf / 1e12, 2)}
    out = solve(z, rain, man, model, tol=a.tol, frac=a.frac, max_iter=a.max_iter, log=log)
    res["solve"] = out
    res["criterion"] = {"tol_m": a.tol, "frac": a.frac, "check_every": 100,
                        "meaning": "at most frac of the cells change by more than tol over 100 iterations"}
    if out["cold"]["converged"] and out.get("warm", {}).get("converged"):
        res["warm_over_cold_time"] = round(out["warm"]["seconds"] / out["cold"]["seconds"], 3)
        res["warm_over_cold_iterations"] = round(out["warm"]["iterations"] / out["cold"]["iterations"], 3)
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main()
