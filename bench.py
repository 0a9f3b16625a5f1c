#!/usr/bin/env python3
"""Benchmark: receivers + lakeflow + accumulation on the 4096^2 depressions
workload (periodic east-west, 2000 Gaussian pits, dx=10 m), one MI355X.

One step = compute_receivers -> reroute_flow(carve=True) ->
accumulate_constant_Q(1e-5, area=True) over the whole grid, with the DEM
already resident in HBM.  Prints one JSON line (see the keys below).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 4096] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mcells/s route+accumulate at 4096² DEM; graphflood s-to-steady warm vs cold"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--pits", type=int, default=2000)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and full-size parity check")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the ragged-batch line (256 tiles, sides 128-1024, routed in one launch sequence)")
    p.add_argument("--no-graphflood", action="store_true",
                   help="skip the GraphFlood line (2048^2, iterations of the reference schedule)")
    p.add_argument("--steady-budget", type=int, default=10000,
                   help="iterations of the 2048^2 GraphFlood run towards max |dh| < 1e-6 m (checked every 100)")
    p.add_argument("--no-warm", action="store_true",
                   help="skip the surrogate warm start against the cold start (2048^2, two solves)")
    p.add_argument("--no-surrogate", action="store_true",
                   help="skip the surrogate's matrix-core line (2048^2 inference, one training step)")
    p.add_argument("--pmc-calibrate", action="store_true",
                   help="after the steps, run a 1 GiB read + 1 GiB write elementwise pass 3 times "
                        "(scale for rocprofv3 FETCH/WRITE_SIZE)")
    return p.parse_args()


def kernel_report(_lib, ctx, names):
    """Per-kernel launches, total ms and algorithmic GB/s of the instrumented
    launches since pff_ctx_profile(ctx, 1), minus the event pair's own cost."""
    import ctypes

    ovh = ctypes.c_double()
    _lib.call("pff_ctx_launch_overhead", ctx, 50, ctypes.byref(ovh))
    out = {}
    for name in names:
        n_l, t_ms, t_b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _lib.call("pff_ctx_profile_report", ctx, name.encode(), ctypes.byref(n_l), ctypes.byref(t_ms),
                  ctypes.byref(t_b))
        if n_l.value > 0:
            t = max(t_ms.value - n_l.value * ovh.value * 1e-3, 1e-6)
            gbs = t_b.value / (t * 1e-3) / 1e9
            out[name] = {"launches": int(n_l.value), "avg_us": round(1e3 * t / n_l.value, 2),
                         "bytes_per_launch": round(t_b.value / n_l.value), "achieved_GBs": round(gbs, 1),
                         "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def fp64_drift(synth, iters=60, n=128):
    """The f32 GraphFlood method against its FP64 restatement on a 128^2
    grid synthesised like the 2048^2 GraphFlood workload (beta 2.2, 300 m
    relief, dx 10 m), cold start: error per result (max, p99, argmax) after
    `iters` iterations of the reference schedule.  Run on the CPU by the
    oracle; the GPU gives the same bits as the f32 oracle (checked by the
    caller), so this is also the GPU's error."""
    from oracle import pff_oracle as O

    z = synth.dem(n, n, 2.2, 300.0, 303, n_pits=2, dx=10.0).ravel()
    g = O.Grid(n, n, 10.0, 0)
    hy = O.hydro(prec=1e-5, manning=0.033, edge_slope=1e-2, dt=5e-3, update=0)
    zh, h, Q = z.copy(), np.zeros(z.size, np.float32), np.zeros(z.size, np.float32)
    zh64, h64, Q64 = z.astype(np.float64), np.zeros(z.size), np.zeros(z.size)
    for _ in range(iters):
        zh, h, Q, _ = O.graphflood_iteration(g, hy, zh, h, Q)
        zh64, h64, Q64, _ = O.graphflood_iteration(g, hy, zh64, h64, Q64, f64=True)
    out = {}
    for name, a, b in (("h", h, h64), ("zh", zh, zh64), ("Q", Q, Q64)):
        e = np.abs(a.astype(np.float64) - b)
        bad = (e > 1e-4) & (e > 1e-5 * np.abs(b))
        out[name] = {"max_abs": float(e.max()), "p99_abs": float(np.quantile(e, 0.99)), "argmax_cell": int(np.argmax(e)),
                     "cells_outside_1e-5rel_1e-4abs": int(bad.sum())}
    return z, (zh, h, Q), out


def gpu_drift(pf, synth, n_small, it_small):
    z, (zh32, h32, Q32), drift = fp64_drift(synth, it_small, n_small)
    rs = pf.flow.FlowRouter(n_small, n_small, 10.0, boundary_mode="normal")
    rs.set_z(z)
    fs = pf.graphflood.Flooder(rs, precipitation_rates=1e-5, manning=0.033, edge_slope=1e-2)
    fs.run_graphflood(it_small)
    drift["gpu_bit_exact_with_f32_oracle"] = bool(
        np.array_equal(fs.h.to_numpy().view(np.uint32), h32.view(np.uint32))
        and np.array_equal(rs.z.to_numpy().view(np.uint32), zh32.view(np.uint32))
        and np.array_equal(rs.Q.to_numpy().view(np.uint32), Q32.view(np.uint32)))
    drift["workload"] = (f"{n_small}x{n_small} beta 2.2, relief 300 m, 2 pits, seed 303, {it_small} iterations from "
                         "h = Q = 0, reference update; f32 method vs its FP64 restatement")
    return drift


def graphflood_block(pf, _lib, synth, s, args):
    import torch

    cg = synth.graphflood_plain()
    N = cg["nx"] * cg["ny"]
    rg = pf.flow.FlowRouter(cg["nx"], cg["ny"], cg["dx"], boundary_mode="normal")
    rg.set_z(cg["z"])
    fl = pf.graphflood.Flooder(rg, precipitation_rates=cg["precipitation"], manning=cg["manning"], edge_slope=1e-2)
    fl.run_graphflood(5)
    torch.cuda.synchronize()
    n_it = 200
    eg = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    eg[0].record(s)
    fl.run_graphflood(n_it)
    eg[1].record(s)
    torch.cuda.synchronize()
    gms = eg[0].elapsed_time(eg[1]) / n_it
    h = fl.get_h()
    # the fused hydro kernel and the fill passes over 20 more iterations
    ctx = _lib.ctx_for()
    _lib.call("pff_ctx_profile", ctx, 1)
    fl.run_graphflood(20)
    kern = kernel_report(_lib, ctx, ("k_hydro", "k_topofill"))
    _lib.call("pff_ctx_profile", ctx, 0)
    del rg, fl
    out = {"workload": "2048x2048 normal dx=10 m, beta=2.2, relief 300 m, 500 pits, seed 303; P=1e-5 m/s, "
                       "n=0.033, edge_slope=1e-2, dt=5e-3 s, fill eps=1e-3, 5 diffusions; reference update",
           "iterations_timed": n_it, "ms_per_iteration": round(gms, 4),
           "Mcells_per_s": round(N / (gms * 1e-3) / 1e6, 1),
           "h_mean_m": float(np.nanmean(h)), "h_nan_cells": int(np.isnan(h).sum()),
           "kernels": {"k_hydro (fill apply + 5 diffusions + Manning, 32 B/cell)": kern.get("k_hydro"),
                       "k_topofill (one fill pass, 28 B/cell)": kern.get("k_topofill")}}
    # FP64 drift of the f32 method (oracle, CPU) and the GPU's bits against it
    n_small, it_small = 128, 60
    if args.no_cpu:
        out["fp64_drift"] = None
    else:
        out["fp64_drift"] = gpu_drift(pf, synth, n_small, it_small)
    # towards steady state: cold start (h = 0, Q = 0), reference update,
    # max |h(t) - h(t-100)| < 1e-6 m checked every 100 iterations
    if args.steady_budget <= 0:
        out["steady_state"] = None
        torch.cuda.empty_cache()
        return out
    rg = pf.flow.FlowRouter(cg["nx"], cg["ny"], cg["dx"], boundary_mode="normal")
    rg.set_z(cg["z"])
    fl = pf.graphflood.Flooder(rg, precipitation_rates=cg["precipitation"], manning=cg["manning"], edge_slope=1e-2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its, median_it, nan_it, stop = 0, None, None, "budget"
    while its < args.steady_budget:
        st = fl.run_to_steady(tol=1e-6, check_every=100, max_iter=100, frac=0.0)
        its += int(st.iterations)
        if median_it is None and st.q50 < 1e-6:
            median_it = its
        if st.converged:
            stop = "converged"
            break
        if st.nan_cells:
            stop, nan_it = "NaN depth", its
            break
    torch.cuda.synchronize()
    out["steady_state"] = {
        "criterion": "max |h(t) - h(t-100)| < 1e-6 m, checked every 100 iterations (cold start h = 0, Q = 0)",
        "reached": stop == "converged", "stop": stop, "iterations": its,
        "seconds": round(time.perf_counter() - t0, 2),
        "at_stop_m": {"max": st.residual, "mean": st.mean_change, "q50": st.q50, "q90": st.q90, "q99": st.q99,
                      "frac_above_1e-6": st.above_frac},
        "nan_iteration": nan_it,
        "median_below_1e-6_at_iteration": median_it}
    del rg, fl
    torch.cuda.empty_cache()
    return out


def surrogate_block():
    """Matrix-core throughput of the surrogate's BF16 MFMA kernels: 2048^2
    inference (fused convolutions) and one training step (forward + backward
    of a batch of 8 256^2 tiles), FLOPs of the 3x3 convolutions and the head
    (UNet.flops_per_pixel; backward = 2x forward) against the 2.5 PFLOP/s
    dense BF16 peak."""
    import torch

    from pyfastflow_amd.surrogate import UNet
    from pyfastflow_amd.surrogate.warmstart import load_weights

    peak = 2500.0  # TFLOP/s, dense BF16
    m = load_weights(UNet().cuda()).eval()
    out = {}
    n = 2048
    x = torch.randn(1, m.in_planes, n, n, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    with torch.no_grad():
        for _ in range(2):
            m(x)
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(10):
            m(x)
        ev[1].record()
        torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 10
    fl = m.flops_per_pixel() * n * n
    out["inference_2048"] = {"ms": round(ms, 3), "TFLOP": round(fl / 1e12, 3), "TFLOP_per_s": round(fl / ms / 1e9, 1),
                             "frac_of_bf16_peak": round(fl / ms / 1e9 / peak, 4)}
    m.train()
    xb = torch.randn(8, m.in_planes, 256, 256, device="cuda")
    yb = torch.randn(8, 256, 256, device="cuda")
    opt = torch.optim.Adam(m.parameters(), lr=1e-4)

    def step():
        loss = ((m(xb) - yb) ** 2).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(10):
        step()
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 10
    fl = 3 * m.flops_per_pixel() * 8 * 256 * 256
    out["training_step_8x256"] = {"ms": round(ms, 3), "TFLOP_per_s": round(fl / ms / 1e9, 1),
                                  "frac_of_bf16_peak": round(fl / ms / 1e9 / peak, 4)}
    del m, opt
    torch.cuda.empty_cache()
    return out


def warm_start_block(pf, synth, args):
    """Surrogate warm start against the cold start on a held-out 2048^2
    config-2-style DEM (beta 2.0, relief 1000 m, 500 pits, seed 9091),
    P = 1e-5 m/s, n = 0.033, single depth update; the reference update's
    first NaN on the same DEM is measured beside it."""
    import torch

    from pyfastflow_amd.surrogate import UNet
    from pyfastflow_amd.surrogate.warmstart import load_weights, solve

    n = 2048
    z = synth.dem(n, n, 2.0, 1000.0, 9091, n_pits=500, dx=10.0)
    model = load_weights(UNet().cuda()).eval()
    out = solve(z, 1e-5, 0.033, model, tol=1e-6, frac=0.5, max_iter=40000, update="single", log=log)
    res = {"workload": "2048x2048 held-out DEM: beta 2.0, relief 1000 m, 500 pits, seed 9091; P=1e-5 m/s, n=0.033, "
                       "single depth update; cold = h 0, warm = network prediction (inference timed in)",
           "criterion": "at most 50% of the cells change by more than 1e-6 m over 100 iterations",
           "cold": out["cold"], "warm": out["warm"], "used": out["used"], "agreement": out.get("agreement")}
    if out["cold"]["converged"] and out["warm"]["converged"]:
        res["warm_over_cold_seconds"] = round(out["warm"]["seconds"] / out["cold"]["seconds"], 3)
        res["warm_over_cold_iterations"] = round(out["warm"]["iterations"] / out["cold"]["iterations"], 3)
    # the reference's update on the same DEM, cold: first NaN depth
    r = pf.flow.FlowRouter(n, n, 10.0, boundary_mode="normal")
    r.set_z(z)
    fl = pf.graphflood.Flooder(r, precipitation_rates=1e-5, manning=0.033, edge_slope=1e-2, update="reference")
    its, nan_it = 0, None
    while its < 2000:
        st = fl.run_to_steady(tol=1e-6, check_every=100, max_iter=100, frac=0.0)
        its += int(st.iterations)
        if st.nan_cells:
            nan_it = its
            break
    res["reference_update_first_nan_check"] = nan_it
    del r, fl, model
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the routing path runs on a single MI355X here "
              "(multi-GPU sharding is a design only, not run)", file=sys.stderr)
        sys.exit(2)

    import torch

    import pyfastflow_amd as pf
    from pyfastflow_amd import _lib, synth

    n = args.size
    pits = args.pits if n == 4096 else int(round(args.pits * n * n / 4096.0 ** 2))
    log(f"[bench] generating {n}x{n} depressions DEM ({pits} pits)")
    c = synth.depressions(n=n, n_pits=pits)
    nx, ny, dx = c["nx"], c["ny"], c["dx"]
    N = nx * ny
    router = pf.flow.FlowRouter(nx, ny, dx, boundary_mode=c["boundary_mode"])
    router.set_z(c["z"])
    torch.cuda.synchronize()

    def step():
        router.compute_receivers()
        router.reroute_flow(carve=True)
        router.accumulate_constant_Q(c["precipitation"], area=True)

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # ---- timed region ---------------------------------------------------------
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        step()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    log(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step (event), {1e3 * wall / args.steps:.3f} ms/step (wall)")
    st = router.last_lake_status
    lake = dict(ndep0=int(st.ndep0), trips=int(st.n_iter), ndep=[int(st.ndep[i]) for i in range(min(st.n_loop, 64))])
    log(f"[bench] lakeflow: {lake}")

    # ---- per-stage times ------------------------------------------------------
    def time_stage(fn, reps=5):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record(s)
        for _ in range(reps):
            fn()
        ev[1].record(s)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps

    z_copy = router.z.data.clone()
    t_rcv = time_stage(router.compute_receivers, 20)
    rec_after = router.receivers.data.clone()

    def lake_only():
        router.compute_receivers()
        router.reroute_flow(carve=True)

    t_lake = time_stage(lake_only, 3) - t_rcv
    t_acc = time_stage(lambda: router.accumulate_constant_Q(c["precipitation"], area=True), 5)
    stages = {"receivers_ms": round(t_rcv, 4), "lakeflow_ms": round(t_lake, 4), "accumulate_ms": round(t_acc, 4)}
    log(f"[bench] stages: {stages}")

    # ---- roofline of the dominant kernel: every instrumented kernel is timed
    #      live during one extra step with HIP events that its launch records
    #      at the kernel's own start and end (hipExtLaunchKernelGGL, the times
    #      a kernel trace reports), and counts the algorithmic bytes of the
    #      units it processed (cells of active tiles, or listed nodes); the
    #      kernel with the largest total time is the dominant one
    import ctypes

    ctx = _lib.ctx_for()
    _lib.call("pff_ctx_profile", ctx, 1)
    step()
    kernels = {}
    for name in ("k_flag_scatter2", "k_zprime_min_t", "k_tag_local", "k_tag_final", "k_tile"):
        n_l, t_ms, t_b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _lib.call("pff_ctx_profile_report", ctx, name.encode(), ctypes.byref(n_l), ctypes.byref(t_ms),
                  ctypes.byref(t_b))
        if n_l.value > 0:
            kernels[name] = (int(n_l.value), t_ms.value, t_b.value)
    _lib.call("pff_ctx_profile", ctx, 0)
    # what the event pair adds to a launch beyond the kernel's own run
    # (median over an empty kernel), taken off every launch so that the
    # durations are those a kernel trace reports
    ovh = ctypes.c_double()
    _lib.call("pff_ctx_launch_overhead", ctx, 50, ctypes.byref(ovh))
    ovh_ms = ovh.value * 1e-3
    kernels = {k: (n_, max(t_ - n_ * ovh_ms, 1e-6), b_) for k, (n_, t_, b_) in kernels.items()}
    log(f"[bench] event pair on an empty kernel: {ovh.value:.2f} us per launch (subtracted)")
    dom = max(kernels, key=lambda k: kernels[k][1])
    launches, t_tot, b_tot = kernels[dom]
    avg_us = 1e3 * t_tot / launches
    bytes_launch = b_tot / launches
    achieved = b_tot / (t_tot * 1e-3) / 1e9
    # HBM bytes per launch of that kernel: not measured in this run (counters
    # need a profiler run of their own); read from the stored rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of the same workload
    traffic, traffic_src = None, None
    tfile = os.path.join(HERE, "profiles", f"{dom}_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        traffic_src = f"stored profile profiles/{dom}_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": dom,
                "launches_per_step": launches, "avg_launch_us": round(avg_us, 2),
                "bytes_per_launch": round(bytes_launch), "event_overhead_us": round(ovh.value, 2)}
    log(f"[bench] roofline {roofline}")
    kernel_times = {k: {"launches": v[0], "ms_per_step": round(v[1], 4),
                        "GBs": round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[1] > 0 else None}
                    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1][1])}
    log(f"[bench] instrumented kernels: {kernel_times}")
    bytes_rcv = 12.0 * N
    rcv_roof = {"kernel": "k_receivers", "us": round(t_rcv * 1e3, 2),
                "achieved_GBs": round(bytes_rcv / (t_rcv * 1e-3) / 1e9, 1)}

    if args.pmc_calibrate:
        # an elementwise kernel that reads 1 GiB and writes 1 GiB (four times
        # the last-level cache, so the read comes from HBM): its FETCH_SIZE and
        # WRITE_SIZE per dispatch give the counters' scale to bytes
        a = torch.ones(256 * 2**20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        log("[bench] calibration: 3 elementwise passes reading and writing 1 GiB each")
        del a, b

    # ---- ragged batch: 256 'normal' tiles, sides log-uniform in [128, 1024],
    #      routed together (receivers + reroute_flow + accumulate per tile)
    batch = None
    c4 = None
    batch_out = None
    if not args.no_batch:
        c4 = synth.ragged_batch(256)
        tiles = [t["z"] for t in c4]
        zt = [torch.from_numpy(z).cuda() for z in tiles]
        bcells = sum(z.size for z in tiles)
        res = pf.flow.route_batch(zt, dx=10.0, value=1e-5)
        torch.cuda.synchronize()
        eb = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = 3
        eb[0].record(s)
        for _ in range(reps):
            res = pf.flow.route_batch(zt, dx=10.0, value=1e-5, out=res)
        eb[1].record(s)
        torch.cuda.synchronize()
        bms = eb[0].elapsed_time(eb[1]) / reps
        # the dominant kernel of one more batch call, timed live as above
        _lib.call("pff_ctx_profile", ctx, 1)
        res = pf.flow.route_batch(zt, dx=10.0, value=1e-5, out=res)
        bk = kernel_report(_lib, ctx, ("k_flag_scatter2", "k_zprime_min_t", "k_tag_local", "k_tag_final", "k_tile"))
        _lib.call("pff_ctx_profile", ctx, 0)
        torch.cuda.synchronize()
        bdom = max(bk, key=lambda k: bk[k]["launches"] * bk[k]["avg_us"])
        batch = {"workload": "256 tiles, sides round(128*8^u), u~U[0,1), seed 404, normal, dx=10 m",
                 "cells": int(bcells), "ms": round(bms, 3), "Mcells_per_s": round(bcells / (bms * 1e-3) / 1e6, 1),
                 "status": res.status,
                 "roofline": {"bound": "hbm", "kernel": bdom, "achieved": bk[bdom]["achieved_GBs"],
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": bk[bdom]["frac_of_hbm_peak"],
                              "launches_per_call": bk[bdom]["launches"], "avg_launch_us": bk[bdom]["avg_us"],
                              "bytes_per_launch": bk[bdom]["bytes_per_launch"]},
                 "instrumented_kernels": bk}
        log(f"[bench] batch {batch}")
        batch_out = ([r.cpu().numpy() for r in res.receivers], [q.cpu().numpy() for q in res.Q])
        del zt, res

    # ---- GraphFlood: 2048^2 open-edge DEM, cold start, run_graphflood
    #      iterations (receivers -> reroute_flow -> fill -> 5 diffusions ->
    #      Manning) with the reference's depth update; then the same DEM
    #      iterated towards steady state, and the warm-start comparison
    gflood = None
    warm = None
    mfma = None
    if not args.no_graphflood:
        gflood = graphflood_block(pf, _lib, synth, s, args)
        log(f"[bench] graphflood {gflood}")
    if not args.no_warm:
        warm = warm_start_block(pf, synth, args)
        log(f"[bench] warm start {warm}")
    if not args.no_surrogate:
        mfma = surrogate_block()
        log(f"[bench] surrogate MFMA {mfma}")

    # ---- CPU baseline (oracle, the reference algorithm restated in C) and
    #      full-size parity of the GPU result against it ----------------------
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, HERE)
        from oracle import pff_oracle as O

        g = O.Grid(nx, ny, dx, c["boundary_mode"])
        nthr = O.num_threads()
        log(f"[bench] CPU baseline: oracle with {nthr} threads, one warm-up then the median of 5 full {n}x{n} steps")
        times = []
        for k in range(6):
            t0 = time.perf_counter()
            rcv0, _ = O.compute_receivers(g, c["z"])
            lk = O.reroute_flow(g, c["z"], rcv0, carve=True)
            Qc, done = O.accumulate_constant_Q(g, lk["rec"], c["precipitation"], True)
            times.append(time.perf_counter() - t0)
            log(f"[bench] CPU step {k}: {times[-1]:.2f} s")
        times = times[1:]  # the first run warms the caches and the thread pool
        tc = float(np.median(times))
        cpu = {"value": round(N / tc / 1e6, 3), "unit": "Mcells/s", "cores": nthr, "kind": "port",
               "sample": f"median of 5 full {n}x{n} steps after one warm-up (receivers, reroute_flow carve, accumulate_constant_Q "
                         f"with the reference's rake-and-compress schedule): {', '.join(f'{t:.2f}' for t in times)} s; "
                         f"partly serial: receivers, labels, the saddle pass 1 and accumulation rounds use {nthr} "
                         f"OpenMP threads, saddle passes 2-5 and the carve's tag loop run on one thread"}
        log(f"[bench] CPU baseline {cpu}")
        # parity at full size: one GPU step with the first trip's labels and
        # the per-trip lake outlets recorded, against the oracle's logged run
        router.z.data.copy_(z_copy)
        router.compute_receivers()
        bf = torch.empty(N, dtype=torch.int32, device="cuda")
        router.reroute_flow(carve=True, bid_first=bf, log_outlets=True)
        router.accumulate_constant_Q(c["precipitation"], area=True)
        torch.cuda.synchronize()
        rg = router.receivers.to_numpy()
        Qg = router.Q.to_numpy().astype(np.float64)
        lkl = O.reroute_flow(g, c["z"], rcv0, carve=True, log_iters=16)
        trips = router.last_lake_outlets
        inv_key = 0x4CBEBC200000002A  # key(1e8f, 42), lakeflow.py:127
        outlets_ok = len(trips) == lkl["n_iter"]
        n_logged = 0
        for t, (b, o, sn) in enumerate(trips[:lkl["n_iter"]]):
            ol, snl = lkl["outlet_log"][t], lkl["saddlenode_log"][t]
            rest = np.ones(ol.size, bool)
            rest[b] = False
            outlets_ok &= bool(np.array_equal(o, ol[b]) and np.array_equal(sn, snl[b])
                               and (ol[rest] == inv_key).all() and (snl[rest] == -1).all())
            n_logged += int(b.size)
        del lkl
        cnt = O.upstream_count(g, lk["rec"])
        q0 = float(np.float32(c["precipitation"] * dx * dx))
        ref64 = cnt.astype(np.float64) * q0
        rel = np.abs(Qg - ref64) / ref64
        amax = int(np.argmax(rel))
        parity = {"receivers_bit_exact": bool(np.array_equal(rg, lk["rec"])),
                  "bid_first_bit_exact": bool(np.array_equal(bf.cpu().numpy(), lk["bid_first"])),
                  "outlets_per_trip_bit_exact": bool(outlets_ok), "trips_logged": len(trips),
                  "outlet_entries_logged": n_logged,
                  "ndep_per_trip_equal": [int(router.last_lake_status.ndep[i])
                                          for i in range(router.last_lake_status.n_loop)] == lk["ndep"],
                  "Q_max_rel_err_vs_f64": float(rel.max()), "Q_p99_rel_err_vs_f64": float(np.quantile(rel, 0.99)),
                  "Q_argmax_cell": amax, "Q_max_abs_err_vs_f64": float(np.abs(Qg - ref64).max()),
                  "Q_vs_reference_schedule_max_rel": float((np.abs(Qg - Qc) / np.abs(Qc)).max()),
                  "reference_schedule_converged": bool(done),
                  "div_ambiguity": {"depressions_4096": {
                      "receivers": O.receivers_div_ambiguity(g, c["z"]),
                      "gradient_bits": O.gradient_div_ambiguity(g, c["z"])}}}
        if gflood is not None:
            cg = synth.graphflood_plain()
            g3 = O.Grid(cg["nx"], cg["ny"], cg["dx"], 0)
            parity["div_ambiguity"]["graphflood_2048"] = {"receivers": O.receivers_div_ambiguity(g3, cg["z"]),
                                                          "gradient_bits": O.gradient_div_ambiguity(g3, cg["z"])}
        if batch is not None:
            # every tile of the batch against the oracle's single-grid run
            ra, gb, bad_r, bad_q, worst = 0, 0, 0, 0, 0.0
            q4 = np.float32(1e-5 * 10.0 * 10.0)
            t0 = time.perf_counter()
            for k, t in enumerate(c4):
                g4 = O.Grid(t["nx"], t["ny"], t["dx"], 0)
                ra += O.receivers_div_ambiguity(g4, t["z"])
                gb += O.gradient_div_ambiguity(g4, t["z"])
                r4, _ = O.compute_receivers(g4, t["z"])
                l4 = O.reroute_flow(g4, t["z"], r4, carve=True)
                c4n = O.upstream_count(g4, l4["rec"])
                bad_r += int(not np.array_equal(batch_out[0][k], l4["rec"]))
                bad_q += int(not np.array_equal(batch_out[1][k], q4 * c4n.astype(np.float32)))
                worst = max(worst, float((np.abs(batch_out[1][k].astype(np.float64) - c4n * float(q4))
                                          / (c4n * float(q4))).max()))
            parity["div_ambiguity"]["ragged_256"] = {"receivers": ra, "gradient_bits": gb}
            parity["ragged_256"] = {"tiles": len(c4), "tiles_receivers_differ": bad_r, "tiles_Q_differ": bad_q,
                                    "Q_max_rel_err_vs_f64": worst,
                                    "oracle_seconds": round(time.perf_counter() - t0, 1)}
        log(f"[bench] parity {parity}")

    out = {
        "metric": METRIC,
        "value": round(N / (ms_step * 1e-3) / 1e6, 3),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"depressions {n}x{n} periodic_EW dx=10 m, {pits} Gaussian pits, spectral beta=2.0 "
                               f"relief 1000 m seed 202; receivers + reroute_flow(carve) + accumulate_constant_Q(1e-5)",
                   "cells": N, "stages": stages, "lakeflow": lake, "parity": parity, "receivers_kernel": rcv_roof,
                   "instrumented_kernels": kernel_times, "ragged_batch": batch, "graphflood": gflood,
                   "warm_start": warm, "surrogate_mfma": mfma},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
