#!/usr/bin/env python3
"""Throughput of the DuFort-Frankel time step on one MI355X.

A step is one full DF step (volume march, upward continuation of Bz(0), surface level, and the
receiver gathers the sampling schedule asks for) over the marine_3d workload: 256 x 256 x 160
cells, seeded bathymetry and resistive slab, one x-dipole, 101 receivers.  Inputs are resident in
HBM before the timed region.  The value is FP32 Gcell-updates/s of the single model; the same line
carries the plain-FP64 and the FP32 transmitter-batched throughput, the roofline of the volume
march (HIP events on the library's stream) and the FP64 CPU oracle timed on this host as the
baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM roofline, FP32-batched vs plain FP64 at 1e-4 trace err"
PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, AMD published
WORDS_PER_CELL = 13            # read Ex Ey Ez Bx By Bz sigma, write the six fields


def gpu_run(T, deck, precision, nbatch, steps, warmup):
    import numpy as np
    st = T.make_setup(**deck.setup)
    with T.Context(st, deck.cond, precision) as c:
        r = deck.rxtime
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx)
        txs = [T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(deck.txs)]
        c.set_transmitters(txs)
        c.begin(0, nbatch)
        c.step(warmup)
        c.sync()
        # kernel times come from HIP events recorded on the library's streams inside the timed
        # region (read after it)
        c.enable_timing(True)
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        t1 = time.perf_counter()
        tm = c.timing()
        c.enable_timing(False)
        ex = c.get(T.EX, 0)
    cells = c.L * c.M * c.N
    sec = t1 - t0
    return {
        "gcell_s": cells * nbatch * steps / sec / 1e9,
        "ms_per_step": 1000.0 * sec / steps,
        "timing": tm, "L": c.L, "M": c.M, "N": c.N, "ex_norm": float(np.linalg.norm(ex)),
        "ex": ex,
    }


def roofline(res, esz, nbatch, traffic):
    # the march advances levels 1..N-1 of every column: 13 words per cell, sigma shared by the
    # batch ((12 + 1/B) words per cell-update)
    cells = res["L"] * res["M"] * (res["N"] - 1) * nbatch
    words = (12.0 + 1.0 / nbatch) if nbatch > 1 else WORDS_PER_CELL
    algo_bytes = cells * words * esz
    ms = res["timing"]["march_ms"]
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "kernel": "k_march", "algorithmic_bytes_per_launch": int(algo_bytes),
            "avg_launch_ms": round(ms, 5)}


def cpu_baseline(deck, budget_s=12.0, max_steps=30):
    """The FP64 oracle (C + OpenMP) on this host: whole steps of the same workload."""
    from oracle import oracle as O
    so = O.make_setup(**deck.setup)
    m = O.Model(so, deck.cond)
    r = deck.rxtime
    rt = O.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"])
    lam, dte, dtb, amp = O.schedule(so, m.mindx(), rt, 400)
    tx = O.make_tx(**deck.txs[0])
    tp = m.locate_tx(tx)
    m.step(lam[0], dte[0], amp[0], tp, dtb[0])     # warm-up
    n, t0 = 0, time.perf_counter()
    while n < max_steps:
        m.step(lam[n + 1], dte[n + 1], amp[n + 1], tp, dtb[n + 1])
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    sec = time.perf_counter() - t0
    cells = m.L * m.M * m.N
    return {"value": round(cells * n / sec / 1e9, 5), "unit": "Gcell-updates/s",
            "cores": O.threads(), "kind": "port",
            "sample": "%d FP64 steps of marine_3d %dx%dx%d after 1 warm-up step (oracle, C/OpenMP)"
                      % (n, m.L, m.M, m.N)}


def load_traffic(key):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (multi-GPU sharding is design-only); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import tcsem_fdtd_amd as T
    from tcsem_fdtd_amd import workloads as W

    deck = W.marine_3d(T)
    r32 = gpu_run(T, deck, T.FP32, 1, args.steps, args.warmup)
    r64 = gpu_run(T, deck, T.FP64, 1, args.steps, args.warmup)
    field_err = float(np.linalg.norm(r32["ex"] - r64["ex"]) / max(np.linalg.norm(r64["ex"]), 1e-300))
    tow = W.towline(T, ntx=args.batch)
    rb = gpu_run(T, tow, T.FP32, args.batch, max(args.steps // 5, 1), max(args.warmup // 5, 1))
    line = {
        "metric": METRIC,
        "value": round(r32["gcell_s"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded model, bathymetry and slab; see tcsem_fdtd_amd/workloads.py)",
        "config": {"workload": "marine_3d 256x256x160, 1 x-dipole, 101 receivers, FP32 fields",
                   "cells": r32["L"] * r32["M"] * r32["N"], "air_plane": "512x512",
                   "batch": 1, "parallelism": "single GPU"},
        "roofline": roofline(r32, 4, 1, load_traffic("marine_3d_fp32")),
        "kernel_ms": {k: round(v, 5) for k, v in r32["timing"].items() if k.endswith("_ms")},
        "fp64_single": {"value": round(r64["gcell_s"], 3), "ms_per_step": round(r64["ms_per_step"], 5),
                        "roofline": roofline(r64, 8, 1, load_traffic("marine_3d_fp64")),
                        "kernel_ms": {k: round(v, 5) for k, v in r64["timing"].items() if k.endswith("_ms")}},
        "fp32_batched": {"value": round(rb["gcell_s"], 3), "batch": args.batch,
                         "workload": "towline %d x-dipoles, marine_3d grid" % args.batch,
                         "ms_per_step": round(rb["ms_per_step"], 5),
                         "roofline": roofline(rb, 4, args.batch, load_traffic("towline32_fp32")),
                         "kernel_ms": {k: round(v, 5) for k, v in rb["timing"].items() if k.endswith("_ms")}},
        "ratio_fp32_batched_vs_fp64": round(rb["gcell_s"] / r64["gcell_s"], 3),
        "fp32_vs_fp64_ex_rel_l2_after_run": field_err,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(deck)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
