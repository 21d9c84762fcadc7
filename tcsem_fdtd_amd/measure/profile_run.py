"""Fixed-step run of one workload for rocprofv3 (kernel traces, counters).

    python -m tcsem_fdtd_amd.measure.profile_run --workload marine_3d --precision 32 --batch 1 --steps 20

Builds the deck, runs the library's march autotuner for the batch size (as bench.py does; its
trial steps precede the field resets of tcsem_begin, the last k_fill_zero dispatches), advances
through the source wavelet (the first 2 (n_src/2) steps, which also run the source pass) and
`warmup` more steps untimed, then `steps` steps, and prints one JSON line with the chosen tiling
and the wall time per step.  --pairs rows,pf (or 0) fixes the two-step pass over the library's
choice.
--workload uneven runs the uneven-model workload as one model batch (its autotuner, the field
reset it leaves, then the same source wavelet, warm-up and timed steps; --pairs rows, or 0, fixes
the batch's two-step pass).  --every-step-rx replaces the deck's
sample times by one sample per step from the first step on,
so that every step runs the receiver gathers of all receivers and components.  Used by the
rocprof summaries under profiles/.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))


def main():
    import numpy as np
    import tcsem_fdtd_amd as T
    from tcsem_fdtd_amd import workloads as W
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="marine_3d",
                    choices=["marine_1d", "marine_3d", "towline", "large_5", "uneven"])
    ap.add_argument("--models", type=int, default=64, help="uneven: number of models in the model batch")
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64])
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--every-step-rx", action="store_true")
    ap.add_argument("--pairs", default=None, help="two steps per pass: rows,pf, or 0 for one step at a time")
    ap.add_argument("--tiling", default=None,
                    help="fixed march tiling rows,pf[,zchunks[,pass|fused[,1|0]]] instead of the autotuner "
                         "(the last field: the 8-bit conductivity index on or off)")
    a = ap.parse_args()
    if a.tiling:
        f = a.tiling.split(",")
        os.environ["TCSEM_MARCH_ROWS"], os.environ["TCSEM_MARCH_PF"] = f[0], f[1]
        if len(f) > 2:
            os.environ["TCSEM_MARCH_ZCHUNKS"] = f[2]
        if len(f) > 3:
            os.environ["TCSEM_MARCH_ZSUM"] = f[3]
        if len(f) > 4:
            os.environ["TCSEM_SIGMA8"] = f[4]
        a.no_tune = True
    if a.workload == "uneven":
        return models(T, W, a)
    deck = W.towline(T, ntx=a.batch) if a.workload == "towline" else getattr(W, a.workload)(T)
    with T.Context(T.make_setup(**deck.setup), deck.cond, a.precision) as c:
        if a.pairs is not None:
            f = [int(x) for x in a.pairs.split(",")]
            c.set_pairs(f[0], f[1] if len(f) > 1 else 1)
        tiling = None if a.no_tune else c.autotune(a.batch, 20)
        r = dict(deck.rxtime)
        if a.every_step_rx:
            # linear samples about four per step over the run (the source sub-steps), so that
            # (nearly) every step gathers E and B at every receiver
            n = a.warmup + a.steps
            ns = deck.setup["n_src"] // 2
            if n > 2 * ns:
                raise SystemExit("--every-step-rx needs warmup + steps <= %d (the source sub-steps)" % (2 * ns))
            t_end = float(np.cumsum(T.waveform(ns, deck.setup["sigma_t"])[1])[n - 1])
            r.update(T0=1e-12, Tn=t_end, Nt=4 * n, is_log=0)
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx)
        c.set_transmitters([T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(deck.txs)])
        c.begin(0, a.batch)
        lead = 0 if a.every_step_rx else 2 * (deck.setup["n_src"] // 2)
        c.step(lead + a.warmup)
        c.sync()
        p0 = c.pair_info()[2]
        t0 = time.perf_counter()
        c.step(a.steps)
        c.sync()
        dt = time.perf_counter() - t0
        rows, pf, p1 = c.pair_info()
        print(json.dumps({"workload": deck.name, "precision": a.precision, "batch": a.batch,
                          "steps": a.steps, "ms_per_step": 1000 * dt / a.steps,
                          "cells": c.L * c.M * c.N, "march_tiling": tiling,
                          "pairs": [rows, pf, p1 - p0] if rows else None}), flush=True)


def models(T, W, a):
    """The uneven-model workload as one model batch (one context and one transmitter per model)."""
    decks = W.uneven_models(T, n=a.models)
    ctxs = []
    for d in decks:
        c = T.Context(T.make_setup(**d.setup), d.cond, a.precision)
        r = d.rxtime
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), d.rx)
        c.set_transmitters([T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(d.txs)])
        ctxs.append(c)
    with T.Batch(ctxs) as b:
        if a.pairs is not None:
            b.set_pairs(int(a.pairs.split(",")[0]))
        tiling = None if a.no_tune else b.autotune(10)
        # the autotuner zeroes the members' fields; restart their runs
        for c in ctxs:
            c.reset(0, 1)
        lead = max(2 * (d.setup["n_src"] // 2) for d in decks)
        b.step(lead + a.warmup)
        b.sync()
        p0 = b.pair_info()[2]
        t0 = time.perf_counter()
        b.step(a.steps)
        b.sync()
        dt = time.perf_counter() - t0
        rows, _, p1 = b.pair_info()
    cells = sum(c.L * c.M * c.N for c in ctxs)
    for c in ctxs:
        c.close()
    print(json.dumps({"workload": "uneven_models(%d)" % a.models, "precision": a.precision,
                      "steps": a.steps, "ms_per_step": 1000 * dt / a.steps, "cells": cells,
                      "march_tiling": tiling, "pairs": [rows, p1 - p0] if rows else None}), flush=True)


if __name__ == "__main__":
    main()
