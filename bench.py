#!/usr/bin/env python3
"""Throughput of the DuFort-Frankel time step on one MI355X.

A step is one full DF step (volume march with the source cells, upward continuation of Bz(0),
surface level, and the receiver gathers the sampling schedule asks for).  The headline workload is
the largest single-GPU configuration, large_5: 512 x 512 x 256 cells, a seeded layered model with
two reservoirs under rough bathymetry, the first of its three x-dipoles, 401 receivers.  Inputs
are resident in HBM before the timed region.  The value is its FP32 Gcell-updates/s over exactly
`--steps` steady-state steps: the source wavelet (the first 200 steps, which also run the source
pass) and then `--warmup` steps run untimed before them.  The same line carries its FP64
throughput over the same window,
the 256 x 256 x 160 marine model in FP32 and FP64, the FP32 transmitter batch (32 x-dipoles on
that grid), 64 conductivity models of uneven grids advanced as one model batch (FP32 and FP64),
complete runs of the small 64 x 64 x 48 model, the roofline of the volume march (HIP events on the
library's stream; with two steps per pass, the pass kernel's), the LDS and HBM rates of the
air-plane chain and of the receiver gathers, and the FP64 CPU oracle timed on this host's cores as
the baseline.  Every GPU run first lets the library's autotuner pick the march tiling for its
batch size, and whether steady-state steps run two per pass.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu] [--no-tune]
    python bench.py --accuracy   # complete batched runs, FP32 vs FP64 trace errors
"""
import argparse
import contextlib
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM roofline, FP32-batched vs plain FP64 at 1e-4 trace err"
PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, AMD published
WORDS_PER_CELL = 13            # read Ex Ey Ez Bx By Bz sigma, write the six fields
# LDS rate assumed for the air-plane FFT figures: 128 B per clock per CU (the CDNA LDS read rate)
# x 256 CUs x 2.4 GHz
PEAK_LDS_GBS = 128 * 256 * 2.4


_roctx = []


@contextlib.contextmanager
def timed_range(label):
    """A rocTX range around a timed region, which `rocprofv3 --marker-trace` records next to the
    kernel dispatches (so a trace of this command can be cut to the timed steps); a no-op when the
    rocTX library is absent."""
    if not _roctx:
        try:
            lib = ctypes.CDLL("librocprofiler-sdk-roctx.so.1")
            lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
            _roctx.append(lib)
        except OSError:
            _roctx.append(None)
    lib = _roctx[0]
    if lib is not None:
        lib.roctxRangePushA(label.encode())
    try:
        yield
    finally:
        if lib is not None:
            lib.roctxRangePop()


def gpu_run(T, deck, precision, nbatch, steps, warmup, tune=True, one_step_too=False):
    import numpy as np
    st = T.make_setup(**deck.setup)
    with T.Context(st, deck.cond, precision) as c:
        # the library's autotuner times every march tiling for this batch size and keeps the
        # fastest (outside the timed region, on zeroed fields)
        tiling = c.autotune(nbatch, 20) if tune else None
        r = deck.rxtime
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx)
        txs = [T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(deck.txs)]
        c.set_transmitters(txs)
        c.begin(0, nbatch)
        # the source wavelet (its 2 (n_src/2) steps also run the source pass) before the warm-up,
        # untimed: the timed steps are steady-state steps, 99 % of a run's steps
        c.step(source_steps(deck))
        c.step(warmup)
        one = None
        rows0, pf0, _ = c.pair_info()
        if one_step_too and rows0:
            # the same model one step at a time (same one-step tiling, two steps per pass off),
            # timed before the window, then the library's choice is restored
            c.set_pairs(0, 1)
            c.step(warmup)
            c.sync()
            u0 = time.perf_counter()
            c.step(steps)
            c.sync()
            u1 = time.perf_counter()
            one = {"value": round(c.L * c.M * c.N * nbatch * steps / (u1 - u0) / 1e9, 3),
                   "ms_per_step": round(1000.0 * (u1 - u0) / steps, 5)}
            c.set_pairs(rows0, pf0)
            c.step(warmup)
        c.sync()
        # kernel times come from HIP events recorded on the library's streams inside the timed
        # region (read after it)
        c.enable_timing(True)
        pairs0 = c.pair_info()[2]
        with timed_range("timed %s fp%d x%d" % (deck.name, precision, nbatch)):
            t0 = time.perf_counter()
            c.step(steps)
            c.sync()
            t1 = time.perf_counter()
        tm = c.timing()
        pass_ms, npass = c.pair_timing()
        rows, pf, pairs1 = c.pair_info()
        c.enable_timing(False)
        ex = c.get(T.EX, 0)
    cells = c.L * c.M * c.N
    sec = t1 - t0
    return {
        "one_step": one,
        "gcell_s": cells * nbatch * steps / sec / 1e9,
        "ms_per_step": 1000.0 * sec / steps,
        "timing": tm, "L": c.L, "M": c.M, "N": c.N, "ex_norm": float(np.linalg.norm(ex)),
        "ex": ex, "tiling": tiling,
        "pairs": {"rows": rows, "pf": pf, "pairs_in_window": pairs1 - pairs0, "pass_ms": pass_ms,
                  "passes_timed": npass} if rows and pairs1 > pairs0 else None,
    }


def source_steps(deck):
    """Steps of the source wavelet (main.cu:223: amp from the waveform while n < 2 (n_src/2))."""
    return 2 * (deck.setup["n_src"] // 2)


def full_run(T, deck, precision, graph=None):
    """A complete run (tcsem_run: every step until the last sample, receivers and all), timed by the
    wall clock around it; graph None keeps the library's default (graph replays of 16 steps for
    grids under 2^21 cells), 0 runs plain steps, k > 0 replays sample-free stretches of k steps."""
    with T.Context(T.make_setup(**deck.setup), deck.cond, precision) as c:
        if graph is not None:
            c.set_graph(graph)
        r = deck.rxtime
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx)
        c.set_transmitters([T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(deck.txs)])
        c.run(0, 1)  # first run: code objects, allocations
        t0 = time.perf_counter()
        c.run(0, 1)
        sec = time.perf_counter() - t0
        steps = c.steps_done
        cells = c.L * c.M * c.N
        replays = c.graph_replays()
    out = {"gcell_s": round(cells * steps / sec / 1e9, 3), "ms_per_step": round(1000.0 * sec / steps, 5),
           "steps": steps, "seconds": round(sec, 4)}
    if replays:
        out.update(graph_replays_per_run=replays // 2)
    return out


def models_run(T, decks, precision, steps, warmup, tune=True, single_steps=0):
    """A model batch: every deck its own context (one transmitter), all advanced together by
    tcsem_batch_step.  single_steps > 0 first times the same contexts stepped one by one
    (tcsem_step per context, each on its own streams, heuristic tiling) for comparison."""
    import numpy as np
    ctxs = []
    for d in decks:
        c = T.Context(T.make_setup(**d.setup), d.cond, precision)
        r = d.rxtime
        c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), d.rx)
        c.set_transmitters([T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(d.txs)])
        ctxs.append(c)
    cells = sum(c.L * c.M * c.N for c in ctxs)
    out = {"models": len(ctxs), "cells": cells}
    try:
        lead = max(source_steps(d) for d in decks)  # the source wavelet, untimed
        if single_steps > 0:
            for c in ctxs:
                c.begin(0, 1)
                c.step(lead + warmup)
            for c in ctxs:
                c.sync()
            t0 = time.perf_counter()
            for c in ctxs:
                c.step(single_steps)
            for c in ctxs:
                c.sync()
            sec = time.perf_counter() - t0
            out["unbatched"] = {"gcell_s": cells * single_steps / sec / 1e9,
                                "ms_per_step": 1000.0 * sec / single_steps, "steps": single_steps}
        with T.Batch(ctxs) as b:
            out["tiling"] = b.autotune(5) if tune else None
            b.step(lead + warmup)
            b.sync()
            b.enable_timing(True)
            with timed_range("timed model batch fp%d x%d" % (precision, len(ctxs))):
                t0 = time.perf_counter()
                b.step(steps)
                b.sync()
                sec = time.perf_counter() - t0
            out["timing"] = b.timing()
            pass_ms, npass = b.pair_timing()
            rows, _, pairs = b.pair_info()
            out["pairs"] = {"rows": rows, "pass_ms": pass_ms, "passes_timed": npass} if rows and npass else None
            b.enable_timing(False)
            out["ex"] = np.concatenate([c.get(T.EX, 0).ravel() for c in ctxs])
        out["gcell_s"] = cells * steps / sec / 1e9
        out["ms_per_step"] = 1000.0 * sec / steps
        # march bytes: levels 1..N-1 of every member, 13 words per cell
        out["march_cells"] = sum(c.L * c.M * (c.N - 1) for c in ctxs)
    finally:
        for c in ctxs:
            c.close()
    return out


def models_roofline(res, esz, traffic_entry=(None, None)):
    step_ms = res["timing"]["march_ms"]
    pr = res.get("pairs")
    if pr:
        # the model batch's two-step pass: 11 words per cell of every member per launch (each
        # member reads its own sigma), for two cell-updates
        algo = res["march_cells"] * 11 * esz
        ms = pr["pass_ms"]
        gbs = algo / (ms * 1e-3) / 1e9
        equiv = res["march_cells"] * WORDS_PER_CELL * esz / (step_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": traffic_entry[0],
                "traffic_tiling": traffic_entry[1],
                "kernel": "k_march2 (model batch, two steps per pass, rows %d)" % pr["rows"],
                "algorithmic_bytes_per_launch": int(algo), "cell_updates_per_launch": 2 * res["march_cells"],
                "avg_launch_ms": round(ms, 5), "march_ms_per_step": round(step_ms, 5),
                "one_step_equivalent": {"GBs": round(equiv, 1), "frac": round(equiv / PEAK_HBM_GBS, 4)}}
    algo = res["march_cells"] * WORDS_PER_CELL * esz
    gbs = algo / (step_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": traffic_entry[0],
            "traffic_tiling": traffic_entry[1], "kernel": "k_march (model batch)",
            "algorithmic_bytes_per_launch": int(algo), "avg_launch_ms": round(step_ms, 5)}


def roofline(res, esz, nbatch, traffic_entry):
    # the one-step march advances levels 1..N-1 of every column: 13 words per cell, sigma shared by
    # the batch ((12 + 1/B) words per cell-update)
    cells = res["L"] * res["M"] * (res["N"] - 1) * nbatch
    words = (12.0 + 1.0 / nbatch) if nbatch > 1 else WORDS_PER_CELL
    traffic, traffic_tiling = traffic_entry if isinstance(traffic_entry, tuple) else (traffic_entry, None)
    pr = res.get("pairs")
    if pr and pr["passes_timed"] > 0:
        # the two-step pass (k_march2) dominates: per launch it reads sigma and the old state and
        # writes the new state of every cell once for two cell-updates, 10 words + sigma (11, or
        # 10 + 1/B shared by the batch); its top launches (levels 2 and 1) are separate launches
        pwords = 10.0 + 1.0 / nbatch
        algo = cells * pwords * esz
        ms = pr["pass_ms"]
        achieved = algo / (ms * 1e-3) / 1e9
        step_ms = res["timing"]["march_ms"]
        equiv = cells * words * esz / (step_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "traffic_tiling": traffic_tiling,
                "kernel": "k_march2 (two steps per pass, rows %d, prefetch %d)" % (pr["rows"], pr["pf"]),
                "algorithmic_bytes_per_launch": int(algo), "cell_updates_per_launch": 2 * cells,
                "avg_launch_ms": round(ms, 5),
                "march_ms_per_step": round(step_ms, 5),
                "one_step_equivalent": {"GBs": round(equiv, 1), "frac": round(equiv / PEAK_HBM_GBS, 4),
                                        "note": "the one-step march's %.4g words per cell-update over the "
                                                "march's time per step (pass and top launches, halved)" % words},
                "span": "per launch of k_march2; march_ms_per_step adds the top launches of both steps"}
    algo_bytes = cells * words * esz
    ms = res["timing"]["march_ms"]
    if not ms > 0:  # no step of the window was timed
        return None
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "traffic_tiling": traffic_tiling,
            "kernel": "k_march", "algorithmic_bytes_per_launch": int(algo_bytes),
            "avg_launch_ms": round(ms, 5),
            "span": "per step: the lower and the top launch of k_march, the lower one from the start of its k_zsum pass when the tiling runs one"}


def traffic_key(key, res):
    """The counter summary's key for a run: the two-step pass has entries of its own, one per pass
    shape where that shape was counted (key_pairs_rows,pf), else one for the workload."""
    pr = res.get("pairs")
    if not pr:
        return key
    if pr.get("pf"):
        exact = "%s_pairs_%d,%d" % (key, pr["rows"], pr["pf"])
        if load_traffic(exact)[0] is not None:
            return exact
    return key + "_pairs"


def _fft_len(n):
    """Working length of a line FFT: n itself for powers of two, else the Bluestein length."""
    need = n if n & (n - 1) == 0 else 2 * n - 1
    m = 1
    while m < need:
        m <<= 1
    return m


def _fft_lds_entries(n, m):
    """LDS entries (complex) read and written by one line FFT of length n on working length m:
    every Stockham pass reads and writes the line, passes after the first read (R-1)/R m
    twiddles; Bluestein adds the chirp products around two transforms of length m."""
    def pow2(m):
        tot, left, first = 0, m.bit_length() - 1, True
        while left > 0:
            r = 8 if left >= 3 else (4 if left == 2 else 2)
            left -= {8: 3, 4: 2, 2: 1}[r]
            tot += 2 * m + (0 if first else (r - 1) * m // r)
            first = False
        return tot
    if n == m:
        return pow2(m)
    return 2 * pow2(m) + (n + m) + 2 * m + 2 * n


def upcont_model(deck, esz):
    """Algorithmic LDS and HBM bytes of one upward continuation (tiled chain, in the field
    precision esz), one slot."""
    import tcsem_fdtd_amd as T
    s = deck.setup
    g = T.grid_info(T.make_setup(**s))
    L, M, L1, M1 = g.L, g.M, g.L1, g.M1
    H = L1 // 2 + 1
    mx, my = _fft_len(L1), _fft_len(M1)
    n1 = s["ALx1"] + s["ALx2"] + 4
    n2 = s["ARx1"] + s["ARx2"] + 4
    n3 = s["Ay1"] + s["Ay2"] + 4
    aL, aR, aY, Nx, Ny = g.AnLx_New, g.AnRx_New, g.Any_New, s["Nx"], s["Ny"]
    npair = (M + 1) // 2
    c = 2 * esz  # bytes per complex entry: the chain computes in the field precision
    k1 = npair * (2 * L * esz + 2 * (Nx + aL * n1 + aR * n2) * esz + L1 * c + _fft_lds_entries(L1, mx) * c + 6 * H * c)
    k2 = H * (M * c + (Ny + 2 * aY * n3) * c + M1 * c + 3 * _fft_lds_entries(M1, my) * c + 5 * M1 * c)
    k3 = M1 * (L1 * c + _fft_lds_entries(L1, mx) * c + L1 * c)
    nair = M * (L + 1) + (M + 1) * L
    hbm = (M * L * esz + H * M * c) + (H * M * c + H * M1 * 2 * c) + (H * M1 * 2 * c + M1 * L1 * c) + \
          (4 * nair * esz + nair * esz)
    return {"lds_bytes": int(k1 + k2 + k3), "hbm_bytes": int(hbm), "plane": "%dx%d" % (L1, M1)}


def gather_model(deck, esz):
    """Algorithmic bytes of one receiver-gather launch pair (E group and B group) over all the
    deck's receivers: 4 x 4 x 4 stencil values per component, and for Bz the Bx/By columns
    under the 4 x 4 stencil columns (5 x 4 + 4 x 5 values per level) from the bottom."""
    import numpy as np
    import tcsem_fdtd_amd as T
    st = T.make_setup(**deck.setup)
    g = T.grid_info(st)
    Z = T.grid_arrays(st)["Z"]
    comps = deck.rxtime["comps"]
    ne = sum(comps[:3])
    nb2 = sum(comps[3:5])
    total = 0
    for x, y, z in np.asarray(deck.rx):
        k = int(np.clip(np.searchsorted(Z, z) - 2, 0, g.N - 4))
        total += (ne + nb2) * 64 * esz
        if comps[5]:
            total += 64 * esz + 40 * (g.N - k) * esz
    return total


def side_kernels(res, deck, esz, nbatch):
    """Upward continuation (LDS and HBM rate of the chain) and receiver gathers (HBM rate) from
    the measured in-situ kernel times and the algorithmic byte models."""
    tm = res["timing"]
    up = upcont_model(deck, esz)
    out = {}
    if tm["upcont_ms"] > 0:
        sec = tm["upcont_ms"] * 1e-3
        out["upcont"] = {"plane": up["plane"], "ms": round(tm["upcont_ms"], 5),
                         "lds_bytes": up["lds_bytes"] * nbatch, "hbm_bytes": up["hbm_bytes"] * nbatch,
                         "lds_GBs": round(up["lds_bytes"] * nbatch / sec / 1e9, 1),
                         "lds_frac": round(up["lds_bytes"] * nbatch / sec / 1e9 / PEAK_LDS_GBS, 4),
                         "hbm_GBs": round(up["hbm_bytes"] * nbatch / sec / 1e9, 1)}
    if tm["gather_ms"] > 0:
        per = gather_model(deck, esz) * nbatch / 2.0  # an E launch and a B launch per sample event
        gb = per / (tm["gather_ms"] * 1e-3) / 1e9
        out["gather"] = {"ms_per_launch": round(tm["gather_ms"], 5), "bytes_per_launch": int(per),
                         "hbm_GBs": round(gb, 1), "hbm_frac": round(gb / PEAK_HBM_GBS, 4)}
    return out


def host_info():
    """nproc, the CPU model name and the physical core count of this host (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        cps, sockets = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sockets:
            info["physical_cores"] = int(cps) * int(sockets)
        info["threads_per_core"] = kv.get("Thread(s) per core")
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        info["nproc_affinity"] = len(os.sched_getaffinity(0))
    except OSError:
        pass
    return info


def cpu_window(O, deck, nsteps, warmup=5):
    """The FP64 oracle on this host: `warmup` untimed then `nsteps` timed DuFort-Frankel steps of
    the deck's first transmitter from the start of its run, with the reference's CPU formulation of
    the air boundary (SplineGridConvIR2R_cpu, the UpContinuation filter, BilinearGridConvR2IR)."""
    so = O.make_setup(**deck.setup)
    m = O.Model(so, deck.cond)
    m.set_cpu_twins(True)
    r = deck.rxtime
    lam, dte, dtb, amp = O.schedule(so, m.mindx(), O.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]),
                                    warmup + nsteps)
    tp = m.locate_tx(O.make_tx(**deck.txs[0]))
    for n in range(warmup):
        m.step(lam[n], dte[n], amp[n], tp, dtb[n])
    t0 = time.perf_counter()
    for n in range(warmup, warmup + nsteps):
        m.step(lam[n], dte[n], amp[n], tp, dtb[n])
    sec = time.perf_counter() - t0
    cells = m.L * m.M * m.N
    return {"gcell_s": round(cells * nsteps / sec / 1e9, 5), "steps": nsteps, "warmup": warmup,
            "seconds": round(sec, 3), "cells": cells}


def cpu_full_run(O, deck):
    """A complete run of the deck's first transmitter on the FP64 oracle (every step to the last
    sample, receivers and all), CPU-formulation air boundary, timed by the wall clock."""
    so = O.make_setup(**deck.setup)
    m = O.Model(so, deck.cond)
    m.set_cpu_twins(True)
    r = deck.rxtime
    t0 = time.perf_counter()
    _, steps = m.run(O.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), deck.rx,
                     O.make_tx(**deck.txs[0]))
    sec = time.perf_counter() - t0
    cells = m.L * m.M * m.N
    return {"gcell_s": round(cells * steps / sec / 1e9, 5), "steps": steps, "seconds": round(sec, 3),
            "cells": cells}


def cpu_baseline(W, T, window=50):
    """The CPU path timed on this host's cores: the FP64 C/OpenMP oracle (a restatement of the
    reference's kernels; the reference's own CUDA path cannot run here) with OpenMP threads bound
    close.  The small 64 x 64 x 48 model complete; the 256 x 256 x 160 and 512 x 512 x 256 models
    over `window` steps after 5 warm-up steps.  `value` is the 512 x 512 x 256 window (the
    headline's workload)."""
    host = host_info()  # before the first OpenMP region binds this thread to one core
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    out = {"marine_1d_complete": cpu_full_run(O, W.marine_1d(T)),
           "marine_3d_window": cpu_window(O, W.marine_3d(T), window),
           "large_5_window": cpu_window(O, W.large_5(T), window)}
    return {"value": out["large_5_window"]["gcell_s"], "unit": "Gcell-updates/s",
            "cores": O.threads(), "kind": "port",
            "sample": "FP64 oracle (C/OpenMP, OMP_PROC_BIND=%s), CPU-formulation air boundary: "
                      "large_5 512x512x256 %d steps after 5 warm-up (value); marine_3d 256x256x160 %d steps "
                      "after 5 warm-up; marine_1d 64x64x48 complete run" % (os.environ["OMP_PROC_BIND"], window,
                                                                          window),
            "host": host, "configs": out}


def trace_error(got, ref, floor=1e-6):
    """Per receiver: ||a-b|| / ||b|| over all times and components, components below floor x the
    receiver's largest component norm excluded (symmetry nulls)."""
    import numpy as np
    out = []
    for r in range(ref.shape[0]):
        norms = np.linalg.norm(ref[r], axis=1)
        keep = norms >= floor * norms.max()
        nb = np.linalg.norm(ref[r][keep])
        out.append(np.linalg.norm(got[r][keep] - ref[r][keep]) / (nb if nb > 0 else 1.0))
    return np.array(out)


def _progress(msg):
    print("[accuracy] " + msg, file=sys.stderr, flush=True)


def accuracy(T, W):
    """Complete runs of the batched workloads (32-transmitter towline, 64 uneven models) with FP32
    and with FP64 fields: per-receiver relative trace error of FP32 against FP64, and the wall
    time of each complete run.  Progress goes to stderr every few hundred steps."""
    import numpy as np
    out = {}
    tow = W.towline(T, ntx=32)
    res, sec = {}, {}
    for prec in (T.FP64, T.FP32):
        with T.Context(T.make_setup(**tow.setup), tow.cond, prec) as c:
            r = tow.rxtime
            c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), tow.rx)
            c.set_transmitters([T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(tow.txs)])
            c.autotune(32, 10)
            t0 = time.perf_counter()
            c.begin(0, 32)
            steps = 0
            while T.lib().tcsem_steps_remaining(c.h) > 0:  # chunks of 200 steps up to the last sample
                c.step(200)
                c.sync()
                steps += 200
                _progress("towline FP%d: %d steps, %.0f s" % (prec, steps, time.perf_counter() - t0))
            res[prec] = [c.traces(b) for b in range(32)]
            sec[prec] = time.perf_counter() - t0
    e = np.concatenate([trace_error(res[T.FP32][b], res[T.FP64][b]) for b in range(32)])
    out["towline32"] = {"steps_stepped": steps, "receivers": int(e.size), "max": float(e.max()), "median": float(np.median(e)),
                        "fp32_run_s": round(sec[T.FP32], 1), "fp64_run_s": round(sec[T.FP64], 1)}
    models = W.uneven_models(T)
    res, sec = {}, {}
    for prec in (T.FP64, T.FP32):
        ctxs = []
        for d in models:
            c = T.Context(T.make_setup(**d.setup), d.cond, prec)
            r = d.rxtime
            c.set_receivers(T.make_rxtime(r["T0"], r["Tn"], r["Nt"], r["is_log"], r["comps"]), d.rx)
            c.set_transmitters([T.make_tx(**dict(t, txid=i + 1)) for i, t in enumerate(d.txs)])
            ctxs.append(c)
        with T.Batch(ctxs) as b:
            b.autotune(5)
            t0 = time.perf_counter()
            res[prec] = b.run()
            sec[prec] = time.perf_counter() - t0
            steps = b.steps_done
            _progress("uneven models FP%d: %d steps, %.0f s" % (prec, steps, sec[prec]))
        for c in ctxs:
            c.close()
    e = np.concatenate([trace_error(res[T.FP32][q], res[T.FP64][q]) for q in range(len(models))])
    out["uneven_models64"] = {"steps": steps, "receivers": int(e.size), "max": float(e.max()),
                              "median": float(np.median(e)), "fp32_run_s": round(sec[T.FP32], 1),
                              "fp64_run_s": round(sec[T.FP64], 1)}
    return out


def load_traffic(key):
    """HBM bytes per step of k_march from the committed counter summary (profiles/traffic.json),
    with the march tiling they were counted for; (None, None) when absent."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            v = json.load(f).get(key)
    except (OSError, ValueError):
        return None, None
    if isinstance(v, dict):
        return v.get("bytes"), v.get("tiling")
    return v, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    # untimed steps after the source wavelet (the first n_src = 200 steps of these decks, where a
    # step also runs the source passes, which always run untimed first): the timed steps are the
    # steady-state steps that make up 99 % of a run (18 802 steps to the last sample of the
    # 256 x 256 x 160 model)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="heuristic march tiling instead of the autotuner")
    ap.add_argument("--accuracy", action="store_true",
                    help="instead of the throughput line: complete batched runs, FP32 vs FP64 trace errors (minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (multi-GPU sharding is design-only); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import tcsem_fdtd_amd as T
    from tcsem_fdtd_amd import workloads as W

    if args.accuracy:
        print(json.dumps({"accuracy": "FP32 vs FP64 per-receiver relative trace error, complete runs",
                          **accuracy(T, W)}), flush=True)
        return
    tune = not args.no_tune
    kern = lambda r: {k: round(v, 5) for k, v in r["timing"].items() if k.endswith("_ms")}
    # the headline: the large single-GPU model (512 x 512 x 256 cells, 1024 x 1024 air plane) in
    # FP32, exactly `steps` timed steps after `warmup`; the same window in FP64 beside it
    big = W.large_5(T)
    rl32 = gpu_run(T, big, T.FP32, 1, args.steps, args.warmup, tune)
    rl64 = gpu_run(T, big, T.FP64, 1, args.steps, args.warmup, tune)
    big_err = float(np.linalg.norm(rl32["ex"] - rl64["ex"]) / max(np.linalg.norm(rl64["ex"]), 1e-300))
    rl32.pop("ex"), rl64.pop("ex")
    # the 256 x 256 x 160 marine model, FP32 and FP64, one transmitter
    deck = W.marine_3d(T)
    r32 = gpu_run(T, deck, T.FP32, 1, args.steps, args.warmup, tune)
    r64 = gpu_run(T, deck, T.FP64, 1, args.steps, args.warmup, tune, one_step_too=True)
    field_err = float(np.linalg.norm(r32["ex"] - r64["ex"]) / max(np.linalg.norm(r64["ex"]), 1e-300))
    r32.pop("ex"), r64.pop("ex")
    # 32 x-dipoles along a towline on the same grid, FP32 batched
    tow = W.towline(T, ntx=args.batch)
    nb = max(args.steps // 2, 8)  # at least one timed pass or step (every 8th step is timed)
    rb = gpu_run(T, tow, T.FP32, args.batch, nb, args.warmup, tune)
    rb.pop("ex")
    # 64 conductivity models of uneven grids (96^3 .. 224x224x160), one x-dipole each, in one
    # model batch: FP32 and FP64, FP32 also stepped model by model for comparison
    models = W.uneven_models(T)
    nm, wm = max(args.steps // 2, 8), args.warmup  # at least one timed pass or step
    m32 = models_run(T, models, T.FP32, nm, wm, tune, single_steps=max(nm // 2, 1))
    m64 = models_run(T, models, T.FP64, nm, wm, tune)
    models_err = float(np.linalg.norm(m32["ex"] - m64["ex"]) / max(np.linalg.norm(m64["ex"]), 1e-300))
    m32.pop("ex"), m64.pop("ex")
    # the small 64 x 64 x 48 marine model: complete runs (host-launch-bound: a few us of GPU work
    # per step)
    small = W.marine_1d(T)
    s32, s64 = full_run(T, small, T.FP32), full_run(T, small, T.FP64)
    s32p = full_run(T, small, T.FP32, graph=0)
    line = {
        "metric": METRIC,
        "value": round(rl32["gcell_s"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rl32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model, reservoirs and rough bathymetry; see tcsem_fdtd_amd/workloads.py)",
        "config": {"workload": "large_5 512x512x256, 1024x1024 air plane, x-dipole 1 of 3, 401 receivers, FP32 fields",
                   "cells": rl32["L"] * rl32["M"] * rl32["N"], "batch": 1, "parallelism": "single GPU",
                   "march_tiling": rl32["tiling"]},
        "roofline": roofline(rl32, 4, 1, load_traffic(traffic_key("large_5_fp32", rl32))),
        "kernel_ms": kern(rl32),
        "side_kernels": side_kernels(rl32, big, 4, 1),
        "fp64_single": {"workload": "large_5 512x512x256, FP64 fields, the same window",
                        "value": round(rl64["gcell_s"], 3), "ms_per_step": round(rl64["ms_per_step"], 5),
                        "roofline": roofline(rl64, 8, 1, load_traffic(traffic_key("large_5_fp64", rl64))),
                        "march_tiling": rl64["tiling"], "kernel_ms": kern(rl64),
                        "fp32_vs_fp64_ex_rel_l2_after_window": big_err},
        "marine_3d": {"workload": "marine_3d 256x256x160, 512x512 air plane, 1 x-dipole, 101 receivers",
                      "steps": args.steps, "warmup": args.warmup,
                      "fp32": {"value": round(r32["gcell_s"], 3), "ms_per_step": round(r32["ms_per_step"], 5),
                               "roofline": roofline(r32, 4, 1, load_traffic(traffic_key("marine_3d_fp32", r32))),
                               "march_tiling": r32["tiling"], "kernel_ms": kern(r32),
                               "side_kernels": side_kernels(r32, deck, 4, 1)},
                      "fp64": {"value": round(r64["gcell_s"], 3), "ms_per_step": round(r64["ms_per_step"], 5),
                               "roofline": roofline(r64, 8, 1, load_traffic(traffic_key("marine_3d_fp64", r64))),
                               "march_tiling": r64["tiling"], "kernel_ms": kern(r64),
                               "one_step": r64["one_step"]},
                      "fp32_vs_fp64_ex_rel_l2_after_window": field_err},
        "fp32_batched": {"value": round(rb["gcell_s"], 3), "batch": args.batch,
                         "workload": "towline %d x-dipoles, marine_3d grid" % args.batch,
                         "steps": nb, "warmup": args.warmup,
                         "ms_per_step": round(rb["ms_per_step"], 5),
                         "roofline": roofline(rb, 4, args.batch, load_traffic(traffic_key("towline32_fp32", rb))),
                         "march_tiling": rb["tiling"], "kernel_ms": kern(rb),
                         "side_kernels": side_kernels(rb, tow, 4, args.batch)},
        "uneven_models": {"workload": "64 models, grids 96^3 .. 224x224x160 (%d cells), 1 x-dipole and 41 receivers each"
                                      % m32["cells"], "steps": nm, "warmup": wm,
                          "fp32": {"value": round(m32["gcell_s"], 3), "ms_per_step": round(m32["ms_per_step"], 5),
                                   "roofline": models_roofline(m32, 4, load_traffic(traffic_key("uneven64_fp32", m32))),
                                   "march_tiling": m32["tiling"], "kernel_ms": kern(m32),
                                   "unbatched": {k: round(v, 5) if isinstance(v, float) else v
                                                 for k, v in m32["unbatched"].items()}},
                          "fp64": {"value": round(m64["gcell_s"], 3), "ms_per_step": round(m64["ms_per_step"], 5),
                                   "roofline": models_roofline(m64, 8, load_traffic(traffic_key("uneven64_fp64", m64))),
                                   "march_tiling": m64["tiling"],
                                   "kernel_ms": kern(m64)},
                          "fp32_vs_fp64_ex_rel_l2_after_window": models_err},
        "marine_1d": {"workload": "marine_1d 64x64x48, 1 x-dipole, 21 receivers, complete run (tcsem_run, "
                                  "graph replays of 16 steps: the library's default for this grid)",
                      "fp32": s32, "fp64": s64, "fp32_plain_steps": s32p},
        # the north-star ratio: FP32 transmitter batch against the plain FP64 single model (the
        # 256 x 256 x 160 grid they share, and the large model's FP64 window)
        "ratio_fp32_batched_vs_fp64": round(rb["gcell_s"] / r64["gcell_s"], 3),
        "ratio_fp32_batched_vs_fp64_large_5": round(rb["gcell_s"] / rl64["gcell_s"], 3),
        # and against the same FP64 model one step at a time, when it takes two steps per pass
        "ratio_fp32_batched_vs_fp64_one_step": (round(rb["gcell_s"] / r64["one_step"]["value"], 3)
                                                if r64["one_step"] else None),
    }
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(W, T)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
