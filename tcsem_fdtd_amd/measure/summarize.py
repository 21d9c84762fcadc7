"""Markdown summaries of rocprofv3 runs of the march (kernel trace, then HBM counters).

    python -m tcsem_fdtd_amd.measure.summarize march --title "..." --trace DIR --trace-cmd "..." \
        --cells N --esz 4 [--fetch DIR --write DIR --pmc-cmd "..." \
        --copy-fetch DIR --copy-write DIR --copy-mib 256 --key large_5_fp32 --tiling 10,2,1,fused \
        --traffic profiles/traffic.json] [--counters DIR ...] > profiles/k_march_large_5.md

    python -m tcsem_fdtd_amd.measure.summarize bench --trace DIR --cmd "..." --line BENCH_JSON \
        [--compare "LABEL=json.path"] > profiles/bench_large_5.md

The kernels form prints a trace table and, per kernel-name pattern, the launches, their average
duration and calibrated HBM bytes per launch.

The bench form cuts a `--kernel-trace --marker-trace` run of bench.py to the rocTX ranges around
its timed regions (one per workload) and sets each range's march per step beside the bench line's
own figure at the given JSON path (its avg_launch_ms).

DIR is a rocprofv3 output folder (`-d DIR -o run --output-format csv`).  Dispatches before the
last k_fill_zero (the field reset of tcsem_begin, after the autotuner's trial steps) are left out.
Steps are counted as k_surface dispatches (one per step: the warm-up steps after the reset and
the timed ones).  The march per step is the sum of its lower and top launches; its algorithmic bytes
are the canonical L*M*(N-1) cell-updates x 13 words x element size.  When the run takes two steps
per pass (k_march2), the pass is also reported per launch: its algorithmic bytes are L*M*(N-1)
cells x --pass-words (10 + 1/B: sigma and the old state read, the new state written, once for two
cell-updates) x element size, and with --key its calibrated HBM bytes per launch go to the
traffic file (the counters' k_march total per step is still printed).  FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) are scaled by what they read for a copy kernel of known size (stream_probe copy).
"""
import argparse
import csv
import glob
import gzip
import json
import os

WORDS = 13


def _csv(d, suffix):
    """Rows of the first *suffix (or *suffix.gz) file under d."""
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True) +
                  glob.glob(os.path.join(d, "**", "*" + suffix + ".gz"), recursive=True))
    if not hits:
        raise SystemExit("no *%s under %s" % (suffix, d))
    opener = gzip.open if hits[0].endswith(".gz") else open
    with opener(hits[0], "rt") as f:
        return list(csv.DictReader(f))


def _after_fill(rows, name_key, start_key):
    rows = sorted(rows, key=lambda r: int(r[start_key]))
    last = -1
    for q, r in enumerate(rows):
        if "k_fill_zero" in r[name_key]:
            last = q
    return rows[last + 1:]


def trace_table(d):
    """Per-kernel calls, average and total duration [us] of the dispatches after the last fill."""
    rows = _after_fill(_csv(d, "kernel_trace.csv"), "Kernel_Name", "Start_Timestamp")
    acc = {}
    for r in rows:
        ns = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        a = acc.setdefault(r["Kernel_Name"], [0, 0])
        a[0] += 1
        a[1] += ns
    return sorted(((k, n, t / n / 1e3, t / 1e3) for k, (n, t) in acc.items()), key=lambda x: -x[3])


def counter_sum(d, counter, match):
    """Sum of `counter` over the dispatches after the last fill whose kernel name contains `match`;
    the number of such dispatches."""
    rows = [r for r in _csv(d, "counter_collection.csv") if r["Counter_Name"] == counter]
    rows = _after_fill(rows, "Kernel_Name", "Start_Timestamp")
    hit = [float(r["Counter_Value"]) for r in rows if match in r["Kernel_Name"]]
    return sum(hit), len(hit)


def counter_table(d):
    """Every counter summed per kernel over the dispatches after the last fill."""
    rows = _after_fill(_csv(d, "counter_collection.csv"), "Kernel_Name", "Start_Timestamp")
    acc = {}
    for r in rows:
        key = (r["Kernel_Name"], r["Counter_Name"])
        acc[key] = acc.get(key, 0.0) + float(r["Counter_Value"])
    return acc


def _short(name, n=100):
    return name if len(name) <= n else name[:n]


def march(a):
    out = ["# " + a.title, ""]
    algo = a.cells * a.words * a.esz
    if a.trace:
        out.append("Kernel trace: `%s`" % a.trace_cmd)
        out.append("")
        out.append("| kernel | calls | avg us | total us |")
        out.append("|---|---|---|---|")
        tab = trace_table(a.trace)
        for k, n, avg, tot in tab:
            out.append("| `%s` | %d | %.2f | %.1f |" % (_short(k), n, avg, tot))
        out.append("")
        steps = sum(n for k, n, avg, tot in tab if "k_surface" in k)
        march_us = sum(tot for k, n, avg, tot in tab if "k_march" in k) / steps
        zsum_us = sum(tot for k, n, avg, tot in tab if "k_zsum" in k) / steps
        span = march_us + zsum_us
        out.append("- k_march per step (lower + top launch%s): %.2f us over %d steps"
                   % (", and the k_zsum passes of the steps that run one" if zsum_us else "", span, steps))
        out.append("- algorithmic bytes per step: %d cell-updates x %g words x %d B = %.1f MB -> %.0f GB/s "
                   "(%.3f of 8000 GB/s)" % (a.cells, a.words, a.esz, algo / 1e6, algo / span / 1e3,
                                             algo / span / 1e3 / 8000.0))
        pn = sum(n for k, n, avg, tot in tab if "k_march2" in k)
        if pn:
            pavg = sum(tot for k, n, avg, tot in tab if "k_march2" in k) / pn
            palgo = a.cells * a.pass_words * a.esz
            out.append("- k_march2 (two steps per pass): %d launches, average %.2f us; algorithmic bytes per "
                       "launch: %d cells x %g words x %d B = %.1f MB -> %.0f GB/s (%.3f of 8000 GB/s)"
                       % (pn, pavg, a.cells, a.pass_words, a.esz, palgo / 1e6, palgo / pavg / 1e3,
                          palgo / pavg / 1e3 / 8000.0))
    else:
        out.append("- algorithmic bytes per step: %d cell-updates x %g words x %d B = %.1f MB"
                   % (a.cells, a.words, a.esz, algo / 1e6))
    if a.fetch and a.write:
        cf, nf = counter_sum(a.copy_fetch, "FETCH_SIZE", "k_copy")
        cw, nw = counter_sum(a.copy_write, "WRITE_SIZE", "k_copy")
        kib = a.copy_mib * 1024.0
        ff, fw = kib / (cf / nf), kib / (cw / nw)
        f, _ = counter_sum(a.fetch, "FETCH_SIZE", "k_march")
        w, _ = counter_sum(a.write, "WRITE_SIZE", "k_march")
        zf, _ = counter_sum(a.fetch, "FETCH_SIZE", "k_zsum")
        zw, _ = counter_sum(a.write, "WRITE_SIZE", "k_zsum")
        _, pmc_steps = counter_sum(a.fetch, "FETCH_SIZE", "k_surface")
        f, w = (f + zf) / pmc_steps, (w + zw) / pmc_steps
        traffic = (f * ff + w * fw) * 1024.0
        out.append("- counter calibration (copy of %d MiB per dispatch, `stream_probe copy`): FETCH_SIZE "
                   "reads %.0f KiB (factor %.4f), WRITE_SIZE %.0f KiB (factor %.4f)"
                   % (a.copy_mib, cf / nf, ff, cw / nw, fw))
        out.append("- HBM traffic per step of k_march (`%s`, %d steps, tiling %s): FETCH %.0f KiB, WRITE %.0f KiB "
                   "raw; calibrated %.1f MB = %.3f x algorithmic" % (a.pmc_cmd, pmc_steps, a.tiling, f, w,
                                                                   traffic / 1e6, traffic / algo))
        pf_, npf = counter_sum(a.fetch, "FETCH_SIZE", "k_march2")
        pw_, npw = counter_sum(a.write, "WRITE_SIZE", "k_march2")
        if npf and npw:
            palgo = a.cells * a.pass_words * a.esz
            traffic = (pf_ / npf * ff + pw_ / npw * fw) * 1024.0
            out.append("- HBM traffic per launch of k_march2 (%d launches): FETCH %.0f KiB, WRITE %.0f KiB raw; "
                       "calibrated %.1f MB = %.3f x its algorithmic %.1f MB" % (npf, pf_ / npf, pw_ / npw,
                                                                           traffic / 1e6, traffic / palgo,
                                                                           palgo / 1e6))
        if a.traffic and a.key:
            try:
                with open(a.traffic) as fh:
                    tj = json.load(fh)
            except (OSError, ValueError):
                tj = {}
            tj[a.key] = {"bytes": int(round(traffic)), "tiling": a.tiling}
            with open(a.traffic, "w") as fh:
                json.dump(tj, fh, indent=1, sort_keys=True)
                fh.write("\n")
    for d, cmd in zip(a.counters or [], a.counters_cmd or []):
        out.append("")
        out.append("Counters: `%s`" % cmd)
        out.append("")
        acc = counter_table(d)
        names = sorted({c for _, c in acc})
        kernels = sorted({k for k, _ in acc}, key=lambda k: -max(acc.get((k, c), 0.0) for c in names))
        out.append("| kernel | " + " | ".join(names) + " |")
        out.append("|---|" + "---|" * len(names))
        for k in kernels:
            out.append("| `%s` | " % _short(k, 60) + " | ".join("%.4g" % acc.get((k, c), 0.0) for c in names) + " |")
    print("\n".join(out))


def _marker_ranges(d):
    """(label, start, end) of the rocTX ranges whose label starts with "timed "."""
    out = []
    for r in _csv(d, "marker_api_trace.csv"):
        lab = next((v for v in r.values() if isinstance(v, str) and v.startswith("timed ")), None)
        if lab:
            out.append((lab, int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    return sorted(out, key=lambda x: x[1])


def bench(a):
    line = None
    with open(a.line) as fh:
        for ln in fh:
            if ln.startswith("{"):
                line = json.loads(ln)
    comp = dict(c.split("=", 1) for c in (a.compare or []))
    rows = _csv(a.trace, "kernel_trace.csv")
    out = ["# " + a.title, "", "Command: `%s`" % a.cmd, "",
           "Each section holds the dispatches that start inside one rocTX range around a timed region of "
           "the bench (its timed steps only: the autotuner and the warm-up steps precede the range).", ""]
    for lab, t0, t1 in _marker_ranges(a.trace):
        acc = {}
        for r in rows:
            s0, s1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            if t0 <= s0 <= t1:
                v = acc.setdefault(r["Kernel_Name"], [0, 0])
                v[0] += 1
                v[1] += s1 - s0
        if not acc:
            continue
        steps = sum(n for k, (n, t) in acc.items() if "k_surface" in k)
        out += ["## " + lab[len("timed "):], "", "| kernel | calls | avg us | total us |", "|---|---|---|---|"]
        for k, (n, t) in sorted(acc.items(), key=lambda x: -x[1][1]):
            out.append("| `%s` | %d | %.2f | %.1f |" % (_short(k), n, t / n / 1e3, t / 1e3))
        march = sum(t for k, (n, t) in acc.items() if "k_march" in k or "k_zsum" in k) / max(steps, 1) / 1e3
        out += ["", "- steps in the range: %d; k_march per step (lower + top launch, and any k_zsum pass; "
                "with two steps per pass, half the pass and both top launches): %.2f us" % (steps, march)]
        pn = sum(n for k, (n, t) in acc.items() if "k_march2" in k)
        pavg = sum(t for k, (n, t) in acc.items() if "k_march2" in k) / max(pn, 1) / 1e3
        if pn:
            out.append("- k_march2 (two steps per pass): %d launches, average %.2f us" % (pn, pavg))
        path = comp.get(lab[len("timed "):])
        if path and line:
            v = line
            for key in path.split("."):
                v = v[key]
            ref, what = (pavg, "the k_march2 average") if pn else (march, "the rocprof figure")
            out.append("- the bench line's own figure (`%s`, HIP events recorded by the launches' dispatch "
                       "packets): %.2f us, %.3f x %s" % (path, 1e3 * v, 1e3 * v / ref, what))
            if pn:
                v2 = line
                for key in path.rsplit(".", 1)[0].split(".") + ["march_ms_per_step"]:
                    v2 = v2[key]
                out.append("- the bench line's march per step (`march_ms_per_step` beside it): %.2f us, %.3f x "
                           "the rocprof k_march per step" % (1e3 * v2, 1e3 * v2 / march))
        out.append("")
    print("\n".join(out))


def kernels(a):
    """Trace table plus, per kernel-name pattern, launches, average duration and calibrated HBM
    bytes per launch."""
    out = ["# " + a.title, "", "Kernel trace: `%s`" % a.trace_cmd, "", "| kernel | calls | avg us | total us |",
           "|---|---|---|---|"]
    tab = trace_table(a.trace)
    for k, n, avg, tot in tab:
        out.append("| `%s` | %d | %.2f | %.1f |" % (_short(k), n, avg, tot))
    out.append("")
    ff = fw = None
    if a.fetch and a.write:
        cf, nf = counter_sum(a.copy_fetch, "FETCH_SIZE", "k_copy")
        cw, nw = counter_sum(a.copy_write, "WRITE_SIZE", "k_copy")
        ff, fw = a.copy_mib * 1024.0 / (cf / nf), a.copy_mib * 1024.0 / (cw / nw)
        out.append("Counters: `%s` (calibration factors FETCH %.4f, WRITE %.4f from `stream_probe copy %d`)"
                   % (a.pmc_cmd, ff, fw, a.copy_mib))
        out.append("")
    for pat in a.match or []:
        n = sum(x[1] for x in tab if pat in x[0])
        tot = sum(x[3] for x in tab if pat in x[0])
        if not n:
            continue
        line = "- `%s`: %d launches, average %.2f us" % (pat, n, tot / n)
        if ff is not None:
            f, nfd = counter_sum(a.fetch, "FETCH_SIZE", pat)
            w, nwd = counter_sum(a.write, "WRITE_SIZE", pat)
            if nfd and nwd:
                b = (f / nfd * ff + w / nwd * fw) * 1024.0
                line += "; HBM %.3f MB per launch -> %.0f GB/s (%.4f of 8000 GB/s)" % (b / 1e6, b / (tot / n) / 1e3,
                                                                                     b / (tot / n) / 1e3 / 8000.0)
        out.append(line)
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    m = sub.add_parser("march")
    m.add_argument("--title", required=True)
    m.add_argument("--trace")
    m.add_argument("--trace-cmd", default="")
    m.add_argument("--words", type=float, default=WORDS,
                   help="algorithmic words per cell-update (13; 12 + 1/B for a B-transmitter batch)")
    m.add_argument("--cells", type=int, required=True, help="cell-updates per step: L*M*(N-1) x batch")
    m.add_argument("--pass-words", type=float, default=11.0,
                   help="algorithmic words per cell of a two-step pass launch (11; 10 + 1/B for a batch)")
    m.add_argument("--esz", type=int, required=True)
    m.add_argument("--fetch")
    m.add_argument("--write")
    m.add_argument("--pmc-cmd", default="")
    m.add_argument("--copy-fetch")
    m.add_argument("--copy-write")
    m.add_argument("--copy-mib", type=int, default=256)
    m.add_argument("--tiling", default="")
    m.add_argument("--key")
    m.add_argument("--traffic")
    m.add_argument("--counters", action="append")
    m.add_argument("--counters-cmd", action="append")
    b = sub.add_parser("bench")
    b.add_argument("--title", default="Kernels of the bench's timed steps (MI355X, gfx950)")
    b.add_argument("--trace", required=True)
    b.add_argument("--cmd", required=True)
    b.add_argument("--line", required=True)
    b.add_argument("--compare", action="append")
    k = sub.add_parser("kernels")
    k.add_argument("--title", required=True)
    k.add_argument("--trace", required=True)
    k.add_argument("--trace-cmd", required=True)
    k.add_argument("--fetch")
    k.add_argument("--write")
    k.add_argument("--pmc-cmd", default="")
    k.add_argument("--copy-fetch")
    k.add_argument("--copy-write")
    k.add_argument("--copy-mib", type=int, default=256)
    k.add_argument("--match", action="append")
    a = ap.parse_args()
    if a.mode == "march":
        march(a)
    elif a.mode == "kernels":
        kernels(a)
    else:
        bench(a)


if __name__ == "__main__":
    main()
