"""Measurement commands behind the committed profiles.

    python -m eggpu_amd.probe copy   [--bytes B] [--reps R]
    python -m eggpu_amd.probe gather [--n N] [--reps R]
    python -m eggpu_amd.probe calib  [--scale S] [--iters K] [--bytes B] [--bins JSON]
        one plain copy of B bytes, then K PageRank iterations on R-MAT
        scale S: the workload rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
        use, so the copy's known bytes calibrate the SpMV's counters
    python -m eggpu_amd.probe stats   TRACE [--out FILE] [--markdown] [--min-us U]
        per-kernel duration summary of a rocprofv3 --kernel-trace database
        or *_kernel_trace.csv
    python -m eggpu_amd.probe traffic --fetch DB --write DB --bytes B [--kernel K] [--out FILE]
        HBM bytes per launch of kernel K from the two --pmc databases of
        `calib`, scaled by what the counters read for the B-byte copy

Each command prints JSON.
"""
from __future__ import annotations

import argparse
import ctypes
import glob
import json
import os
import sqlite3
import sys

from . import native


def _copy(nbytes, reps):
    sec = ctypes.c_double(0.0)
    native._check(native.lib().egx_probe_copy(int(nbytes), int(reps), ctypes.byref(sec)))
    return sec.value


def _gather(ws, n, reps):
    sec = ctypes.c_double(0.0)
    native._check(native.lib().egx_probe_gather(int(ws), int(n), int(reps), ctypes.byref(sec)))
    return sec.value


def cmd_copy(a):
    s = _copy(a.bytes, a.reps)
    print(json.dumps({"probe": "copy", "bytes": a.bytes, "seconds": s,
                      "read_plus_write_TBps": 2 * a.bytes / s / 1e12}))


def cmd_gather(a):
    rows = []
    for ws in (1 << 12, 1 << 14, 1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24, 1 << 26):
        s = _gather(ws, a.n, a.reps)
        rows.append({"working_set_bytes": ws * 4, "gathers": a.n, "seconds": s,
                     "Ggathers_per_s": a.n / s / 1e9})
    print(json.dumps({"probe": "gather", "rows": rows}))


def cmd_calib(a):
    _copy(a.bytes, 1)
    g = native.DeviceGraph.rmat(a.scale, 16, seed=1, weighted=a.weighted, wseed=2)
    if a.bins:
        g.set_bins(**json.loads(a.bins))
    inf = g.info()
    _, st = g.pagerank_stats(0.85, a.iters, 0.0, time_each_launch=False)
    print(json.dumps({"probe": "calib", "copy_bytes": a.bytes, "n": inf["n"], "m": inf["m"],
                      "iterations": st["iterations"]}))


def _dbs(path):
    if os.path.isfile(path):
        return [path]
    return sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))


def _kernel_rows(db):
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(kernels)")]
    rows = [dict(zip(cols, r)) for r in c.execute("select * from kernels")]
    c.close()
    return rows


def _trace_durations(path):
    """{kernel name: [duration ns per dispatch]} from a rocprofv3 kernel
    trace: a --kernel-trace database or a *_kernel_trace.csv."""
    per = {}
    if path.endswith(".csv"):
        import csv

        with open(path) as fh:
            for r in csv.DictReader(fh):
                per.setdefault(r["Kernel_Name"], []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        return per
    for db in _dbs(path):
        for r in _kernel_rows(db):
            per.setdefault(r["name"], []).append(r["duration"])
    return per


def _short(name):
    name = name.replace("egx::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0] if "(" in name else name


def cmd_stats(a):
    """Per-kernel durations.  Launches that find the device stop flag set
    return in a few microseconds; --min-us separates them from the launches
    that iterate, so the mean of the latter compares with the hipEvent
    timings that bench.py reports."""
    out = []
    for name, durs in _trace_durations(a.db).items():
        work = [d for d in durs if d >= a.min_us * 1e3]
        out.append({"kernel": _short(name), "calls": len(durs), "total_us": sum(durs) / 1e3,
                    "mean_us": sum(durs) / len(durs) / 1e3, "min_us": min(durs) / 1e3, "max_us": max(durs) / 1e3,
                    "calls_ge_min": len(work), "mean_ge_min_us": (sum(work) / len(work) / 1e3) if work else None})
    out.sort(key=lambda r: -r["total_us"])
    if a.markdown:
        lines = [f"# {a.title or 'kernel durations'}", ""]
        if a.command:
            lines += ["Command (from the repository root):", "", "```", a.command, "```", ""]
        lines += [f"Source: `{a.db}`; launches shorter than {a.min_us:g} us (stop flag already set) are "
                  "excluded from the last two columns.", "",
                  f"| kernel | calls | total us | mean us | min us | max us | calls >= {a.min_us:g} us | mean of those us |",
                  "|---|---:|---:|---:|---:|---:|---:|---:|"]
        for r in out[: a.top]:
            m = "-" if r["mean_ge_min_us"] is None else f"{r['mean_ge_min_us']:.1f}"
            lines.append(f"| `{r['kernel']}` | {r['calls']} | {r['total_us']:.1f} | {r['mean_us']:.1f} | "
                         f"{r['min_us']:.1f} | {r['max_us']:.1f} | {r['calls_ge_min']} | {m} |")
        text = "\n".join(lines)
    else:
        text = json.dumps({"source": a.db, "kernels": out}, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text + "\n")
    print(text)


def _pmc(db, counter):
    """{kernel name: [counter value per dispatch]} from a --pmc database."""
    c = sqlite3.connect(db)
    tabs = [r[0] for r in c.execute("select name from sqlite_master where type in ('table','view')")]
    res = {}
    if "counters_collection" in tabs:
        cols = [r[1] for r in c.execute("pragma table_info(counters_collection)")]
        name_col = "kernel_name" if "kernel_name" in cols else "name"
        val_col = "value" if "value" in cols else "counter_value"
        q = f"select {name_col}, dispatch_id, counter_name, {val_col} from counters_collection"
        acc = {}
        for name, disp, cname, val in c.execute(q):
            if cname != counter:
                continue
            acc[(name, disp)] = acc.get((name, disp), 0.0) + float(val)
        for (name, _), v in acc.items():
            res.setdefault(name, []).append(v)
    c.close()
    return res


def cmd_traffic(a):
    fetch = {}
    write = {}
    for db in _dbs(a.fetch):
        for k, v in _pmc(db, "FETCH_SIZE").items():
            fetch.setdefault(k, []).extend(v)
    for db in _dbs(a.write):
        for k, v in _pmc(db, "WRITE_SIZE").items():
            write.setdefault(k, []).extend(v)

    def pick(d, part):
        for k, v in d.items():
            if part in k and v:
                return sum(v) / len(v), len(v), k
        return None, 0, None

    cf, _, _ = pick(fetch, "k_probe_copy")
    cw, _, _ = pick(write, "k_probe_copy")
    if None in (cf, cw):
        raise SystemExit(f"missing copy counters: fetch {cf} write {cw}")
    f_scale = a.bytes / cf  # bytes per counter unit, from the copy's known read
    w_scale = a.bytes / cw
    # --kernel a,b: the launches of one iteration, summed (e.g. a scatter + sum pair)
    per = []
    for part in [k for k in a.kernel.split(",") if k]:
        kf, nf, kname = pick(fetch, part)
        kw, nw, _ = pick(write, part)
        if None in (kf, kw):
            raise SystemExit(f"missing counters for {part}: fetch {kf} write {kw}")
        per.append({"kernel": kname, "dispatches": {"fetch": nf, "write": nw}, "FETCH_SIZE_mean": kf,
                    "WRITE_SIZE_mean": kw, "hbm_read_bytes": kf * f_scale, "hbm_write_bytes": kw * w_scale})
    rd = sum(p["hbm_read_bytes"] for p in per)
    wr = sum(p["hbm_write_bytes"] for p in per)
    out = {
        "kernel": " + ".join(p["kernel"] for p in per),
        "dispatches": per[0]["dispatches"],
        "copy_bytes": a.bytes,
        "copy_FETCH_SIZE": cf,
        "copy_WRITE_SIZE": cw,
        "fetch_scale_bytes_per_unit": f_scale,
        "write_scale_bytes_per_unit": w_scale,
        "kernel_FETCH_SIZE_mean": sum(p["FETCH_SIZE_mean"] for p in per),
        "kernel_WRITE_SIZE_mean": sum(p["WRITE_SIZE_mean"] for p in per),
        "hbm_read_bytes_per_launch": rd,
        "hbm_write_bytes_per_launch": wr,
        "hbm_bytes_per_launch": rd + wr,
        "per_kernel": per if len(per) > 1 else None,
        "fetch_db": a.fetch,
        "write_db": a.write,
    }
    text = json.dumps(out, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text + "\n")
    print(text)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m eggpu_amd.probe")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("copy")
    p.add_argument("--bytes", type=int, default=1 << 30)
    p.add_argument("--reps", type=int, default=10)
    p = sub.add_parser("gather")
    p.add_argument("--n", type=int, default=1 << 27)
    p.add_argument("--reps", type=int, default=3)
    p = sub.add_parser("calib")
    p.add_argument("--scale", type=int, default=24)
    p.add_argument("--iters", type=int, default=3)
    p.add_argument("--bytes", type=int, default=1 << 30)
    p.add_argument("--bins", help='JSON of egx_graph_set_bins fields, e.g. {"tile_sort": 1}')
    p.add_argument("--weighted", action="store_true", help="U(0.1,1) arc weights, as bench.py --weighted")
    p = sub.add_parser("stats")
    p.add_argument("db")
    p.add_argument("--out")
    p.add_argument("--min-us", type=float, default=100.0)
    p.add_argument("--markdown", action="store_true")
    p.add_argument("--title")
    p.add_argument("--command")
    p.add_argument("--top", type=int, default=20)
    p = sub.add_parser("traffic")
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--bytes", type=int, default=1 << 30)
    p.add_argument("--kernel", default="k_pr_spmv", help="kernel name part, or several joined by commas")
    p.add_argument("--out")
    a = ap.parse_args(argv)
    {"copy": cmd_copy, "gather": cmd_gather, "calib": cmd_calib, "stats": cmd_stats,
     "traffic": cmd_traffic}[a.cmd](a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
