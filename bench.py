"""PageRank throughput on one MI355X.

Workload (the plain one-graph GPU method): directed R-MAT scale 24 (16.7M
nodes, 16 x 2^24 draws, a=0.57 b=0.19 c=0.19, seed 1, self-loops and
duplicates removed), unweighted, alpha 0.85, generated and built in HBM.
A step is one power iteration over the whole graph (tol = 0 so every step
runs; the tol=1e-6 solve is reported beside it).  value = GTEPS =
arcs x steps / seconds of the timed region, which is bracketed by device
synchronisation on both sides.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--scale 24]

Prints one JSON line.  ``roofline`` uses the algorithmic bytes of one
iteration, 4m + 16n (int32 column index per arc; row pointer, out-degree,
x read and x' write per node), over the mean duration of the SpMV launch
measured with hipEvents on the graph's stream (the scatter + sum launch pair
when the tuner picks the propagation-blocked iteration).  ``cpu_baseline`` times the
OpenMP push of pagerank.cpp:21-122 (oracle/egx_oracle.c) on this host, on
the same graph, for the same K iterations; its ranks and the sequential
FP64 restatement of pagerank.py:75-139 at tol=1e-6 check the GPU's ranks,
and the max-abs and L1 errors go to ``config.parity``.

    python bench.py --calib-copy BYTES ...
        one device copy of BYTES before the workload: run under
        rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE), the copy calibrates
        the counters
    python bench.py --pmc-traffic FETCH_DIR WRITE_DIR [--calib-copy BYTES]
                    [--kernels k_pb_scatter,k_pb_sum]
        no GPU work: HBM bytes per launch of the kernels from those runs
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12

from eggpu_amd.profiling import calib_copy, pmc_traffic  # noqa: E402


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(blocked, weighted=False):
    """HBM bytes per SpMV launch (per scatter + sum pair when blocked) from
    the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration); none is kept for the weighted pull SpMV."""
    if weighted and not blocked:
        return None, None
    name = ("pmc_traffic_rmat24_weighted_blocked.json" if weighted else "pmc_traffic_rmat24_blocked.json") \
        if blocked else "pmc_traffic_rmat24.json"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch"), os.path.relpath(path, ROOT)


def cpu_baseline(g, n, m, iters, threads, gpu_k, gpu_tol, weighted=False):
    """The OpenMP push of pagerank.cpp:21-122 timed over `iters` fixed
    iterations on this host; its result is also the check of the GPU's
    K-iteration ranks (gpu_k), and the sequential FP64 restatement of
    pagerank.py:75-139 at tol=1e-6 checks the GPU's tol=1e-6 solve
    (gpu_tol = (ranks, iterations)).  Errors go to `parity`."""
    from oracle import oracle as O

    V, E, W = g.download_csr(with_weights=weighted)
    Wc = W if weighted else None
    t0 = time.perf_counter()
    ref_k, it = O.pagerank_omp(V, E, Wc, 0.85, iters, 0.0, threads=threads)
    dt = time.perf_counter() - t0
    d = np.abs(gpu_k - ref_k)
    parity = {"fixed_K": {"K": iters, "oracle": "orc_pagerank_omp (FP64)", "max_abs": float(d.max()),
                          "l1": float(d.sum())}}
    t1 = time.perf_counter()
    ref_t, it_t, _ = O.pagerank_csr(V, E, Wc, 0.85, 100, 1e-6)
    d = np.abs(gpu_tol[0] - ref_t)
    parity["tol1e-6"] = {"oracle": "orc_pagerank_push (FP64, pagerank.py:75-139 order)", "iterations": gpu_tol[1],
                         "oracle_iterations": it_t, "max_abs": float(d.max()), "l1": float(d.sum()),
                         "oracle_seconds": round(time.perf_counter() - t1, 2)}
    return {
        "value": round(m * it / dt / 1e9, 4),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"same R-MAT scale graph (n={n}, m={m}), {it} iterations at fixed K, OpenMP push with "
                  f"atomic scatter (pagerank.cpp:21-122 semantics), {dt:.2f} s",
        "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model()},
        "parity": parity,
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_c1(V, E, threads):
    """R-MAT 14 on the host: the sequential FP64 restatement of
    pagerank.py:75-139 (one core) at tol=1e-6 and at K=20, and the OpenMP
    push of pagerank.cpp:21-122 at K=20."""
    from oracle import oracle as O

    n, m = len(V) - 1, len(E)
    src, dst = O.csr_rows(V, E)
    t0 = time.perf_counter()
    _, it, _ = O.pagerank_rows(n, src, dst, None, 0.85, 100, 1e-6)
    seq_tol = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.pagerank_rows(n, src, dst, None, 0.85, 20, 0.0)
    seq_k = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.pagerank_omp(V, E, None, 0.85, 20, 0.0, threads=threads)
    omp_k = time.perf_counter() - t0
    return {"sequential_tol1e-6_ms": round(seq_tol * 1e3, 3), "sequential_tol1e-6_iterations": it,
            "sequential_K20_gteps": round(m * 20 / seq_k / 1e9, 4), "openmp_K20_gteps": round(m * 20 / omp_k / 1e9, 4),
            "openmp_threads": threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default bins")
    ap.add_argument("--weighted", action="store_true",
                    help="U(0.1,1) arc weights (a secondary measurement; the headline graph is unweighted)")
    ap.add_argument("--bins", help='fixed bins instead of the tuner, JSON of egx_graph_set_bins fields')
    ap.add_argument("--sections", default="batched_ppr,multi_graph,serving,katz,eigenvector,ingest",
                    help="secondary workloads reported beside the headline line ('' for none)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="BYTES",
                    help="run one device copy of BYTES first (counter calibration for --pmc-traffic)")
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no GPU work: summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this "
                         "script with --calib-copy into HBM bytes per launch of --kernels")
    ap.add_argument("--kernels", default="k_pb_scatter,k_pb_sum")
    args = ap.parse_args()
    if args.pmc_traffic:
        out = pmc_traffic(args.pmc_traffic[0], args.pmc_traffic[1], args.calib_copy or (1 << 30),
                          [k for k in args.kernels.split(",") if k])
        print(json.dumps(out, indent=1), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import eggpu_amd
    from eggpu_amd import native

    native.lib()
    if native.device_count() < 1:
        raise SystemExit("no HIP device visible")

    if args.calib_copy > 0:
        calib_copy(args.calib_copy)
    log(f"building R-MAT scale {args.scale} on the device")
    t0 = time.perf_counter()
    g = native.DeviceGraph.rmat(args.scale, 16, a=0.57, b=0.19, c=0.19, seed=1, keep_csr=not args.no_cpu,
                                weighted=args.weighted, wseed=2)
    build_wall = time.perf_counter() - t0
    info = g.info()
    n, m = int(info["n"]), int(info["m"])
    log(f"n={n} m={m} dangling={info['n_dangling']} build {build_wall:.3f}s")

    # bins chosen by the autotuner (coordinate descent from the defaults)
    tuned = None
    if args.bins:
        g.set_bins(**json.loads(args.bins))
        info = g.info()
        tuned = {"bins": json.loads(args.bins), "fixed": True}
    elif not args.no_tune:
        from eggpu_amd import autotune

        t1 = time.perf_counter()
        cfg, sec, trials = autotune.tune(g, passes=1, iters=4)
        tuned = {"bins": cfg, "us_per_iter": round(sec * 1e6, 1), "trials": len(trials),
                 "seconds": round(time.perf_counter() - t1, 2)}
        log(f"autotuned bins {cfg} ({sec * 1e3:.3f} ms/iter, {len(trials)} trials)")
        info = g.info()

    # warmup
    if args.warmup > 0:
        g.pagerank_stats(0.85, args.warmup, 0.0, time_each_launch=False)

    # timed region: K iterations, device synchronised before and after
    K = args.steps
    t0 = time.perf_counter()
    _, st = g.pagerank_stats(0.85, K, 0.0, time_each_launch=False)  # returns after a stream sync
    wall = time.perf_counter() - t0
    assert st["iterations"] == K
    value = m * K / wall / 1e9

    # per-launch timing of the SpMV kernel for the roofline (separate run)
    _, st_t = g.pagerank_stats(0.85, K, 0.0, time_each_launch=True)
    spmv_mean = st_t["spmv_seconds"] / st_t["spmv_launches"]
    reduce_mean = st_t["reduce_seconds"] / st_t["spmv_launches"]
    # weighted graphs also read a 4-byte arc value per arc
    bytes_iter = (8 if args.weighted else 4) * m + 16 * n
    achieved = bytes_iter / spmv_mean
    blocked = bool(info["blocked"])
    traffic, traffic_src = read_traffic(blocked, args.weighted) if args.scale == 24 else (None, None)

    # tol = 1e-6 solve (reference semantics), the K-iteration ranks and the
    # FP64 variant, beside
    pr, it_tol, _ = g.pagerank(0.85, 100, 1e-6)
    pr_k, _, _ = g.pagerank(0.85, K, 0.0)
    _, st64 = g.pagerank_stats(0.85, K, 0.0, precision=native.PRECISION_FP64, time_each_launch=False)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {K} iterations, {threads} threads; oracle checks of the GPU ranks")
        cpu = cpu_baseline(g, n, m, K, threads, pr_k, (pr, it_tol), weighted=args.weighted)
        log(f"parity: {json.dumps(cpu['parity'])}")

    # the CPU-runnable configuration (R-MAT 14) on the GPU, with its CPU legs
    from eggpu_amd import benchmarks

    c1, V1, E1 = benchmarks.c1_small()
    if cpu is not None:
        cpu["c1"] = cpu_c1(V1, E1, min(16, os.cpu_count() or 1))

    secondary = {}
    wanted = [x for x in args.sections.split(",") if x]
    if wanted:
        from eggpu_amd import benchmarks

        g.close()  # the secondary workloads build their own graphs
        for name in wanted:
            log(f"secondary workload: {name}")
            t1 = time.perf_counter()
            try:
                secondary[name] = getattr(benchmarks, name)()
            except Exception as exc:  # reported in the line; the headline stands on its own
                secondary[name] = {"error": f"{type(exc).__name__}: {exc}"}
            secondary[name]["section_wall_s"] = round(time.perf_counter() - t1, 1)
            log(f"{name}: {json.dumps(secondary[name])}")

    line = {
        "metric": "PageRank GTEPS & % HBM roofline; batched solves/s; p50/p99 latency @QPS",
        "value": round(value, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT generated on the device)",
        "config": {
            "workload": f"pagerank R-MAT scale {args.scale}{' weighted U(0.1,1)' if args.weighted else ''} plain "
                        f"method, fixed K={K} (tol=0)",
            "n": n,
            "m": m,
            "alpha": 0.85,
            "iterate_dtype": "f32",
            "reductions": "f64 (dangling mass, L1 residual, teleport, row sums)",
            "bins": {k: info[k] for k in ("thread_row_max", "wave_row_max", "hub_segment", "tile_items", "tile_sort",
                                                  "hot_sources", "blocked")},
            "autotune": tuned,
            "hub_rows": info["rows_hub"],
            "tiles": info["tiles"],
            "build_seconds": round(info["build_seconds"], 4),
            "kernel_ms_per_iter": round(st["kernel_seconds"] / K * 1e3, 4),
            "spmv_ms_mean": round(spmv_mean * 1e3, 4),
            "reduce_ms_mean": round(reduce_mean * 1e3, 4),
            "tol1e-6_iterations": it_tol,
            "fp64_ms_per_iter": round(st64["kernel_seconds"] / K * 1e3, 4),
            "fp64_gteps": round(m * K / st64["kernel_seconds"] / 1e9, 3),
            "parity": cpu["parity"] if cpu else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved / 1e9, 2),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": traffic,
            "kernel": "k_pb_scatter + k_pb_sum" if blocked else "k_pr_spmv",
            "algorithmic_bytes_per_launch": bytes_iter,
            "traffic_source": traffic_src,
        },
        "cpu_baseline": cpu,
        "secondary": dict(secondary, c1_small=c1),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
