"""Autotuner for the SpMV work split (degree bins and tile shape).

The knobs are those of egx_graph_set_bins:

* thread_row_max -- a tile whose mean row length is at most this reduces its
  rows one thread per row, otherwise one wavefront per row;
* wave_row_max   -- rows with more in-arcs are hub rows (split into
  segments; they also lead the relabelled iteration layout);
* hub_segment    -- arcs per hub segment (one workgroup each);
* tile_items     -- rows + arcs per merge-path tile (LDS-staged);
* tile_sort      -- gather a tile's arcs in source order (1) or row order (0);
* hot_sources    -- sources past this relabelled id are gathered with
  non-temporal loads, keeping the hot prefix resident (0: off);
* blocked        -- propagation-blocked iterations instead of the pull
  SpMV (FP32 iterates): tried once, after the pull knobs, on the best pull
  setting's relabelling.

``tune`` runs a coordinate descent over candidate values, timing the SpMV and
reduce launches of a few fixed iterations with hipEvents on the graph's own
stream, and keeps the fastest setting.  Results are cached in a JSON file
keyed by a coarse graph signature (weightedness, log2 n, mean degree and
log2 of the largest in-degree), so graphs of one family are tuned once.  Bin
settings never change what is computed, only how the work is split: results
stay within the parity bounds for every setting.
"""
from __future__ import annotations

import json
import math
import os
import time

from . import native

DEFAULT = dict(thread_row_max=12, wave_row_max=2048, hub_segment=8192, tile_items=5888, tile_sort=0, hot_sources=0,
               blocked=0)

SPACE = {
    "thread_row_max": [4, 8, 12, 16, 24, 48],
    "wave_row_max": [512, 1024, 2048, 4096],
    "hub_segment": [2048, 4096, 8192, 16384, 32768],
    "tile_items": [1536, 2944, 4416, 5888, 7424],
    "tile_sort": [0, 1],
    "hot_sources": [0, 1 << 18, 1 << 19, 1 << 20, 1 << 21],
}

# tile_items + wave_row_max must fit the LDS staging tile of the SpMV (7936 terms)
_LDS_TERMS = 7936


def _valid(cfg):
    return cfg["tile_items"] + cfg["wave_row_max"] <= _LDS_TERMS


def signature(info, precision=native.PRECISION_FP32):
    n = max(int(info["n"]), 1)
    m = int(info["m"])
    deg = m / n
    return "w%d-p%d-n%d-d%d-h%d" % (int(bool(info["weighted"])), int(precision), round(math.log2(n)),
                                   round(2 * math.log2(max(deg, 1.0))), round(math.log2(max(info["max_in_degree"], 1))))


def set_bins(g, cfg):
    """egx_graph_set_bins with ``cfg``; hot_sources 0 means off here."""
    args = dict(cfg)
    if "hot_sources" in args and args["hot_sources"] <= 0:
        args["hot_sources"] = -1
    g.set_bins(**args)


def measure(g, cfg, iters=5, precision=native.PRECISION_FP32, repeats=2):
    """Mean seconds per iteration (SpMV + reduce launches) under ``cfg``."""
    set_bins(g, cfg)
    g.pagerank_stats(0.85, 2, 0.0, precision=precision, time_each_launch=False)
    best = float("inf")
    for _ in range(repeats):
        _, st = g.pagerank_stats(0.85, iters, 0.0, precision=precision, time_each_launch=True)
        best = min(best, (st["spmv_seconds"] + st["reduce_seconds"]) / max(st["spmv_launches"], 1))
    return best


def tune(g, precision=native.PRECISION_FP32, iters=5, passes=2, space=None, start=None, log=None):
    """Coordinate descent from ``start`` (default: the library defaults).
    Returns (best cfg, seconds per iteration, trials).  Leaves ``g`` binned
    with the best setting."""
    space = space or SPACE
    cur = dict(DEFAULT, **(start or {}))
    cur["blocked"] = 0
    trials = []
    seen = {}

    def cost(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = measure(g, cfg, iters, precision)
            trials.append((dict(cfg), seen[key]))
            if log:
                log(f"  {cfg} -> {seen[key] * 1e6:.1f} us/iter")
        return seen[key]

    best = cost(cur)
    for _ in range(passes):
        improved = False
        for name, values in space.items():
            for v in values:
                if v == cur[name]:
                    continue
                cand = dict(cur, **{name: v})
                if not _valid(cand):
                    continue
                c = cost(cand)
                if c < best * 0.99:  # ignore timing noise
                    best, cur, improved = c, cand, True
        if not improved:
            break
    info = g.info()
    if precision == native.PRECISION_FP32 and info["m"] > 0:
        cand = dict(cur, blocked=1)
        c = cost(cand)
        if c < best * 0.99:
            best, cur = c, cand
    set_bins(g, cur)
    return cur, best, trials


class TuneCache:
    """JSON cache {signature: {"bins": cfg, "us_per_iter": t, "tuned_on": {...}}}."""

    def __init__(self, path=None):
        self.path = path or os.environ.get("EGGPU_AMD_TUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "eggpu_amd", "bins.json")
        self.data = {}
        if os.path.exists(self.path):
            try:
                with open(self.path) as fh:
                    self.data = json.load(fh)
            except (OSError, ValueError):
                self.data = {}

    def get(self, sig):
        e = self.data.get(sig)
        return dict(e["bins"]) if e else None

    def put(self, sig, cfg, seconds, info):
        self.data[sig] = {"bins": dict(cfg), "us_per_iter": round(seconds * 1e6, 2),
                          "tuned_on": {"n": int(info["n"]), "m": int(info["m"])}, "when": int(time.time())}
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(self.data, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def apply(g, cache=None, precision=native.PRECISION_FP32, tune_if_missing=True, **kw):
    """Bin ``g`` with the cached setting of its signature, tuning (and
    caching) it first when missing.  Returns the setting used."""
    cache = cache if cache is not None else TuneCache()
    info = g.info()
    sig = signature(info, precision)
    cfg = cache.get(sig)
    if cfg is None:
        if not tune_if_missing:
            return None
        cfg, sec, _ = tune(g, precision=precision, **kw)
        cache.put(sig, cfg, sec, info)
    else:
        set_bins(g, cfg)
    return cfg


def apply_to_set(graphs, cache=None, precision=native.PRECISION_FP32, log=None, **kw):
    """Tune one representative (the largest) per signature and bin every
    graph of the set with its signature's setting.  Returns {signature: cfg}."""
    cache = cache if cache is not None else TuneCache()
    groups = {}
    for g in graphs:
        info = g.info()
        groups.setdefault(signature(info, precision), []).append((int(info["m"]), g))
    chosen = {}
    for sig, members in groups.items():
        members.sort(key=lambda t: -t[0])
        cfg = cache.get(sig)
        if cfg is None:
            rep = members[0][1]
            cfg, sec, _ = tune(rep, precision=precision, **kw)
            cache.put(sig, cfg, sec, rep.info())
            if log:
                log(f"tuned {sig}: {cfg} ({sec * 1e6:.1f} us/iter on m={members[0][0]})")
        for _, g in members:
            set_bins(g, cfg)
        chosen[sig] = cfg
    return chosen
