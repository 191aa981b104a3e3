"""ctypes binding of ``libegx.so`` (the C ABI in ``include/egx.h``).

This module is the only door to the GPU.  It raises
:class:`NativeLibraryMissing` when the library has not been built; nothing in
the package computes PageRank on the CPU instead.

It also exposes the dict-returning shim the backend calls, mirroring
``cpp_easygraph.cpp_gpu_pagerank_dense`` (``cpp_easygraph.cpp:120``,
``functions/pagerank/pagerank.cpp:176-223``): ``{"values_dense": float64[n],
"kernel_seconds": float}``.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libegx.so")

EGX_SUCC = 0
EGX_FAILED_TO_ALLOCATE_HOST_MEM = 1
EGX_FAILED_TO_ALLOCATE_DEVICE_MEM = 2
EGX_DEVICE_ERR = 3
EGX_UNKNOW_ERROR = 4
EGX_UNSUPPORTED_GRAPH = 5
EGX_NEGATIVE_CYCLE = 6

PRECISION_FP32 = 0
PRECISION_FP64 = 1
SOLVE_TIME_LAUNCHES = 1
SOLVE_HIP_GRAPH = 2
SOLVE_XCD_PLACEMENT = 4


class NativeLibraryMissing(RuntimeError):
    """libegx.so is absent or failed to load: the HIP path cannot run."""


class EgxError(RuntimeError):
    """A non-zero status from the C ABI (``RuntimeError`` like the
    reference's ``pybind11_fail``, ``pagerank.cpp:205-207``)."""

    def __init__(self, status, detail, last):
        msg = detail if not last else f"{detail} ({last})"
        super().__init__(msg)
        self.status = status


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("m", ctypes.c_int64),
        ("n_dangling", ctypes.c_int64),
        ("weighted", ctypes.c_int32),
        ("max_in_degree", ctypes.c_int32),
        ("max_out_degree", ctypes.c_int32),
        ("thread_row_max", ctypes.c_int32),
        ("wave_row_max", ctypes.c_int32),
        ("hub_segment", ctypes.c_int32),
        ("tile_items", ctypes.c_int32),
        ("reordered", ctypes.c_int32),
        ("hub_rows_first", ctypes.c_int32),
        ("tile_sort", ctypes.c_int32),
        ("hot_sources", ctypes.c_int32),
        ("blocked", ctypes.c_int32),
        ("rows_thread", ctypes.c_int64),
        ("rows_wave", ctypes.c_int64),
        ("rows_hub", ctypes.c_int64),
        ("hub_segments", ctypes.c_int64),
        ("tiles", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("build_seconds", ctypes.c_double),
        ("layout_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BinOpts(ctypes.Structure):
    _fields_ = [
        ("thread_row_max", ctypes.c_int32),
        ("wave_row_max", ctypes.c_int32),
        ("hub_segment", ctypes.c_int32),
        ("tile_items", ctypes.c_int32),
        ("reorder", ctypes.c_int32),
        ("tile_sort", ctypes.c_int32),
        ("hot_sources", ctypes.c_int32),
        ("blocked", ctypes.c_int32),
        ("blocked_shared", ctypes.c_int32),
    ]


class SolveOpts(ctypes.Structure):
    _fields_ = [
        ("alpha", ctypes.c_double),
        ("max_iter", ctypes.c_int32),
        ("tol", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("check_every", ctypes.c_int32),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("kernel_seconds", ctypes.c_double),
        ("spmv_seconds", ctypes.c_double),
        ("reduce_seconds", ctypes.c_double),
        ("iterations", ctypes.c_int32),
        ("spmv_launches", ctypes.c_int32),
        ("spmv_grid", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class IngestReport(ctypes.Structure):
    _fields_ = [("in_range", ctypes.c_int32), ("finite", ctypes.c_int32), ("simple_sorted", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("bytes", ctypes.c_int64), ("read_seconds", ctypes.c_double),
                ("copy_seconds", ctypes.c_double), ("check_seconds", ctypes.c_double),
                ("build_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_sessions", ctypes.c_int32),
                ("step_iters", ctypes.c_int32), ("coalesce_off", ctypes.c_int32), ("pad", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("width", ctypes.c_int32), ("queue_seconds", ctypes.c_double),
                ("service_seconds", ctypes.c_double), ("latency_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "failed", "steps", "sessions_opened",
                                              "sessions_closed", "peak_active_columns", "peak_session_bytes",
                                              "budget_bytes", "largest_session_bytes", "coalesced")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class EigOpts(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int32),
        ("delta", ctypes.c_int32),
        ("tol", ctypes.c_double),
        ("norm_floor", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("start_norm", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class CacheStats(ctypes.Structure):
    _fields_ = [
        ("hits", ctypes.c_uint64),
        ("misses", ctypes.c_uint64),
        ("builds", ctypes.c_uint64),
        ("value_refreshes", ctypes.c_uint64),
        ("resident_bytes", ctypes.c_uint64),
        ("evictions", ctypes.c_uint64),
        ("entries", ctypes.c_uint64),
    ]


_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every function declared in include/egx.h
SIGNATURES = {
    "egx_err_code_detail": (ctypes.c_char_p, [ctypes.c_int]),
    "egx_last_error": (ctypes.c_char_p, []),
    "egx_version": (ctypes.c_char_p, []),
    "egx_device_count": (ctypes.c_int, []),
    "egx_pagerank": (ctypes.c_int, [_i32p, _i32p, _f64p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                    ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, _f64p, _i32p, _f64p]),
    "egx_pagerank_cache_reset": (None, []),
    "egx_pagerank_cache_stats": (ctypes.c_int, [ctypes.POINTER(CacheStats)]),
    "egx_graph_from_csr": (ctypes.c_int, [_i32p, _i32p, _f64p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_int32, ctypes.POINTER(_vp)]),
    "egx_graph_rmat": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int32,
                                      ctypes.POINTER(_vp)]),
    "egx_graph_rmat_n": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int32,
                                        ctypes.POINTER(_vp)]),
    "egx_graph_from_csr_files": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32,
                                                ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.POINTER(IngestReport), ctypes.POINTER(_vp)]),
    "egx_graph_er": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int32,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                    ctypes.c_int32, ctypes.POINTER(_vp)]),
    "egx_graph_info_get": (ctypes.c_int, [_vp, ctypes.POINTER(GraphInfo)]),
    "egx_graph_set_bins": (ctypes.c_int, [_vp, ctypes.POINTER(BinOpts)]),
    "egx_graph_download_csc": (ctypes.c_int, [_vp, _i32p, _i32p, _f32p, _i32p]),
    "egx_graph_download_csr": (ctypes.c_int, [_vp, _i32p, _i32p, _f64p, _i64p]),
    "egx_graph_pagerank": (ctypes.c_int, [_vp, ctypes.POINTER(SolveOpts), _f64p, _i32p, _f64p]),
    "egx_graph_pagerank_ex": (ctypes.c_int, [_vp, ctypes.POINTER(SolveOpts), _f64p, ctypes.POINTER(SolveStats),
                                             ctypes.c_int32]),
    "egx_graph_pagerank_noalloc": (ctypes.c_int, [_vp, ctypes.POINTER(SolveOpts), ctypes.c_void_p, _i32p]),
    "egx_graph_ppr_single": (ctypes.c_int, [_vp, _i32p, ctypes.c_int32, ctypes.POINTER(SolveOpts), _f64p,
                                            ctypes.POINTER(SolveStats), ctypes.c_int32]),
    "egx_multi_pagerank": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.POINTER(SolveOpts),
                                          ctypes.POINTER(_f64p), _i32p, ctypes.POINTER(SolveStats), ctypes.c_int32]),
    "egx_graph_free": (None, [_vp]),
    "egx_server_open": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int32, ctypes.POINTER(ServerOpts),
                                       ctypes.POINTER(_vp)]),
    "egx_server_submit": (ctypes.c_int, [_vp, ctypes.c_int32, _i32p, ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_int32, ctypes.c_double, _f64p, _i64p]),
    "egx_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(RequestInfo)]),
    "egx_server_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    "egx_server_close": (None, [_vp]),
    "egx_batch_open": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "egx_batch_assign": (ctypes.c_int, [_vp, ctypes.c_int32, _i32p, ctypes.c_int32, ctypes.c_double,
                                        ctypes.c_int32, ctypes.c_double]),
    "egx_batch_release": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "egx_batch_step": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "egx_batch_poll": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _f64p]),
    "egx_batch_ready": (ctypes.c_int, [_vp]),
    "egx_batch_run": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _f64p]),
    "egx_batch_result": (ctypes.c_int, [_vp, ctypes.c_int32, _f64p]),
    "egx_batch_close": (None, [_vp]),
    "egx_graph_ppr": (ctypes.c_int, [_vp, ctypes.c_int32, _i32p, _i32p, _f64p, ctypes.c_double, ctypes.c_int32,
                                     ctypes.c_double, _f64p, _i32p, _f64p]),
    "egx_katz_csr": (ctypes.c_int, [_i32p, _i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _f64p,
                                    ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _f64p, _i32p,
                                    _i32p, _f64p]),
    "egx_eigenvector_csr": (ctypes.c_int, [_i32p, _i32p, _f64p, ctypes.c_int32, ctypes.c_int32, _f64p,
                                           ctypes.POINTER(EigOpts), _f64p, _i32p, _i32p, _f64p]),
}

_lib = None
_lib_lock = threading.Lock()


def library_path():
    return LIB_PATH


def lib():
    """Load libegx.so once; raise NativeLibraryMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m eggpu_amd.build` (hipcc, gfx950)"
            )
        try:
            handle = ctypes.CDLL(LIB_PATH)
        except OSError as exc:
            raise NativeLibraryMissing(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
        return _lib


def err_code_detail(status):
    return lib().egx_err_code_detail(int(status)).decode()


def _check(rc):
    if rc != EGX_SUCC:
        L = lib()
        raise EgxError(rc, L.egx_err_code_detail(rc).decode(), L.egx_last_error().decode())


def device_count():
    return int(lib().egx_device_count())


def _ptr(arr, ptype):
    return None if arr is None else arr.ctypes.data_as(ptype)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class DeviceGraph:
    """A graph resident in HBM: pull CSC, bins and solve workspace.

    Prepared once and reused across calls, the role the reference's
    thread-local PagerankWorkspace plays (``pagerank.cu:44-195``)."""

    def __init__(self, handle, keep=None):
        self._h = _vp(handle)
        self._keep = keep
        self._closed = False

    # -- construction ---------------------------------------------------
    @classmethod
    def from_csr(cls, V, E, W=None, alpha=0.85, keep_csr=False):
        V = _i32(V)
        E = _i32(E)
        W = _f64(W)
        n = len(V) - 1
        h = _vp()
        _check(lib().egx_graph_from_csr(_ptr(V, _i32p), _ptr(E, _i32p), _ptr(W, _f64p), n, len(E),
                                        float(alpha), int(bool(keep_csr)), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr_files(cls, offsets_path, indices_path, weights_path=None, n=None, m=None, alpha=0.85,
                       validate=True, keep_csr=False):
        """eggpu-csr-v1 files streamed straight into HBM and validated there.
        Returns (DeviceGraph, ingest report dict)."""
        import os

        if n is None:
            n = os.path.getsize(offsets_path) // 4 - 1
        if m is None:
            m = os.path.getsize(indices_path) // 4
        rep = IngestReport()
        h = _vp()
        w = os.fsencode(str(weights_path)) if weights_path else None
        _check(lib().egx_graph_from_csr_files(os.fsencode(str(offsets_path)), os.fsencode(str(indices_path)), w,
                                              int(n), int(m), float(alpha), int(bool(validate)),
                                              int(bool(keep_csr)), ctypes.byref(rep), ctypes.byref(h)))
        return cls(h.value), rep.as_dict()

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weighted=False, wlo=0.1, whi=1.0,
             wseed=0, alpha=0.85, keep_csr=False):
        h = _vp()
        _check(lib().egx_graph_rmat(int(scale), int(edge_factor), float(a), float(b), float(c),
                                    int(seed) & (2**64 - 1), int(bool(weighted)), float(wlo), float(whi),
                                    int(wseed) & (2**64 - 1), float(alpha), int(bool(keep_csr)),
                                    ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat_n(cls, n, nedges, a=0.57, b=0.19, c=0.19, seed=1, weighted=False, wlo=0.1, whi=1.0, wseed=0,
               alpha=0.85, keep_csr=False):
        """R-MAT on any n >= 2 nodes with ``nedges`` draws (endpoints >= n dropped)."""
        h = _vp()
        _check(lib().egx_graph_rmat_n(int(n), int(nedges), float(a), float(b), float(c),
                                      int(seed) & (2**64 - 1), int(bool(weighted)), float(wlo), float(whi),
                                      int(wseed) & (2**64 - 1), float(alpha), int(bool(keep_csr)),
                                      ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def er(cls, n, nedges, seed=1, weighted=False, wlo=0.1, whi=1.0, wseed=0, alpha=0.85, keep_csr=False):
        h = _vp()
        _check(lib().egx_graph_er(int(n), int(nedges), int(seed) & (2**64 - 1), int(bool(weighted)),
                                  float(wlo), float(whi), int(wseed) & (2**64 - 1), float(alpha),
                                  int(bool(keep_csr)), ctypes.byref(h)))
        return cls(h.value)

    # -- queries ----------------------------------------------------------
    @property
    def handle(self):
        if self._closed:
            raise ValueError("device graph is closed")
        return self._h

    def info(self):
        out = GraphInfo()
        _check(lib().egx_graph_info_get(self.handle, ctypes.byref(out)))
        return out.as_dict()

    @property
    def n(self):
        return int(self.info()["n"])

    def set_bins(self, thread_row_max=0, wave_row_max=0, hub_segment=0, tile_items=0, reorder=-1, tile_sort=-1,
                 hot_sources=0, blocked=-1, blocked_shared=-1):
        """Re-bin rows (values <= 0 keep the current setting; reorder,
        tile_sort, blocked and blocked_shared < 0 keep theirs; hot_sources 0
        keeps it, < 0 turns non-temporal cold gathers off).  blocked_shared=1
        sizes the blocked layout's work items for graphs solved together in
        multi_pagerank (the largest items); 0 for a graph solved alone."""
        o = BinOpts(int(thread_row_max), int(wave_row_max), int(hub_segment), int(tile_items), int(reorder),
                    int(tile_sort), int(hot_sources), int(blocked), int(blocked_shared))
        _check(lib().egx_graph_set_bins(self.handle, ctypes.byref(o)))

    def download_csc(self):
        inf = self.info()
        n, m = int(inf["n"]), int(inf["m"])
        in_row = np.empty(n + 1, dtype=np.int32)
        in_col = np.empty(max(m, 1), dtype=np.int32)
        in_val = np.empty(max(m, 1), dtype=np.float32)
        dang = np.empty(max(n, 1), dtype=np.int32)
        _check(lib().egx_graph_download_csc(self.handle, _ptr(in_row, _i32p), _ptr(in_col, _i32p),
                                            _ptr(in_val, _f32p), _ptr(dang, _i32p)))
        return in_row, in_col[:m], in_val[:m], dang[:n]

    def download_csr(self, with_weights=True):
        inf = self.info()
        n = int(inf["n"])
        m = ctypes.c_int64(0)
        _check(lib().egx_graph_download_csr(self.handle, None, None, None, ctypes.byref(m)))
        V = np.empty(n + 1, dtype=np.int32)
        E = np.empty(max(m.value, 1), dtype=np.int32)
        W = np.empty(max(m.value, 1), dtype=np.float64) if (with_weights and inf["weighted"]) else None
        _check(lib().egx_graph_download_csr(self.handle, _ptr(V, _i32p), _ptr(E, _i32p), _ptr(W, _f64p),
                                            ctypes.byref(m)))
        return V, E[: m.value], (None if W is None else W[: m.value])

    # -- solves -------------------------------------------------------------
    def pagerank(self, alpha=0.85, max_iter=200, tol=1e-6, precision=PRECISION_FP32, check_every=0,
                 download=True):
        """Returns (float64[n] ranks or None, iterations, kernel_seconds)."""
        n = self.n
        out = np.empty(n, dtype=np.float64) if download else None
        it = ctypes.c_int32(0)
        ks = ctypes.c_double(0.0)
        o = SolveOpts(float(alpha), int(max_iter), float(tol), int(precision), int(check_every))
        _check(lib().egx_graph_pagerank(self.handle, ctypes.byref(o), _ptr(out, _f64p), ctypes.byref(it),
                                        ctypes.byref(ks)))
        return out, int(it.value), float(ks.value)

    def pagerank_stats(self, alpha=0.85, max_iter=200, tol=1e-6, precision=PRECISION_FP32, time_each_launch=True,
                       download=False, hip_graph=False, check_every=0):
        """Like :meth:`pagerank` but returns (ranks or None, stats dict) with
        per-launch kernel timings (hipEvents on the graph's stream) when
        ``time_each_launch``; ``hip_graph`` replays captured runs of
        ``check_every`` iterations instead of launching each kernel."""
        n = self.n
        out = np.empty(n, dtype=np.float64) if download else None
        st = SolveStats()
        o = SolveOpts(float(alpha), int(max_iter), float(tol), int(precision), int(check_every))
        flags = (SOLVE_TIME_LAUNCHES if time_each_launch else 0) | (SOLVE_HIP_GRAPH if hip_graph else 0)
        _check(lib().egx_graph_pagerank_ex(self.handle, ctypes.byref(o), _ptr(out, _f64p), ctypes.byref(st),
                                           flags))
        return out, st.as_dict()

    def ppr_single(self, seeds, alpha=0.85, max_iter=200, tol=1e-6, precision=PRECISION_FP32, download=True):
        """One personalised vector on the plain SpMV (egx_graph_ppr_single):
        teleport and dangling mass uniform over the distinct seeds.
        Returns (float64[n] or None, iterations, kernel_seconds)."""
        seeds = np.ascontiguousarray(seeds, dtype=np.int32)
        out = np.empty(self.n, dtype=np.float64) if download else None
        st = SolveStats()
        o = SolveOpts(float(alpha), int(max_iter), float(tol), int(precision), 0)
        _check(lib().egx_graph_ppr_single(self.handle, _ptr(seeds, _i32p), len(seeds), ctypes.byref(o),
                                          _ptr(out, _f64p), ctypes.byref(st), 0))
        d = st.as_dict()
        return out, int(d["iterations"]), float(d["kernel_seconds"])

    def ppr(self, seed_sets, alpha=0.85, max_iter=200, tol=1e-6, alphas=None, download=True):
        """Batched PageRank: one column per entry of ``seed_sets`` (a list of
        node-id lists; an empty list is a global PageRank), at most 64.
        Returns (float64[k, n] or None, iterations int32[k], kernel_seconds)."""
        k = len(seed_sets)
        if not 1 <= k <= 64:
            raise ValueError("1..64 seed sets per call")
        offs = np.zeros(k + 1, dtype=np.int32)
        offs[1:] = np.cumsum([len(s) for s in seed_sets])
        flat = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.int32) for s in seed_sets])
                                    if offs[-1] else np.zeros(1, np.int32), dtype=np.int32)
        al = None if alphas is None else np.ascontiguousarray(alphas, dtype=np.float64)
        n = self.n
        out = np.empty((k, n), dtype=np.float64) if download else None
        it = np.zeros(k, dtype=np.int32)
        ks = ctypes.c_double(0.0)
        _check(lib().egx_graph_ppr(self.handle, k, _ptr(offs, _i32p), _ptr(flat, _i32p), _ptr(al, _f64p),
                                   float(alpha), int(max_iter), float(tol), _ptr(out, _f64p), _ptr(it, _i32p),
                                   ctypes.byref(ks)))
        return out, it, float(ks.value)

    def batch(self):
        """Open a :class:`Batch` session (64 column slots) on this graph."""
        return Batch(self)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if not self._closed and self._h is not None and _lib is not None:
            _lib.egx_graph_free(self._h)
        self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """64 PageRank column slots resident on one :class:`DeviceGraph`, with
    their own HIP stream: assign / step / poll / result, slots re-assignable
    while other columns keep iterating (continuous batching)."""

    SLOTS = 64

    def __init__(self, graph):
        self.graph = graph
        h = _vp()
        _check(lib().egx_batch_open(graph.handle, ctypes.byref(h)))
        self._h = h
        self._closed = False

    def assign(self, slot, seeds=(), alpha=0.85, max_iter=200, tol=1e-6):
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.int32).reshape(-1))
        _check(lib().egx_batch_assign(self._h, int(slot), _ptr(s, _i32p) if len(s) else None, len(s),
                                      float(alpha), int(max_iter), float(tol)))

    def release(self, slot):
        _check(lib().egx_batch_release(self._h, int(slot)))

    def step(self, iterations=1):
        _check(lib().egx_batch_step(self._h, int(iterations)))

    def poll(self):
        """dict of int32[64] 'active', 'done', 'iterations' and float64[64] 'err'."""
        a = np.zeros(64, np.int32)
        d = np.zeros(64, np.int32)
        it = np.zeros(64, np.int32)
        e = np.zeros(64, np.float64)
        _check(lib().egx_batch_poll(self._h, _ptr(a, _i32p), _ptr(d, _i32p), _ptr(it, _i32p), _ptr(e, _f64p)))
        return {"active": a, "done": d, "iterations": it, "err": e}

    def ready(self):
        """True when all queued work of the session has finished (never blocks)."""
        rc = lib().egx_batch_ready(self._h)
        if rc < 0:
            _check(-rc)
        return rc == 1

    def run(self, max_launches=200, check_every=0):
        ks = ctypes.c_double(0.0)
        _check(lib().egx_batch_run(self._h, int(max_launches), int(check_every), ctypes.byref(ks)))
        return float(ks.value)

    def result(self, slot):
        out = np.empty(self.graph.n, dtype=np.float64)
        _check(lib().egx_batch_result(self._h, int(slot), _ptr(out, _f64p)))
        return out

    def close(self):
        if not self._closed and _lib is not None:
            _lib.egx_batch_close(self._h)
        self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching PageRank server over a pool of DeviceGraphs
    (egx_server_*).  submit() queues a request and returns its id; wait()
    returns (ranks float64[n] or None, info dict).  Identical requests in
    flight together share one column unless coalesce=False."""

    def __init__(self, graphs, hbm_budget_bytes=0, max_sessions=0, step_iters=0, coalesce=True):
        self._graphs = list(graphs)
        k = len(self._graphs)
        handles = (_vp * max(k, 1))(*[g.handle for g in self._graphs])
        o = ServerOpts(int(hbm_budget_bytes), int(max_sessions), int(step_iters), 0 if coalesce else 1, 0)
        h = _vp()
        _check(lib().egx_server_open(handles, k, ctypes.byref(o), ctypes.byref(h)))
        self._h = h.value
        self._outs = {}

    def submit(self, graph, seeds=(), alpha=0.85, max_iter=100, tol=1e-6, download=True):
        seeds = np.ascontiguousarray(seeds, dtype=np.int32)
        out = np.empty(self._graphs[graph].n, dtype=np.float64) if download else None
        rid = ctypes.c_int64(0)
        _check(lib().egx_server_submit(self._h, int(graph), _ptr(seeds, _i32p) if len(seeds) else None, len(seeds),
                                       float(alpha), int(max_iter), float(tol), _ptr(out, _f64p),
                                       ctypes.byref(rid)))
        self._outs[rid.value] = out
        return rid.value

    def wait(self, request_id, timeout=-1.0):
        info = RequestInfo()
        rc = lib().egx_server_wait(self._h, int(request_id), float(timeout), ctypes.byref(info))
        if rc != 0 and timeout >= 0:
            _check(rc)  # may still be running: its buffer stays referenced until close()
        out = self._outs.pop(request_id, None)
        _check(rc)
        return out, info.as_dict()

    def stats(self):
        st = ServerStats()
        _check(lib().egx_server_stats_get(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().egx_server_close(self._h)  # joins the worker: no request writes after this
            self._h = None
            self._outs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def multi_pagerank(graphs, alpha=0.85, max_iter=200, tol=1e-6, precision=PRECISION_FP32, check_every=0,
                   hip_graph=True, download=True, xcd_placement=False):
    """Global PageRank of many device-resident graphs in shared launches.
    Returns (list of float64[n_i] or None, iterations int32[k], stats dict)."""
    k = len(graphs)
    handles = (_vp * max(k, 1))(*[g.handle for g in graphs])
    outs = [np.empty(g.n, dtype=np.float64) for g in graphs] if download else None
    out_ptrs = (_f64p * max(k, 1))(*[_ptr(o, _f64p) for o in outs]) if download else None
    iters = np.zeros(max(k, 1), dtype=np.int32)
    st = SolveStats()
    o = SolveOpts(float(alpha), int(max_iter), float(tol), int(precision), int(check_every))
    flags = (SOLVE_HIP_GRAPH if hip_graph else 0) | (SOLVE_XCD_PLACEMENT if xcd_placement else 0)
    _check(lib().egx_multi_pagerank(handles, k, ctypes.byref(o), out_ptrs, _ptr(iters, _i32p), ctypes.byref(st),
                                    flags))
    return outs, iters[:k], st.as_dict()


def pagerank_csr(V, E, W=None, alpha=0.85, max_iter=200, tol=1e-6, graph_id=0):
    """``egx_pagerank``: the drop-in for ``cuda_pagerank`` with the per-thread
    prepared-graph cache.  Returns (float64[n], iterations, kernel_seconds)."""
    V = _i32(V)
    E = _i32(E)
    W = _f64(W)
    n = len(V) - 1
    out = np.empty(max(n, 0), dtype=np.float64)
    it = ctypes.c_int32(0)
    ks = ctypes.c_double(0.0)
    _check(lib().egx_pagerank(_ptr(V, _i32p), _ptr(E, _i32p), _ptr(W, _f64p), n, len(E), float(alpha),
                              int(max_iter), float(tol), int(graph_id) & (2**64 - 1), _ptr(out, _f64p),
                              ctypes.byref(it), ctypes.byref(ks)))
    return out, int(it.value), float(ks.value)


def katz_csr(V, E, alpha=0.1, beta=1.0, max_iter=1000, tol=1e-6, normalized=True):
    """``egx_katz_csr`` on an out-CSR: (float64[n], iterations, converged,
    kernel_seconds).  beta is a scalar or n values in node order."""
    V = _i32(V)
    E = _i32(E)
    n = len(V) - 1
    vec = None
    scalar = 1.0
    if np.ndim(beta) == 0:
        scalar = float(beta)
    else:
        vec = np.ascontiguousarray(beta, dtype=np.float64)
        if len(vec) != n:
            raise ValueError(f"beta has {len(vec)} values for {n} nodes")
    out = np.empty(max(n, 0), dtype=np.float64)
    it = ctypes.c_int32(0)
    conv = ctypes.c_int32(0)
    ks = ctypes.c_double(0.0)
    _check(lib().egx_katz_csr(_ptr(V, _i32p), _ptr(E, _i32p), n, len(E), float(alpha), _ptr(vec, _f64p), scalar,
                              int(max_iter), float(tol), 1 if normalized else 0, _ptr(out, _f64p),
                              ctypes.byref(it), ctypes.byref(conv), ctypes.byref(ks)))
    return out, int(it.value), bool(conv.value), float(ks.value)


EIG_CONVERGED, EIG_MAX_ITER, EIG_NORM_FLOOR = 1, 2, 3


def eigenvector_csr(V, E, W=None, x0=None, *, max_iter=100, tol=1e-6, delta=2, norm_floor=1e-10, shift=0.0,
                    start_norm=1):
    """``egx_eigenvector_csr`` on an out-CSR (row u lists the heads of u's
    arcs; row v of A x sums over v's in-arcs): (float64[n] last normalised
    iterate, iterations, stop, kernel_seconds).  stop is EIG_CONVERGED,
    EIG_MAX_ITER or EIG_NORM_FLOOR.  x0 defaults to ones."""
    V = _i32(V)
    E = _i32(E)
    n = len(V) - 1
    if W is not None:
        W = np.ascontiguousarray(W, dtype=np.float64)
        if len(W) != len(E):
            raise ValueError(f"{len(W)} weights for {len(E)} arcs")
    x0 = np.ones(max(n, 0), dtype=np.float64) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    if len(x0) != max(n, 0):
        raise ValueError(f"start vector has {len(x0)} values for {n} nodes")
    opts = EigOpts(max_iter=int(max_iter), delta=int(delta), tol=float(tol), norm_floor=float(norm_floor),
                   shift=float(shift), start_norm=int(start_norm), pad=0)
    out = np.empty(max(n, 0), dtype=np.float64)
    it = ctypes.c_int32(0)
    stop = ctypes.c_int32(0)
    ks = ctypes.c_double(0.0)
    _check(lib().egx_eigenvector_csr(_ptr(V, _i32p), _ptr(E, _i32p), _ptr(W, _f64p), n, len(E), _ptr(x0, _f64p),
                                     ctypes.byref(opts), _ptr(out, _f64p), ctypes.byref(it), ctypes.byref(stop),
                                     ctypes.byref(ks)))
    return out, int(it.value), int(stop.value), float(ks.value)


def cache_stats():
    s = CacheStats()
    _check(lib().egx_pagerank_cache_stats(ctypes.byref(s)))
    return {name: int(getattr(s, name)) for name, _ in s._fields_}


def cache_reset():
    lib().egx_pagerank_cache_reset()


def cpp_gpu_pagerank_dense(graph, alpha=0.85, max_iterator=200, threshold=1e-6, weight=None):
    """Shim with the keys of ``cpp_gpu_pagerank_dense`` (``pagerank.cpp:176-223``).

    ``graph`` is a :class:`DeviceGraph` prepared by the backend; ``weight`` is
    accepted for signature compatibility (weights were fixed at build)."""
    values, iters, ks = graph.pagerank(alpha=alpha, max_iter=max_iterator, tol=threshold)
    return {"values_dense": values, "kernel_seconds": ks, "iterations": iters}


def cpp_gpu_pagerank(graph, alpha=0.85, max_iterator=200, threshold=1e-6, weight=None, nodes=None):
    """Shim with the keys of the dict-returning ``cpp_gpu_pagerank``
    (``pagerank.cpp:124-174``): ``{"values": {node: rank}, "kernel_seconds":
    s}``.  Node i of the device graph is ``nodes[i]`` (the backend passes its
    node list; without it the keys are the indices)."""
    values, iters, ks = graph.pagerank(alpha=alpha, max_iter=max_iterator, tol=threshold)
    keys = range(len(values)) if nodes is None else nodes
    if len(keys) != len(values):
        raise ValueError(f"cpp_gpu_pagerank: {len(keys)} nodes for {len(values)} ranks")
    return {"values": dict(zip(keys, values.tolist())), "kernel_seconds": ks, "iterations": iters}
