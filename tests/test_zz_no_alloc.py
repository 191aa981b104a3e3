"""GPU test of egx_graph_pagerank_noalloc, the allocation-free form of
egx_graph_pagerank: called through the C ABI, its ranks must equal the
existing entry point's bit for bit (same kernels, same summation order),
with the same iteration count, on two calls in a row -- the second on a
smaller graph, so the shared scratch array holds the first call's
contents -- in FP32 and FP64, on the pull and the propagation-blocked
iteration.  A graph past the array's size is refused before any device work.

The device output buffers are the test's own (hipMalloc through the HIP
runtime); the library call itself allocates nothing.
"""
import ctypes

import numpy as np
import pytest

from eggpu_amd import native

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


class _DeviceVector:
    """n doubles of device memory owned by the test."""

    def __init__(self, n):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0
        # poison: the entry point must overwrite every entry
        assert self.hip.hipMemset(self.ptr, 0xFF, ctypes.c_size_t(8 * n)) == 0

    def download(self):
        out = np.empty(self.n, dtype=np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _noalloc(g, alpha, max_iter, tol, precision):
    out = _DeviceVector(g.n)
    try:
        it = ctypes.c_int32(-1)
        o = native.SolveOpts(float(alpha), int(max_iter), float(tol), int(precision), 0)
        rc = native.lib().egx_graph_pagerank_noalloc(g.handle, ctypes.byref(o), out.ptr, ctypes.byref(it))
        native._check(rc)
        return out.download(), int(it.value)
    finally:
        out.free()


CASES = [
    # (scale, seed, blocked, precision, tol, max_iter)
    (15, 3, 0, native.PRECISION_FP32, 1e-6, 100),
    (13, 4, 0, native.PRECISION_FP32, 1e-6, 100),
    (15, 5, 1, native.PRECISION_FP32, 0.0, 12),
    (12, 6, 1, native.PRECISION_FP32, 1e-6, 100),
    (14, 7, 0, native.PRECISION_FP64, 1e-8, 100),
    (11, 8, 0, native.PRECISION_FP64, 0.0, 9),
]


def test_noalloc_equals_the_allocating_entry_point(gpu_ready):
    for scale, seed, blocked, prec, tol, max_iter in CASES:
        _say(f"R-MAT {scale} seed {seed} blocked={blocked} precision={prec} tol={tol} K={max_iter}")
        g = native.DeviceGraph.rmat(scale, 16, seed=seed)
        g.set_bins(blocked=blocked)
        assert g.info()["blocked"] == blocked
        a, ia = _noalloc(g, 0.85, max_iter, tol, prec)
        ref, it_ref, _ = g.pagerank(0.85, max_iter, tol, precision=prec)
        assert ia == it_ref
        assert np.array_equal(a.view(np.uint64), ref.view(np.uint64)), f"max-abs {np.abs(a - ref).max():.3e}"
        # a second call on the same graph after the other entry point ran
        b, ib = _noalloc(g, 0.85, max_iter, tol, prec)
        assert ib == ia and np.array_equal(b.view(np.uint64), a.view(np.uint64))
        g.close()


def test_noalloc_weighted_and_alpha_change(gpu_ready):
    src = native.DeviceGraph.rmat(13, 16, seed=9, keep_csr=True)
    V, E, _ = src.download_csr(with_weights=False)
    src.close()
    W = np.random.default_rng(10).uniform(0.1, 1.0, len(E))
    g = native.DeviceGraph.from_csr(V, E, W, alpha=0.85)
    for alpha in (0.85, 0.9):
        a, ia = _noalloc(g, alpha, 100, 1e-6, native.PRECISION_FP32)
        ref, it_ref, _ = g.pagerank(alpha, 100, 1e-6)
        assert ia == it_ref and np.array_equal(a.view(np.uint64), ref.view(np.uint64))
    g.close()


def test_noalloc_refuses_graphs_past_its_array(gpu_ready):
    # R-MAT 22 has 4.2M nodes: 16 bytes per node of FP32 iterates exceed 64 MiB
    g = native.DeviceGraph.rmat(22, 4, seed=2)
    out = _DeviceVector(g.n)  # room for every rank, so even a call that ran would stay in bounds
    try:
        it = ctypes.c_int32(-1)
        o = native.SolveOpts(0.85, 5, 0.0, native.PRECISION_FP32, 0)
        rc = native.lib().egx_graph_pagerank_noalloc(g.handle, ctypes.byref(o), out.ptr, ctypes.byref(it))
        assert rc == native.EGX_UNSUPPORTED_GRAPH
        msg = native.lib().egx_last_error().decode()
        assert "per-solve scratch" in msg and str(64 << 20) in msg
    finally:
        out.free()
        g.close()
