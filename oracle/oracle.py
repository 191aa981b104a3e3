"""CPU oracle for the PageRank path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, the repository's smoke check and ``bench.py``'s
cpu_baseline leg may import this module, and only to check results (or as
the timed CPU baseline).  The ``eggpu_amd`` package never imports it; its HIP path fails
loudly when the native library is missing instead of falling back here.

Python-side helpers around ``egx_oracle.c``:

* :func:`graph_rows` reproduces how the reference oracle turns a graph into
  index rows: node order ``list(G.nodes.keys())`` (``gpu_runtime.py:18-27``),
  edge rows in ``G.edges`` order (``gpu_runtime.py:51-58``), and for
  undirected graphs forward rows without self-loops followed by the reversed
  rows (``pagerank.py:87-91``).
* :func:`pagerank_graph` is ``_pagerank_power_iteration``
  (``pagerank.py:75-139``) on top of the C restatement.
* :func:`csr_from_graph` is the reference's ``Graph::gen_CSR``
  (``graph.cpp:719-886``): rows in node-index order, columns ascending,
  weights aligned; undirected graphs get both directions and no self-loops
  (``graph_convert.cpp:72-108``), directed graphs keep self-loops
  (``graph_convert.cpp:110-134``).
* :func:`rmat_csr` / :func:`er_csr` are this project's seeded generators,
  the definitions the HIP generator must reproduce bit for bit.
* :func:`katz_rows` / :func:`katz_graph` restate the CPU Katz centrality
  ``katz_centrality.py:7-110`` in FP64 with a sparse adjacency.
* :func:`eigenvector_rows` / :func:`eigenvector_graph` restate the CPU
  eigenvector centrality (``eigenvector.py:62-154`` and the power iteration
  of ``eigenvector.cpp:33-97``) in FP64.

Functions with no independent check are marked "parity unpinned" in their
docstring; every other function is pinned in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


def build():
    """Compile ``liboracle.so`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_pagerank_push.restype = ctypes.c_int
    lib.orc_pagerank_push.argtypes = [
        ctypes.c_int64, ctypes.c_int64, _i32p, _i32p, _f64p, ctypes.c_double,
        ctypes.c_int32, ctypes.c_double, _f64p, _i32p, _f64p,
    ]
    lib.orc_ppr_push.restype = ctypes.c_int
    lib.orc_ppr_push.argtypes = [
        ctypes.c_int64, ctypes.c_int64, _i32p, _i32p, _f64p, ctypes.c_double,
        ctypes.c_int32, ctypes.c_double, _f64p, _f64p, _i32p, _f64p,
    ]
    lib.orc_build_inbound.restype = ctypes.c_int64
    lib.orc_build_inbound.argtypes = [
        ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f64p, ctypes.c_double,
        _i32p, _i32p, _f32p, _i32p, _f64p,
    ]
    lib.orc_pagerank_omp.restype = ctypes.c_int
    lib.orc_pagerank_omp.argtypes = [
        ctypes.c_int32, _i32p, _i32p, _f64p, ctypes.c_double, ctypes.c_int32,
        ctypes.c_double, ctypes.c_int32, _f64p, _i32p,
    ]
    lib.orc_omp_max_threads.restype = ctypes.c_int
    lib.orc_rmat_coo.restype = None
    lib.orc_rmat_coo.argtypes = [
        ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _i32p, _i32p,
    ]
    lib.orc_er_coo.restype = None
    lib.orc_er_coo.argtypes = [
        ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, _i32p, _i32p,
    ]
    lib.orc_edge_weights.restype = None
    lib.orc_edge_weights.argtypes = [
        ctypes.c_int64, _i32p, _i32p, ctypes.c_uint64, ctypes.c_double,
        ctypes.c_double, _f64p,
    ]
    _lib = lib
    return lib


def _p(arr, ptype):
    if arr is None:
        return None
    return arr.ctypes.data_as(ptype)


# --------------------------------------------------------------------------
# graph -> rows / CSR, as the reference does it
# --------------------------------------------------------------------------
def graph_nodes(G):
    nodes = G.nodes
    return list(nodes.keys()) if hasattr(nodes, "keys") else list(nodes)


def _edge_triples(G):
    for e in G.edges:
        if len(e) == 3:
            yield e
        else:
            yield (e[0], e[1], {})


def graph_rows(G, weight=None):
    """Index rows of ``pagerank.py:81-91``: (src, dst, w-or-None) arrays."""
    nodes = graph_nodes(G)
    idx = {v: i for i, v in enumerate(nodes)}
    src, dst, w = [], [], []
    for u, v, data in _edge_triples(G):
        src.append(idx[u])
        dst.append(idx[v])
        if weight is not None:
            w.append(float(data.get(weight, 1)))
    src = np.asarray(src, dtype=np.int32)
    dst = np.asarray(dst, dtype=np.int32)
    wa = np.asarray(w, dtype=np.float64) if weight is not None else None
    if not G.is_directed():
        keep = src != dst
        fs, fd = src[keep], dst[keep]
        src = np.concatenate([fs, fd])
        dst = np.concatenate([fd, fs])
        if wa is not None:
            fw = wa[keep]
            wa = np.concatenate([fw, fw])
    return nodes, src, dst, wa


def csr_from_graph(G, weight=None):
    """Out-CSR of ``Graph::gen_CSR`` (``graph.cpp:719-886``): V, E, W, nodes."""
    nodes = graph_nodes(G)
    n = len(nodes)
    idx = {v: i for i, v in enumerate(nodes)}
    src, dst, w = [], [], []
    seen = set()
    for u, v, data in _edge_triples(G):
        iu, iv = idx[u], idx[v]
        if not G.is_directed():
            if iu == iv:
                continue
            key = (min(iu, iv), max(iu, iv))
            if key in seen:
                continue
            seen.add(key)
            wv = float(data.get(weight, 1.0)) if weight is not None else 1.0
            src += [iu, iv]
            dst += [iv, iu]
            w += [wv, wv]
        else:
            src.append(iu)
            dst.append(iv)
            w.append(float(data.get(weight, 1.0)) if weight is not None else 1.0)
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    w = np.asarray(w, dtype=np.float64)
    order = np.lexsort((dst, src))
    E = dst[order].astype(np.int32)
    W = w[order]
    V = np.zeros(n + 1, dtype=np.int32)
    if n:
        V[1:] = np.cumsum(np.bincount(src, minlength=n))
    return V, E, W, nodes


def csr_rows(V, E):
    """Expand CSR to COO rows in CSR order."""
    V = np.asarray(V, dtype=np.int64)
    n = len(V) - 1
    src = np.repeat(np.arange(n, dtype=np.int32), np.diff(V))
    return src, np.asarray(E, dtype=np.int32)


# --------------------------------------------------------------------------
# PageRank oracles
# --------------------------------------------------------------------------
def pagerank_rows(n, src, dst, w=None, alpha=0.85, max_iter=200, tol=1e-6):
    """FP64 restatement of ``pagerank.py:75-139`` on index rows.

    Returns (ranks float64[n], iterations, last L1 residual)."""
    lib = load()
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    it = ctypes.c_int32(0)
    err = ctypes.c_double(0.0)
    rc = lib.orc_pagerank_push(
        n, len(src), _p(src, _i32p), _p(dst, _i32p), _p(wa, _f64p), float(alpha),
        int(max_iter), float(tol), _p(out, _f64p), ctypes.byref(it), ctypes.byref(err),
    )
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out, int(it.value), float(err.value)


def pagerank_graph(G, alpha=0.85, weight=None, max_iter=200, tol=1e-6):
    """``_pagerank_power_iteration`` (``pagerank.py:75-139``) -> (dict, iters)."""
    nodes, src, dst, w = graph_rows(G, weight)
    n = len(nodes)
    if n == 0:
        return {}, 0
    pr, it, _ = pagerank_rows(n, src, dst, w, alpha, max_iter, tol)
    return {nodes[i]: float(pr[i]) for i in range(n)}, it


def pagerank_csr(V, E, W=None, alpha=0.85, max_iter=200, tol=1e-6):
    """The oracle on an out-CSR (rows in CSR order)."""
    src, dst = csr_rows(V, E)
    return pagerank_rows(len(V) - 1, src, dst, W, alpha, max_iter, tol)


def ppr_rows(n, src, dst, w, p, alpha=0.85, max_iter=200, tol=1e-6):
    """Personalised PageRank, networkx 3.4.2 semantics (teleport and dangling
    mass follow ``p``), FP64, push order of :func:`pagerank_rows`."""
    lib = load()
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    pa = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    it = ctypes.c_int32(0)
    err = ctypes.c_double(0.0)
    rc = lib.orc_ppr_push(
        n, len(src), _p(src, _i32p), _p(dst, _i32p), _p(wa, _f64p), float(alpha),
        int(max_iter), float(tol), _p(pa, _f64p), _p(out, _f64p), ctypes.byref(it),
        ctypes.byref(err),
    )
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out, int(it.value), float(err.value)


def build_inbound(V, E, W=None, alpha=0.85):
    """``build_inbound_from_csr`` (``pagerank.cu:286-347``).

    Returns in_row, in_col, in_val (float32), dangling01, out_weight_sum."""
    lib = load()
    V = np.ascontiguousarray(V, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.int32)
    n = len(V) - 1
    m = len(E)
    Wa = None if W is None else np.ascontiguousarray(W, dtype=np.float64)
    # Rows are clamped to [0, m] one by one (pagerank.cu:302-306), so
    # malformed offsets can yield more stored entries than m.
    begin = np.clip(V[:-1].astype(np.int64), 0, m)
    end = np.minimum(V[1:].astype(np.int64), m)
    begin = np.minimum(begin, end)
    cap = int(max((end - begin).sum(), 1)) if n > 0 else 1
    in_row = np.empty(n + 1, dtype=np.int32)
    in_col = np.empty(cap, dtype=np.int32)
    in_val = np.empty(cap, dtype=np.float32)
    dang = np.empty(max(n, 1), dtype=np.int32)
    outw = np.empty(max(n, 1), dtype=np.float64)
    total = lib.orc_build_inbound(
        n, m, _p(V, _i32p), _p(E, _i32p), _p(Wa, _f64p), float(alpha),
        _p(in_row, _i32p), _p(in_col, _i32p), _p(in_val, _f32p), _p(dang, _i32p),
        _p(outw, _f64p),
    )
    if total < 0:
        raise MemoryError("oracle allocation failed")
    return in_row, in_col[:total], in_val[:total], dang[:n], outw[:n]


def pagerank_omp(V, E, W=None, alpha=0.85, max_iter=200, tol=1e-6, threads=0):
    """OpenMP push of ``pagerank.cpp:21-122`` (CPU baseline, 'port').

    Summation order is non-deterministic, as in the reference.  Returns
    (unnormalised ranks, iterations)."""
    lib = load()
    V = np.ascontiguousarray(V, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.int32)
    Wa = None if W is None else np.ascontiguousarray(W, dtype=np.float64)
    n = len(V) - 1
    out = np.empty(n, dtype=np.float64)
    it = ctypes.c_int32(0)
    rc = lib.orc_pagerank_omp(
        n, _p(V, _i32p), _p(E, _i32p), _p(Wa, _f64p), float(alpha), int(max_iter),
        float(tol), int(threads), _p(out, _f64p), ctypes.byref(it),
    )
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out, int(it.value)


def omp_max_threads():
    return int(load().orc_omp_max_threads())


# --------------------------------------------------------------------------
# generators (this project's definitions)
# --------------------------------------------------------------------------
def rmat_thresholds(a, b, c):
    """floor(cumulative probability * 2^32), clamped to uint32."""
    def t(x):
        return int(min(max(np.floor(x * 4294967296.0), 0), 4294967295))
    return t(a), t(a + b), t(a + b + c)


def rmat_coo(scale, nedges, seed, a=0.57, b=0.19, c=0.19):
    lib = load()
    tA, tAB, tABC = rmat_thresholds(a, b, c)
    src = np.empty(nedges, dtype=np.int32)
    dst = np.empty(nedges, dtype=np.int32)
    lib.orc_rmat_coo(int(scale), 0, int(nedges), int(seed) & (2**64 - 1), tA, tAB, tABC,
                     _p(src, _i32p), _p(dst, _i32p))
    return src, dst


def er_coo(n, nedges, seed):
    lib = load()
    src = np.empty(nedges, dtype=np.int32)
    dst = np.empty(nedges, dtype=np.int32)
    lib.orc_er_coo(int(n), 0, int(nedges), int(seed) & (2**64 - 1), _p(src, _i32p), _p(dst, _i32p))
    return src, dst


def simple_csr_from_coo(n, src, dst):
    """Drop self-loops and duplicate arcs; rows ascending, columns ascending."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    keep = src != dst
    keys = np.unique((src[keep] << 32) | dst[keep])
    s = (keys >> 32).astype(np.int64)
    E = (keys & 0xFFFFFFFF).astype(np.int32)
    V = np.zeros(n + 1, dtype=np.int32)
    V[1:] = np.cumsum(np.bincount(s, minlength=n))
    return V, E


def edge_weights(V, E, wseed, lo=0.1, hi=1.0):
    lib = load()
    src, dst = csr_rows(V, E)
    w = np.empty(len(E), dtype=np.float64)
    lib.orc_edge_weights(len(E), _p(src, _i32p), _p(np.ascontiguousarray(dst), _i32p),
                         int(wseed) & (2**64 - 1), float(lo), float(hi), _p(w, _f64p))
    return w


def rmat_csr(scale, edge_factor=16, seed=1, a=0.57, b=0.19, c=0.19):
    """Directed R-MAT: 2^scale nodes, edge_factor*2^scale draws, self-loops
    and duplicates removed, CSR sorted by (src, dst)."""
    n = 1 << scale
    src, dst = rmat_coo(scale, edge_factor * n, seed, a, b, c)
    return simple_csr_from_coo(n, src, dst)


def rmat_n_csr(n, nedges, seed=1, a=0.57, b=0.19, c=0.19):
    """R-MAT on any n >= 2 nodes: nedges draws of ceil(log2 n) quadrant
    levels; draws with an endpoint >= n are dropped, as are self-loops and
    duplicates.  For n = 2^scale this is rmat_csr."""
    scale = max(int(n - 1).bit_length(), 1)
    src, dst = rmat_coo(scale, nedges, seed, a, b, c)
    keep = (src < n) & (dst < n)
    return simple_csr_from_coo(n, src[keep], dst[keep])


def er_csr(n, nedges, seed):
    src, dst = er_coo(n, nedges, seed)
    return simple_csr_from_coo(n, src, dst)


def adjacency_csr_from_graph(G):
    """A of ``katz_centrality.py:73-81``: node order ``list(G.nodes)``, one
    entry A[u][v] = 1 per adjacency key v of u (self-loops included, both
    directions of an undirected edge).  Returns (n, src, dst) index arrays."""
    nodes = list(G.nodes)
    idx = {node: i for i, node in enumerate(nodes)}
    src, dst = [], []
    for u in nodes:
        for v in G.adj[u]:
            src.append(idx[u])
            dst.append(idx[v])
    return len(nodes), np.asarray(src, dtype=np.int64), np.asarray(dst, dtype=np.int64)


def katz_rows(n, src, dst, alpha=0.1, beta=1.0, max_iter=1000, tol=1e-6, normalized=True):
    """``katz_centrality.py:83-110`` in FP64: x = 1; x_new = alpha*A@x + b;
    stop when ||x_new - x||_1 < tol keeping the previous x; L2-normalise.
    beta is a scalar or an array of n values.  Returns (x, iterations) or
    (None, max_iter) when the loop ends without the test holding (the
    reference raises RuntimeError)."""
    import scipy.sparse as sp

    A = sp.csr_matrix((np.ones(len(src)), (np.asarray(src), np.asarray(dst))), shape=(n, n))
    x = np.ones(n, dtype=np.float64)
    b = np.asarray(beta, dtype=np.float64) * np.ones(n) if np.ndim(beta) else np.ones(n) * float(beta)
    for it in range(max_iter):
        x_new = alpha * (A @ x) + b
        if np.linalg.norm(x_new - x, ord=1) < tol:
            break
        x = x_new
    else:
        return None, max_iter
    if normalized:
        norm = np.linalg.norm(x)
        if norm > 0:
            x = x / norm
    return x, it + 1


def katz_graph(G, alpha=0.1, beta=1.0, max_iter=1000, tol=1e-6, normalized=True):
    """:func:`katz_rows` on a graph; beta may be a dict node -> value
    (default 1.0).  Returns ({node: value} or None, iterations)."""
    nodes = list(G.nodes)
    n, src, dst = adjacency_csr_from_graph(G)
    b = np.array([beta.get(v, 1.0) for v in nodes]) if isinstance(beta, dict) else beta
    x, it = katz_rows(n, src, dst, alpha, b, max_iter, tol, normalized)
    if x is None:
        return None, it
    return {v: float(x[i]) for i, v in enumerate(nodes)}, it


def eigenvector_rows(n, src, dst, w=None, x0=None, max_iter=100, tol=1e-6, delta=2, norm_floor=1e-10, shift=0.0,
                     start_norm=1):
    """The power iterations of eigenvector centrality in FP64, on
    A[v][u] = w for every arc u -> v (duplicates summed, as
    ``eigenvector.py:92`` builds it):

    * ``eigenvector.py:110-131`` (delta=2, start_norm=1, norm_floor=1e-10):
      v = x0 / sum|x0|; y = A v; stop keeping v when ||y|| < floor; v' = y/||y||;
      stop when ||v' - v||_2 < tol;
    * ``eigenvector.cpp:33-97`` (delta=1, start_norm=2, norm_floor=1e-12):
      x = x0 * (1/||x0||) (uniform 1/sqrt(n) below 1e-12); x' = y * (1/||y||);
      stop when sum|x' - x| < tol.

    shift > 0 iterates A + shift*I.  Returns (x, iterations, stop): stop 1 the
    delta test held, 2 max_iter iterations ran without it, 3 the norm test
    stopped at that iteration (x is the iterate before it)."""
    import scipy.sparse as sp

    w = np.ones(len(src)) if w is None else np.asarray(w, dtype=np.float64)
    A = sp.csr_matrix((w, (np.asarray(dst), np.asarray(src))), shape=(n, n))
    v = np.ones(n, dtype=np.float64) if x0 is None else np.array(x0, dtype=np.float64)
    if start_norm == 1:
        v = v / np.sum(np.abs(v))
    else:
        nrm = np.sqrt(np.sum(v * v))
        v = np.full(n, 1.0 / np.sqrt(n)) if nrm < 1e-12 else v * (1.0 / nrm)
    for k in range(1, max_iter + 1):
        last = v
        y = A @ last
        if shift:
            y = y + shift * last
        nrm = np.linalg.norm(y)
        if nrm < norm_floor:
            return last, k, 3
        v = y * (1.0 / nrm) if delta == 1 else y / nrm
        d = np.sum(np.abs(v - last)) if delta == 1 else np.linalg.norm(v - last)
        if d < tol:
            return v, k, 1
    return v, max_iter, 2


def eigenvector_graph(G, max_iter=100, tol=1e-6, nstart=None, weight=None):
    """``eigenvector_centrality`` of ``eigenvector.py:62-154`` on a graph:
    A from ``G[u].items()`` in ``list(G.nodes)`` order with weight
    ``attrs.get(weight, 1.0)``, a 1e-4 self-loop on every node whose row and
    column sums are both 0, the power iteration when nstart is given
    (:func:`eigenvector_rows`, raising RuntimeError where the reference
    raises EasyGraphError), else ``scipy.sparse.linalg.eigs(A, k=1,
    which='LR', maxiter=max_iter, tol=tol)`` as the reference calls it; then
    the sign fix, L2 normalisation, and isolated nodes set to 0 with a
    renormalisation.  Raises ValueError for fewer than two nodes."""
    import scipy.sparse as sp
    from scipy.sparse import linalg as sla

    nodes = list(G.nodes)
    n = len(nodes)
    if n < 2:
        raise ValueError("eigenvector centrality needs at least two nodes")
    idx = {node: i for i, node in enumerate(nodes)}
    src, dst, w = [], [], []
    for u in nodes:
        for v, attrs in G[u].items():
            src.append(idx[u])
            dst.append(idx[v])
            w.append(float(attrs.get(weight, 1.0)) if weight else 1.0)
    A = sp.csr_matrix((w, (dst, src)), shape=(n, n))
    isolated = np.where((np.asarray(A.sum(axis=1)).ravel() == 0) & (np.asarray(A.sum(axis=0)).ravel() == 0))[0]
    if len(isolated):
        src = src + [int(i) for i in isolated]
        dst = dst + [int(i) for i in isolated]
        w = w + [1.0e-4] * len(isolated)
        A = sp.csr_matrix((w, (dst, src)), shape=(n, n))
    if nstart is not None:
        x0 = np.array([nstart.get(v, 1.0) for v in nodes], dtype=float)
        c, _, stop = eigenvector_rows(n, np.asarray(src), np.asarray(dst), np.asarray(w), x0, max_iter, tol)
        if stop == 2:
            raise RuntimeError(f"Eigenvector calculation did not converge in {max_iter} iterations")
    else:
        _, vecs = sla.eigs(A, k=1, which="LR", maxiter=max_iter, tol=tol)
        c = np.real(vecs[:, 0])
    if c.sum() < 0:
        c = -c
    c = c / np.linalg.norm(c)
    if len(isolated):
        c[isolated] = 0.0
        if np.sum(c) > 0:
            c = c / np.linalg.norm(c)
    return {nodes[i]: float(c[i]) for i in range(n)}
