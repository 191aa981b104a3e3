"""PageRank slice of the GPU backend, over the MI355X library.

Mirrors the contract of ``easygraph/utils/gpu_eggpu_backend.py`` for
PageRank, so this module can stand where the reference backend stands:

* ``pagerank(G, alpha=0.85, max_iter=200, eps=1e-6, weight=None)``
  (``gpu_eggpu_backend.py:1782-1821``): context, opt-in result cache, native
  call, read-only dense mapping, ``NotImplementedError`` for weighted
  PageRank on a topology-only bulk graph (``:1784-1788``);
* the prepared graph context and its validity rules
  (``:1590-1716``): mutation generation, node count, and an O(|V|) stamp
  re-checked every ``EASYGRAPH_GPU_CTX_VALIDATE_EVERY``-th reuse;
* the dense-result contract (``_normalize_dense_array`` ``:1092-1126``,
  ``_DenseValueDict`` ``:254-348``): exact length or a leading sentinel,
  otherwise fail closed.

What differs is below the native call: instead of a pybind C++ graph and
``cuda_pagerank``, the graph's out-CSR is built with numpy in
``Graph::gen_CSR`` order and handed to ``libegx.so`` once; the pull CSC,
bins and iteration workspace stay resident in HBM on a
:class:`eggpu_amd.native.DeviceGraph` cached in ``G.cache`` until the graph
mutates.
"""
from __future__ import annotations

import copy
import importlib
import os
import time
import weakref
from collections.abc import Mapping

import itertools

import numpy as np

from .gpu_runtime import gpu_runtime_enabled


def _env_true(name, default="FALSE"):
    return os.environ.get(name, default).strip().upper() in {"1", "TRUE", "ON", "YES"}


_LAST_KERNEL_SECONDS = {}
_LAST_ITERATIONS = {}
_PREPARED_CACHE_KEY = "__eggpu_prepared_graph__"
_DEVICE_GRAPH_KEY = "__egx_device_graph__"
_HASH_MASK = (1 << 64) - 1
_RESULT_CACHE_ENABLED = _env_true("EASYGRAPH_GPU_RESULT_CACHE")
_RESULT_CACHE_RETURN_COPY = _env_true("EASYGRAPH_GPU_RESULT_CACHE_RETURN_COPY", "TRUE")
_RESULT_CACHE_MAX_ITEMS = max(1, int(os.environ.get("EASYGRAPH_GPU_RESULT_CACHE_MAX_ITEMS", "8")))
_RESULT_CACHE_MAX_NODES = max(0, int(os.environ.get("EASYGRAPH_GPU_RESULT_CACHE_MAX_NODES", "2000000")))
_GRAPH_CONTEXT_CACHE_DISABLED = _env_true("EASYGRAPH_GPU_DISABLE_GRAPH_CONTEXT_CACHE", "")
_CPP_GRAPH_CACHE_DISABLED = _env_true("EASYGRAPH_GPU_DISABLE_CPP_GRAPH_CACHE", "")
_GRAPH_CTX_VALIDATE_EVERY = max(1, int(os.environ.get("EASYGRAPH_GPU_CTX_VALIDATE_EVERY", "32")))
_FALLBACK_GRAPH_CACHE = weakref.WeakKeyDictionary()


def eggpu_backend_enabled():
    return gpu_runtime_enabled()


def set_last_kernel_time(key, seconds):
    if seconds is None:
        _LAST_KERNEL_SECONDS.pop(key, None)
    else:
        _LAST_KERNEL_SECONDS[key] = float(seconds)


def get_last_kernel_time(key):
    return _LAST_KERNEL_SECONDS.get(key)


def get_last_iterations(key):
    """Iterations the last solve performed (new; the reference does not expose it)."""
    return _LAST_ITERATIONS.get(key)


def _native():
    # Looked up through sys.modules so tests can substitute a fake module,
    # as the reference's tests do for cpp_easygraph.
    return importlib.import_module("eggpu_amd.native")


def _is_bulk_graph(G):
    return bool(getattr(G, "_eggpu_bulk_csr", False))


def _bulk_weight_available(G, weight):
    return (
        _is_bulk_graph(G)
        and weight is not None
        and getattr(G, "weights_path", None) is not None
        and getattr(G, "weight_key", None) == str(weight)
    )


# ---------------------------------------------------------------------------
# dense results
# ---------------------------------------------------------------------------
class _DenseResultShapeError(ValueError):
    """A native dense result violates its documented shape."""


def _dense_shape_message(result_name, expected_shapes, actual_shape):
    expected = " or ".join(str(tuple(s)) for s in expected_shapes)
    return f"EGGPU {result_name} returned invalid dense shape {tuple(actual_shape)}; expected {expected}"


def _normalize_dense_array(raw_values, target_len, dtype="float64", *, allow_leading_sentinel=False,
                           result_name="dense vector"):
    """Accept exactly ``(target_len,)`` or, when allowed, ``(target_len+1,)``
    with a leading sentinel that is dropped; anything else raises."""
    try:
        arr = np.asarray(raw_values, dtype=dtype)
    except Exception as exc:
        raise _DenseResultShapeError(
            f"EGGPU {result_name} could not be converted to a one-dimensional dense result: {exc}"
        ) from exc
    exact = (int(target_len),)
    with_sentinel = (int(target_len) + 1,)
    if arr.ndim == 1 and arr.shape == exact:
        return arr
    if allow_leading_sentinel and arr.ndim == 1 and arr.shape == with_sentinel:
        return arr[1:]
    shapes = [exact] + ([with_sentinel] if allow_leading_sentinel else [])
    raise _DenseResultShapeError(_dense_shape_message(result_name, shapes, arr.shape))


class _DenseValueDict(Mapping):
    """Read-only node -> value mapping over a dense float64 vector.

    A ``Mapping`` and not a ``dict`` subclass: equality, ``dict(result)``,
    pickling and iteration behave like the logical dict, mutation raises
    ``TypeError``, and no per-node Python objects exist until asked for."""

    def __init__(self, nodes, values, node_to_idx=None, dtype=float):
        self._nodes = nodes if isinstance(nodes, range) else list(nodes)
        self._values = _normalize_dense_array(values, len(self._nodes), dtype=dtype,
                                              result_name="dense node-value mapping")
        self._values.setflags(write=False)
        self._node_to_idx = node_to_idx if isinstance(node_to_idx, dict) else None
        self._range_ids = (isinstance(self._nodes, range) and self._nodes.start == 0
                           and self._nodes.step == 1)
        self._cast = int if self._values.dtype.kind in "iu" else float
        self._lazy_index = None

    def _pos(self, key):
        if self._node_to_idx is not None:
            return self._node_to_idx.get(key)
        if self._range_ids:
            if isinstance(key, (int, np.integer)) and not isinstance(key, bool) and 0 <= key < len(self._nodes):
                return int(key)
            return None
        if self._lazy_index is None:
            self._lazy_index = {node: i for i, node in enumerate(self._nodes)}
        try:
            return self._lazy_index.get(key)
        except TypeError:
            return None

    def __len__(self):
        return len(self._nodes)

    def __iter__(self):
        return iter(self._nodes)

    def __contains__(self, key):
        return self._pos(key) is not None

    def __getitem__(self, key):
        i = self._pos(key)
        if i is None:
            raise KeyError(key)
        return self._cast(self._values[int(i)])

    def get(self, key, default=None):
        i = self._pos(key)
        return default if i is None else self._cast(self._values[int(i)])

    def keys(self):
        return self._nodes if self._range_ids else list(self._nodes)

    def values(self):
        return self._values

    def items(self):
        for i, node in enumerate(self._nodes):
            yield node, self._cast(self._values[i])

    def copy(self):
        return dict(self.items())

    def to_dict(self):
        return dict(self.items())

    def to_numpy(self, copy=False):
        return self._values.copy() if copy else self._values

    def tolist(self):
        return self._values.tolist()

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self._values, dtype=dtype)

    def __reduce__(self):
        return (_DenseValueDict, (self._nodes, np.array(self._values), self._node_to_idx))

    def __repr__(self):
        if len(self._nodes) > 1000:
            return (f"{type(self).__name__}(nodes={len(self._nodes)}, dtype={self._values.dtype}, "
                    "materialized=False)")
        return repr(self.to_dict())


# ---------------------------------------------------------------------------
# prepared graph context (gpu_eggpu_backend.py:1590-1716)
# ---------------------------------------------------------------------------
def _graph_cache_dict(G):
    cache = getattr(G, "cache", None)
    if isinstance(cache, dict):
        return cache
    try:
        cache = _FALLBACK_GRAPH_CACHE.get(G)
        if cache is None:
            cache = {}
            _FALLBACK_GRAPH_CACHE[G] = cache
        return cache
    except TypeError:
        return None


def _adj_dict(G):
    if hasattr(G, "_adj"):
        return G._adj
    return G.adj


def _mix64(h, x):
    h ^= (int(x) + 0x9E3779B97F4A7C15 + ((h << 6) & _HASH_MASK) + (h >> 2)) & _HASH_MASK
    return h & _HASH_MASK


def _graph_stamp(G):
    """O(|V|) fingerprint: node count, arc slots, and a position-dependent
    hash of every node's hash and degree plus the first neighbours of the
    first 32 nodes.  Node hashes and degrees are gathered at C speed (map)
    and mixed with wrapping uint64 arithmetic in numpy."""
    adj = _adj_dict(G)
    n = len(adj)
    hu = np.fromiter(map(hash, adj), dtype=np.int64, count=n).view(np.uint64)
    dg = np.fromiter(map(len, adj.values()), dtype=np.int64, count=n).view(np.uint64)
    slots = int(dg.sum(dtype=np.uint64)) if n else 0
    with np.errstate(over="ignore"):
        x = hu * np.uint64(0x9E3779B97F4A7C15) + dg * np.uint64(0xC2B2AE3D27D4EB4F) + \
            np.arange(n, dtype=np.uint64) * np.uint64(0x165667B19E3779F9)
        x ^= x >> np.uint64(31)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(29)
    h = 1469598103934665603
    if n:
        h = _mix64(h, int(np.bitwise_xor.reduce(x)))
        h = _mix64(h, int(x.sum(dtype=np.uint64)))
    for i, (u, nbrs) in enumerate(adj.items()):
        if i >= 32:
            break
        for j, v in enumerate(nbrs):
            if j >= 4:
                break
            h = _mix64(h, hash(v))
    return (bool(G.is_directed()), n, slots, h)


def _prepared_graph_usable(G, prepared):
    if not isinstance(prepared, dict):
        return False
    if bool(G.is_directed()) != bool(prepared.get("directed")):
        return False
    if _is_bulk_graph(G):
        return prepared.get("bulk_signature") == getattr(G, "_eggpu_bulk_signature", None)
    gen = getattr(G, "_mutation_generation", None)
    if gen is not None and prepared.get("mutation_generation") != gen:
        return False
    if len(_adj_dict(G)) != int(prepared.get("node_count", -1)):
        return False
    hits = int(prepared.get("reuse_hits", 0)) + 1
    prepared["reuse_hits"] = hits
    if hits % _GRAPH_CTX_VALIDATE_EVERY != 0:
        return True
    return prepared.get("stamp") == _graph_stamp(G)


def _prepare_graph_cache(G):
    cache = None if _GRAPH_CONTEXT_CACHE_DISABLED else _graph_cache_dict(G)
    prepared = cache.get(_PREPARED_CACHE_KEY) if cache is not None else None
    if prepared is not None and _prepared_graph_usable(G, prepared):
        return prepared
    directed = bool(G.is_directed())
    if _is_bulk_graph(G):
        n = int(len(G))
        prepared = {
            "ctx_id": time.time_ns(),
            "stamp": (directed, n, int(getattr(G, "num_entries", 0)),
                      hash(getattr(G, "_eggpu_bulk_signature", None)) & _HASH_MASK),
            "bulk_csr": True,
            "bulk_signature": getattr(G, "_eggpu_bulk_signature", None),
            "node_count": n,
            "directed": directed,
            "nodes": G.nodes,
            "node_to_idx": None,
            "result_cache": {},
            "result_cache_order": [],
            "reuse_hits": 0,
            "mutation_generation": None,
        }
    else:
        nodes = list(G.nodes.keys()) if hasattr(G.nodes, "keys") else list(G.nodes)
        prepared = {
            "ctx_id": time.time_ns(),
            "stamp": _graph_stamp(G),
            "node_count": len(nodes),
            "directed": directed,
            "nodes": nodes,
            "node_to_idx": {node: i for i, node in enumerate(nodes)},
            "result_cache": {},
            "result_cache_order": [],
            "reuse_hits": 0,
            "mutation_generation": getattr(G, "_mutation_generation", None),
        }
    if cache is not None:
        cache[_PREPARED_CACHE_KEY] = prepared
    return prepared


def _graph_context(G):
    return _prepare_graph_cache(G)


# ---------------------------------------------------------------------------
# opt-in result cache (gpu_eggpu_backend.py:1723-1779)
# ---------------------------------------------------------------------------
def _cache_key(name, *parts):
    return tuple([name] + [round(float(p), 12) if isinstance(p, float) else p for p in parts])


def _allow_result_cache(prepared):
    if not _RESULT_CACHE_ENABLED:
        return False
    nodes = prepared.get("nodes")
    if nodes is None:
        return False
    return not (_RESULT_CACHE_MAX_NODES > 0 and len(nodes) > _RESULT_CACHE_MAX_NODES)


def _copy_for_return(obj):
    return copy.deepcopy(obj) if _RESULT_CACHE_RETURN_COPY else obj


def _result_cache_get(prepared, key):
    cache = prepared.get("result_cache")
    if not isinstance(cache, dict):
        return None
    entry = cache.get(key)
    if entry is None:
        return None
    return _copy_for_return(entry.get("result")), entry.get("kernel")


def _result_cache_put(prepared, key, result, kernel):
    if not _allow_result_cache(prepared):
        return
    cache = prepared.setdefault("result_cache", {})
    order = prepared.setdefault("result_cache_order", [])
    cache[key] = {"result": _copy_for_return(result), "kernel": kernel}
    if key in order:
        order.remove(key)
    order.append(key)
    while len(order) > _RESULT_CACHE_MAX_ITEMS:
        cache.pop(order.pop(0), None)


# ---------------------------------------------------------------------------
# graph -> out-CSR (Graph::gen_CSR order) -> DeviceGraph
# ---------------------------------------------------------------------------
def _identity_labels(prepared):
    """True when the node list is exactly the ints 0..n-1 in order (so a
    node is its own index); computed once per graph context."""
    flag = prepared.get("identity_labels")
    if flag is None:
        nodes = prepared["nodes"]
        flag = all(type(x) is int for x in nodes) and nodes == list(range(len(nodes)))
        prepared["identity_labels"] = flag
    return flag


def adjacency_csr(G, prepared):
    """(V, E) of the adjacency as Katz centrality reads it
    (``katz_centrality.py:73-81``): row u lists every adjacency key of u in
    adjacency order, self-loops included, both directions of an undirected
    edge; no sorting and no weights."""
    if _is_bulk_graph(G):
        V, E, _ = G.csr_arrays(None)
        return V, E
    idx = prepared["node_to_idx"]
    nodes = prepared["nodes"]
    adj = _adj_dict(G)
    identity = _identity_labels(prepared)
    get = idx.__getitem__
    counts, dst = [], []
    for u in nodes:
        nbrs = adj[u]
        counts.append(len(nbrs))
        dst.extend(nbrs if identity else map(get, nbrs))
    V = np.zeros(len(nodes) + 1, dtype=np.int64)
    if counts:
        V[1:] = np.cumsum(np.asarray(counts, dtype=np.int64))
    if V[-1] >= 2**31 - 1:
        raise ValueError("adjacency entry count exceeds the signed 32-bit EGGPU CSR limit")
    return V.astype(np.int32), np.fromiter(dst, dtype=np.int64, count=len(dst)).astype(np.int32)


def adjacency_weights(G, prepared, weight):
    """Weights aligned with :func:`adjacency_csr`: ``attrs.get(weight, 1.0)``
    for every adjacency entry, as eigenvector centrality reads them
    (``eigenvector.py:80-89``)."""
    if _is_bulk_graph(G):
        _, _, W = G.csr_arrays(weight)
        return np.asarray(W, dtype=np.float64)
    adj = _adj_dict(G)
    vals = []
    for u in prepared["nodes"]:
        vals.extend(float(d.get(weight, 1.0)) for d in adj[u].values())
    return np.asarray(vals, dtype=np.float64)


def graph_out_csr(G, prepared, weight=None):
    """Out-CSR of a protocol graph in the reference's order.

    Rows follow the node order, columns ascend within each row
    (``graph.cpp:824-886``, ``:719-822``).  Directed graphs keep self-loops
    (``graph_convert.cpp:110-134``); undirected graphs drop self-loops and
    store each pair in both directions (``graph_convert.cpp:72-108``).
    Missing weight attributes count as 1.0."""
    if _is_bulk_graph(G):
        return G.csr_arrays(weight)
    idx = prepared["node_to_idx"]
    n = len(prepared["nodes"])
    adj = _adj_dict(G)
    src, dst, wts = [], [], []
    weighted = weight is not None
    if prepared["directed"] and not weighted and _identity_labels(prepared) and list(adj) == prepared["nodes"]:
        # nodes 0..n-1 in adjacency order: rows and columns come straight from
        # the dicts at C speed (map / chain), no per-arc Python work
        counts = np.fromiter(map(len, adj.values()), dtype=np.int64, count=n)
        d = np.fromiter(itertools.chain.from_iterable(adj.values()), dtype=np.int64, count=int(counts.sum()))
        return _finish_out_csr(np.repeat(np.arange(n, dtype=np.int64), counts), d, None, n)
    if prepared["directed"]:
        # whole rows at a time: one extend per node instead of one append per arc;
        # nodes 0..n-1 in order are their own indices, so no per-arc lookup
        rows, counts = [], []
        get = idx.__getitem__
        identity = _identity_labels(prepared)
        for u, nbrs in adj.items():
            if not nbrs:
                continue
            rows.append(u if identity else idx[u])
            counts.append(len(nbrs))
            dst.extend(nbrs if identity else map(get, nbrs))
            if weighted:
                wts.extend(float(d.get(weight, 1.0)) for d in nbrs.values())
        src = np.repeat(np.asarray(rows, dtype=np.int64), np.asarray(counts, dtype=np.int64))
    else:
        seen = set()
        for u, nbrs in adj.items():
            iu = idx[u]
            for v, data in nbrs.items():
                iv = idx[v]
                if iu == iv:
                    continue
                key = (iu, iv) if iu < iv else (iv, iu)
                if key in seen:
                    continue
                seen.add(key)
                src.append(iu)
                dst.append(iv)
                src.append(iv)
                dst.append(iu)
                if weighted:
                    w = float(data.get(weight, 1.0))
                    wts.append(w)
                    wts.append(w)
    s = np.asarray(src, dtype=np.int64)
    d = np.fromiter(dst, dtype=np.int64, count=len(dst))
    W = np.fromiter(wts, dtype=np.float64, count=len(wts)) if weighted else None
    return _finish_out_csr(s, d, W, n)


def _finish_out_csr(s, d, W, n):
    """(V, E, W) from arc rows s and columns d: rows ascending, columns
    ascending within a row.  Usually already so (adjacency in node order,
    neighbours inserted in order); else one stable sort on the (row, column)
    key."""
    if len(d) > 1 and not bool(np.all((s[1:] > s[:-1]) | ((s[1:] == s[:-1]) & (d[1:] > d[:-1])))):
        order = np.argsort(s * max(n, 1) + d, kind="stable")
        s = s[order]
        d = d[order]
        if W is not None:
            W = W[order]
    E = d.astype(np.int32)
    V = np.zeros(n + 1, dtype=np.int64)
    if n:
        V[1:] = np.cumsum(np.bincount(s, minlength=n))
    if V[-1] >= 2**31 - 1:
        raise ValueError("adjacency entry count exceeds the signed 32-bit EGGPU CSR limit")
    return V.astype(np.int32), E, W


def _get_cached_cpp_graph(G, prepared=None, weight=None, alpha=0.85):
    """Device-resident graph for (G, weight), rebuilt when the context changes."""
    if prepared is None:
        prepared = _prepare_graph_cache(G)
    cache = None if _CPP_GRAPH_CACHE_DISABLED else _graph_cache_dict(G)
    key = f"{_DEVICE_GRAPH_KEY}:{'' if weight is None else str(weight)}"
    if cache is not None:
        entry = cache.get(key)
        if entry is not None and entry[0] == prepared.get("ctx_id"):
            return entry[1]
    if _is_bulk_graph(G) and hasattr(G, "device_graph"):
        dg = G.device_graph(weight, alpha=alpha)  # artifact files streamed straight into HBM
    else:
        V, E, W = graph_out_csr(G, prepared, weight)
        dg = _native().DeviceGraph.from_csr(V, E, W, alpha=alpha)
    if cache is not None:
        cache[key] = (prepared.get("ctx_id"), dg)
    return dg


def _cpp_gpu_pagerank(prepared, G, alpha, max_iter, eps, weight):
    """Native call + dense wrap (``gpu_eggpu_backend.py:636-704``)."""
    graph = _get_cached_cpp_graph(G, prepared=prepared, weight=weight, alpha=alpha)
    if graph is None:
        return None
    nodes = prepared.get("nodes", [])
    out = _native().cpp_gpu_pagerank_dense(graph, alpha=float(alpha), max_iterator=int(max_iter),
                                           threshold=float(eps), weight=weight)
    if not isinstance(out, dict) or out.get("values_dense") is None:
        return None
    vals = _normalize_dense_array(out["values_dense"], len(nodes), dtype="float64",
                                  allow_leading_sentinel=True, result_name="PageRank values_dense")
    values = _DenseValueDict(nodes, vals, node_to_idx=prepared.get("node_to_idx"), dtype=float)
    ks = out.get("kernel_seconds")
    it = out.get("iterations")
    if it is not None:
        _LAST_ITERATIONS["pagerank"] = int(it)
    return values, (float(ks) if ks is not None else None)


def pagerank(G, alpha=0.85, max_iter=200, eps=1e-6, weight=None):
    """GPU PageRank (``gpu_eggpu_backend.py:1782-1821``)."""
    prepared = _graph_context(G)
    if prepared.get("bulk_csr") and weight is not None and not _bulk_weight_available(G, weight):
        raise NotImplementedError(
            f"weighted PageRank requires an EGGPU CSR artifact containing the {weight!r} edge attribute"
        )
    key = _cache_key("pagerank", bool(prepared["directed"]), alpha, int(max_iter), eps, str(weight))
    cached = _result_cache_get(prepared, key)
    if cached is not None:
        result, ks = cached
        set_last_kernel_time("pagerank", ks)
        return result
    nodes = prepared["nodes"]
    if not len(nodes):
        set_last_kernel_time("pagerank", 0.0)
        return {}
    hit = _cpp_gpu_pagerank(prepared, G, alpha=alpha, max_iter=max_iter, eps=eps, weight=weight)
    if hit is None:
        raise RuntimeError("native HIP PageRank backend unavailable")
    result, ks = hit
    set_last_kernel_time("pagerank", ks)
    _result_cache_put(prepared, key, result, ks)
    return result
