"""PageRank slice of EasyGraph's GPU backend, over the MI355X library.

This module can stand where ``easygraph/utils/gpu_eggpu_backend.py`` stands
for PageRank: the same entry (``pagerank(G, alpha, max_iter, eps, weight)``,
``:1782-1821``), the same switches, the same prepared-context cache key and
dense-result contract (``:254-348,1092-1126``), which the reference's own
contract tests address by name.  Below that surface it is built differently:

* a graph's context is one dict, validated by the mutation generation, the
  node count and, every ``EASYGRAPH_GPU_CTX_VALIDATE_EVERY``-th reuse, a
  64-bit fingerprint of node keys, order, degrees and first neighbours,
  computed in C;
* the adjacency is read into int32/float64 CSR arrays by a small C
  extension (``csrc/egx_host.c``), not by per-arc Python code, and handed to
  ``libegx.so`` once; unweighted graphs skip the per-row column sort, which
  cannot change the pull CSC built from them (sources ascend in every
  destination row whatever the column order inside a source row), while
  weighted graphs keep ``Graph::gen_CSR`` order, since the out-weight sums
  follow it;
* the pull CSC, bins and workspace stay in HBM on a
  :class:`eggpu_amd.native.DeviceGraph` cached beside the context until the
  graph mutates;
* results are read-only mappings over the dense rank vector.
"""
from __future__ import annotations

import copy
import importlib
import itertools
import os
import weakref
from collections.abc import Mapping

import numpy as np

from .gpu_runtime import gpu_runtime_enabled

try:  # the C adjacency reader; absent (not built) -> the same arrays from Python
    from . import _egx_host as _host
except ImportError:  # pragma: no cover - exercised only without the extension
    _host = None


# ---------------------------------------------------------------------------
# switches: names, defaults and parsing are the reference's contract
# (gpu_eggpu_backend.py:12-74)
# ---------------------------------------------------------------------------
_TRUE_WORDS = frozenset({"1", "TRUE", "ON", "YES"})


def _flag(name, default=False):
    raw = os.environ.get(name)
    return bool(default) if raw is None else raw.strip().upper() in _TRUE_WORDS


def _count(name, default, floor):
    return max(floor, int(os.environ.get(name, str(default))))


_RESULT_CACHE_ENABLED = _flag("EASYGRAPH_GPU_RESULT_CACHE")
_RESULT_CACHE_RETURN_COPY = _flag("EASYGRAPH_GPU_RESULT_CACHE_RETURN_COPY", True)
_RESULT_CACHE_MAX_ITEMS = _count("EASYGRAPH_GPU_RESULT_CACHE_MAX_ITEMS", 8, 1)
_RESULT_CACHE_MAX_NODES = _count("EASYGRAPH_GPU_RESULT_CACHE_MAX_NODES", 2000000, 0)
_CONTEXT_CACHE_OFF = _flag("EASYGRAPH_GPU_DISABLE_GRAPH_CONTEXT_CACHE")
_DEVICE_GRAPH_CACHE_OFF = _flag("EASYGRAPH_GPU_DISABLE_CPP_GRAPH_CACHE")
_VALIDATE_EVERY = _count("EASYGRAPH_GPU_CTX_VALIDATE_EVERY", 32, 1)
# threads of the C walk of int-keyed adjacencies (one below 64K arcs)
_HOST_THREADS = _count("EASYGRAPH_GPU_HOST_THREADS", min(8, len(os.sched_getaffinity(0))), 1)
# the dict-returning native route: the PageRank switch, else the global one
# when that is set, else off
_PR_PREFER_DICT = _flag(
    "EASYGRAPH_GPU_PR_PREFER_CPP_DICT",
    _flag("EASYGRAPH_GPU_PREFER_CPP_DICT") if "EASYGRAPH_GPU_PREFER_CPP_DICT" in os.environ else False,
)

_CTX_KEY = "__eggpu_prepared_graph__"  # the reference's key in G.cache
_DEVICE_KEY = "__egx_device_graph__"
_ctx_ids = itertools.count(1)
_side_caches = weakref.WeakKeyDictionary()  # G.cache stand-in for graphs without one
_last_solve = {}  # name -> {"seconds": ..., "iterations": ...}


def eggpu_backend_enabled():
    return gpu_runtime_enabled()


def set_last_kernel_time(key, seconds):
    entry = _last_solve.setdefault(key, {})
    if seconds is None:
        entry.pop("seconds", None)
    else:
        entry["seconds"] = float(seconds)


def get_last_kernel_time(key):
    return _last_solve.get(key, {}).get("seconds")


def get_last_iterations(key):
    """Iterations of the last solve (new; the reference does not expose it)."""
    return _last_solve.get(key, {}).get("iterations")


def _native():
    # resolved through sys.modules on every call, so a test can put a fake
    # native module in place, as the reference's tests do for cpp_easygraph
    return importlib.import_module("eggpu_amd.native")


def _is_bulk_graph(G):
    return bool(getattr(G, "_eggpu_bulk_csr", False))


def _bulk_weight_available(G, weight):
    if not _is_bulk_graph(G) or weight is None:
        return False
    return getattr(G, "weights_path", None) is not None and getattr(G, "weight_key", None) == str(weight)


# ---------------------------------------------------------------------------
# dense results
# ---------------------------------------------------------------------------
class _DenseResultShapeError(ValueError):
    """A native dense result does not have its documented shape."""


def _normalize_dense_array(raw_values, target_len, dtype="float64", *, allow_leading_sentinel=False,
                           result_name="dense vector"):
    """The result as a 1-D array of exactly ``target_len`` entries; with
    ``allow_leading_sentinel`` also ``target_len + 1`` entries, whose first is
    dropped.  Any other shape raises :class:`_DenseResultShapeError`."""
    try:
        arr = np.asarray(raw_values, dtype=dtype)
    except (TypeError, ValueError) as exc:
        raise _DenseResultShapeError(f"EGGPU {result_name} is not a one-dimensional numeric array: {exc}") from exc
    n = int(target_len)
    accepted = ((n,), (n + 1,)) if allow_leading_sentinel else ((n,),)
    if arr.shape == accepted[0]:
        return arr
    if arr.shape in accepted:
        return arr[1:]
    raise _DenseResultShapeError(
        f"EGGPU {result_name} returned invalid dense shape {arr.shape}; expected "
        + " or ".join(str(s) for s in accepted)
    )


class _DenseValueDict(Mapping):
    """Read-only ``node -> value`` view of one dense result vector.

    It is a Mapping, not a dict: comparisons, ``dict(result)``, iteration and
    pickling behave as for the logical dict, writes raise ``TypeError``, and
    no per-node objects exist until a value is read.  A node is located
    through ``node_to_idx`` when given, arithmetically when the nodes are
    ``range(n)``, or through an index built on first use."""

    __slots__ = ("_keys", "_vals", "_index", "_arith")

    def __init__(self, nodes, values, node_to_idx=None, dtype=float):
        self._keys = nodes if isinstance(nodes, range) else list(nodes)
        vals = _normalize_dense_array(values, len(self._keys), dtype=np.dtype(dtype),
                                      result_name="dense node-value mapping")
        vals = vals.view()
        vals.flags.writeable = False
        self._vals = vals
        self._index = node_to_idx if isinstance(node_to_idx, dict) else None
        k = self._keys
        self._arith = self._index is None and isinstance(k, range) and k.start == 0 and k.step == 1

    def _slot(self, key):
        if self._index is not None:
            return self._index.get(key)
        if self._arith:
            if isinstance(key, bool) or not isinstance(key, (int, np.integer)):
                return None
            return int(key) if 0 <= key < len(self._keys) else None
        self._index = dict(zip(self._keys, range(len(self._keys))))
        return self._slot(key)

    def __getitem__(self, key):
        try:
            i = self._slot(key)
        except TypeError:  # unhashable key
            i = None
        if i is None:
            raise KeyError(key)
        return self._vals[i].item()

    def __contains__(self, key):
        try:
            return self._slot(key) is not None
        except TypeError:
            return False

    def __iter__(self):
        return iter(self._keys)

    def __len__(self):
        return len(self._keys)

    def values(self):
        """The dense vector itself (read-only), in node order."""
        return self._vals

    def to_numpy(self, copy=False):
        return self._vals.copy() if copy else self._vals

    def to_dict(self):
        return dict(zip(self._keys, self._vals.tolist()))

    copy = to_dict

    def tolist(self):
        return self._vals.tolist()

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self._vals, dtype=dtype)

    def __reduce__(self):
        return (type(self), (self._keys, np.array(self._vals), self._index, self._vals.dtype))

    def __repr__(self):
        if len(self._keys) > 1000:
            return f"{type(self).__name__}({len(self._keys)} nodes, {self._vals.dtype})"
        return repr(self.to_dict())


# ---------------------------------------------------------------------------
# the prepared graph context (the contract of gpu_eggpu_backend.py:1590-1716)
# ---------------------------------------------------------------------------
def _graph_cache_dict(G):
    cache = getattr(G, "cache", None)
    if isinstance(cache, dict):
        return cache
    try:
        return _side_caches.setdefault(G, {})
    except TypeError:  # neither a cache attribute nor weak-referenceable
        return None


def _adj_dict(G):
    return G._adj if hasattr(G, "_adj") else G.adj


def _node_list(G):
    nodes = G.nodes
    return list(nodes.keys()) if hasattr(nodes, "keys") else list(nodes)


def _fingerprint_py(adj):
    """Python form of _egx_host.stamp (used only without the extension)."""
    import hashlib

    h = hashlib.blake2b(digest_size=8)
    for i, (u, nbrs) in enumerate(adj.items()):
        first = next(iter(nbrs), None)
        h.update(np.array([hash(u), len(nbrs), i, hash(first)], dtype=np.int64).tobytes())
    return int.from_bytes(h.digest(), "little")


def _graph_stamp(G):
    """(directed, nodes, arc slots, fingerprint): what a context re-checks
    every _VALIDATE_EVERY-th reuse, to see changes made behind the mutation
    counter (direct writes into the adjacency dicts)."""
    adj = _adj_dict(G)
    if _host is not None and type(adj) is dict:
        fp, slots = _host.stamp(adj)
    else:
        fp, slots = _fingerprint_py(adj), sum(map(len, adj.values()))
    return (bool(G.is_directed()), len(adj), slots, fp)


def _identity_labels(prepared):
    """True when the nodes are exactly the ints 0..n-1 in order, so that a
    node is its own index (computed once per context)."""
    flag = prepared.get("identity_labels")
    if flag is None:
        nodes = prepared["nodes"]
        if _host is not None and type(nodes) is list:
            flag = bool(_host.identity(nodes))
        else:
            flag = set(map(type, nodes)) <= {int} and nodes == list(range(len(nodes)))
        prepared["identity_labels"] = flag
    return flag


def _context_current(G, ctx):
    if not isinstance(ctx, dict) or bool(ctx.get("directed")) != bool(G.is_directed()):
        return False
    if _is_bulk_graph(G):
        return ctx.get("bulk_signature") == getattr(G, "_eggpu_bulk_signature", None)
    if ctx.get("mutation_generation") != getattr(G, "_mutation_generation", None):
        return False
    if len(_adj_dict(G)) != ctx.get("node_count"):
        return False
    ctx["uses"] = ctx.get("uses", 0) + 1
    if ctx["uses"] % _VALIDATE_EVERY != 0:
        return True
    # a context whose adjacency was never read has no stamp: it is renewed
    return ctx.get("stamp") is not None and ctx["stamp"] == _graph_stamp(G)


def _new_context(G):
    directed = bool(G.is_directed())
    if _is_bulk_graph(G):
        return {"ctx_id": next(_ctx_ids), "directed": directed, "bulk_csr": True,
                "bulk_signature": getattr(G, "_eggpu_bulk_signature", None), "nodes": G.nodes,
                "node_to_idx": None, "node_count": int(len(G)), "mutation_generation": None, "stamp": None,
                "uses": 0, "result_cache": {}}
    nodes = _node_list(G)
    # the stamp is taken by the first read of the adjacency (in the same C
    # pass when it can be), so that it describes what the device graph holds
    ctx = {"ctx_id": next(_ctx_ids), "directed": directed, "nodes": nodes, "node_count": len(nodes),
           "mutation_generation": getattr(G, "_mutation_generation", None), "stamp": None,
           "uses": 0, "result_cache": {}}
    # nodes 0..n-1 in order are their own indices: no index dict is needed
    ctx["node_to_idx"] = None if _identity_labels(ctx) else dict(zip(nodes, range(len(nodes))))
    return ctx


def _prepare_graph_cache(G):
    """The graph's context: the cached one while still current, else a new
    one (stored in G.cache unless the context cache is switched off)."""
    cache = None if _CONTEXT_CACHE_OFF else _graph_cache_dict(G)
    if cache is not None:
        ctx = cache.get(_CTX_KEY)
        if ctx is not None and _context_current(G, ctx):
            return ctx
    ctx = _new_context(G)
    if cache is not None:
        cache[_CTX_KEY] = ctx
    return ctx


_graph_context = _prepare_graph_cache


# ---------------------------------------------------------------------------
# opt-in result cache (the contract of gpu_eggpu_backend.py:1723-1779): an
# insertion-ordered dict per context, least recently stored evicted first,
# holding snapshots so that callers cannot change what later calls receive
# ---------------------------------------------------------------------------
def _snapshot(obj):
    return copy.deepcopy(obj) if _RESULT_CACHE_RETURN_COPY else obj


def _result_cache_get(prepared, key):
    store = prepared.get("result_cache")
    hit = store.get(key) if isinstance(store, dict) else None
    if hit is None:
        return None
    result, kernel = hit
    return _snapshot(result), kernel


def _result_cache_put(prepared, key, result, kernel):
    if not _RESULT_CACHE_ENABLED or prepared.get("nodes") is None:
        return
    if 0 < _RESULT_CACHE_MAX_NODES < len(prepared["nodes"]):
        return
    store = prepared.setdefault("result_cache", {})
    store.pop(key, None)
    store[key] = (_snapshot(result), kernel)
    while len(store) > _RESULT_CACHE_MAX_ITEMS:
        del store[next(iter(store))]


# ---------------------------------------------------------------------------
# graph -> CSR arrays
# ---------------------------------------------------------------------------
def _walk_py(adj, index, n, drop_self, weight):
    """Python form of _egx_host.arcs (used only without the extension)."""
    V = np.zeros(n + 1, dtype=np.int32)
    cols, wts = [], []
    for row, nbrs in enumerate(adj.values()):
        for v, data in nbrs.items():
            col = v if index is None else index[v]
            if drop_self and col == row:
                continue
            cols.append(col)
            if weight is not None:
                wts.append(float(data.get(weight, 1.0)) if isinstance(data, dict) else 1.0)
        V[row + 1] = len(cols)
    E = np.asarray(cols, dtype=np.int32)
    W = np.asarray(wts, dtype=np.float64) if weight is not None else None
    return V, E, W


def _int_arcs(adj, prepared, index, n, drop_self):
    """(V, E) from the threaded C walk when the rows are keyed 0..n-1 and
    every neighbour is an int (no weights), recording the context's stamp
    from the same pass when it has none; None otherwise."""
    if _host is None or index is not None or type(adj) is not dict or not hasattr(_host, "int_arcs"):
        return None
    got = _host.int_arcs(adj, n, drop_self, _HOST_THREADS)
    if got is None:
        return None
    Vb, Eb, m, fp, slots = got
    if prepared.get("stamp") is None:
        prepared["stamp"] = (bool(prepared["directed"]), len(adj), slots, fp)
    return np.frombuffer(Vb, dtype=np.int32), np.frombuffer(Eb, dtype=np.int32)[:m]


def _out_arcs(G, prepared, weight, sort):
    """(V, E, W) of the out-arcs: rows in node order, columns in adjacency
    order or, with sort, ascending; directed graphs keep self-loops
    (graph_convert.cpp:110-134), undirected ones drop them and hold every
    pair in both directions (graph_convert.cpp:72-108)."""
    adj = _adj_dict(G)
    n = len(prepared["nodes"])
    index = None if _identity_labels(prepared) else prepared["node_to_idx"]
    drop_self = not prepared["directed"]
    got = _int_arcs(adj, prepared, index, n, drop_self) if weight is None else None
    if got is not None:
        V, E = got
        W = None
    elif _host is not None and type(adj) is dict:
        stamp = prepared.get("stamp")
        if stamp is None:
            stamp = prepared["stamp"] = _graph_stamp(G)
            cap = stamp[2]
        else:
            cap = sum(map(len, adj.values()))
        if cap >= 2**31 - 1:
            raise ValueError("adjacency entry count exceeds the signed 32-bit EGGPU CSR limit")
        V = np.empty(n + 1, dtype=np.int32)
        E = np.empty(max(cap, 1), dtype=np.int32)
        W = np.empty(max(cap, 1), dtype=np.float64) if weight is not None else None
        m = _host.arcs(adj, index, n, drop_self, V, E, weight, W)
        E = E[:m]
        W = W[:m] if W is not None else None
    else:
        V, E, W = _walk_py(adj, index, n, drop_self, weight)
    if sort and len(E) > 1:
        if _host is not None:
            _host.sort_rows(V, E, W)
        else:
            rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(V))
            order = np.lexsort((E, rows))
            E = E[order]
            W = W[order] if W is not None else None
    return V, E, W


def graph_out_csr(G, prepared, weight=None):
    """Out-CSR in ``Graph::gen_CSR`` order (``graph.cpp:824-886``,
    ``:719-822``): rows follow the node order, columns ascend within a row;
    missing weight attributes count as 1.0."""
    if _is_bulk_graph(G):
        return G.csr_arrays(weight)
    return _out_arcs(G, prepared, weight, sort=True)


def adjacency_csr(G, prepared):
    """(V, E) of the adjacency as Katz centrality reads it
    (``katz_centrality.py:73-81``): row u lists every adjacency key of u in
    adjacency order, self-loops included, both directions of an undirected
    edge; no sorting, no weights."""
    if _is_bulk_graph(G):
        V, E, _ = G.csr_arrays(None)
        return V, E
    adj = _adj_dict(G)
    n = len(prepared["nodes"])
    index = None if _identity_labels(prepared) else prepared["node_to_idx"]
    got = _int_arcs(adj, prepared, index, n, False)
    if got is not None:
        return got
    if _host is not None and type(adj) is dict:
        cap = sum(map(len, adj.values()))
        V = np.empty(n + 1, dtype=np.int32)
        E = np.empty(max(cap, 1), dtype=np.int32)
        m = _host.arcs(adj, index, n, False, V, E, None, None)
        return V, E[:m]
    V, E, _ = _walk_py(adj, index, n, False, None)
    return V, E


def adjacency_weights(G, prepared, weight):
    """Weights aligned with :func:`adjacency_csr`, ``attrs.get(weight, 1.0)``
    per adjacency entry, as eigenvector centrality reads them
    (``eigenvector.py:80-89``)."""
    if _is_bulk_graph(G):
        _, _, W = G.csr_arrays(weight)
        return np.asarray(W, dtype=np.float64)
    adj = _adj_dict(G)
    n = len(prepared["nodes"])
    index = None if _identity_labels(prepared) else prepared["node_to_idx"]
    if _host is not None and type(adj) is dict:
        cap = sum(map(len, adj.values()))
        V = np.empty(n + 1, dtype=np.int32)
        E = np.empty(max(cap, 1), dtype=np.int32)
        W = np.empty(max(cap, 1), dtype=np.float64)
        m = _host.arcs(adj, index, n, False, V, E, weight, W)
        return W[:m]
    return _walk_py(adj, index, n, False, weight)[2]


# ---------------------------------------------------------------------------
# device graph and the native call
# ---------------------------------------------------------------------------
def _get_cached_cpp_graph(G, prepared=None, weight=None, alpha=0.85):
    """The device-resident graph of (G, weight), built when the context is
    new and kept beside it in G.cache."""
    if prepared is None:
        prepared = _prepare_graph_cache(G)
    cache = None if _DEVICE_GRAPH_CACHE_OFF else _graph_cache_dict(G)
    slot = f"{_DEVICE_KEY}:{'' if weight is None else weight}"
    if cache is not None:
        held = cache.get(slot)
        if held is not None and held[0] == prepared.get("ctx_id"):
            return held[1]
    if _is_bulk_graph(G) and hasattr(G, "device_graph"):
        dg = G.device_graph(weight, alpha=alpha)  # artifact files streamed into HBM
    else:
        # weighted: gen_CSR column order, which the out-weight sums follow
        V, E, W = _out_arcs(G, prepared, weight, sort=weight is not None)
        dg = _native().DeviceGraph.from_csr(V, E, W, alpha=alpha)
    if cache is not None:
        cache[slot] = (prepared.get("ctx_id"), dg)
    return dg


def _cpp_gpu_pagerank(prepared, G, alpha, max_iter, eps, weight):
    """The native call and its result (``gpu_eggpu_backend.py:636-704``):
    ``(mapping, kernel_seconds)``, or None when the native side declines.
    With the dict switch on, the dict-returning native function is tried
    first; an empty answer for a non-empty graph without kernel time counts
    as declined and the dense function is used."""
    graph = _get_cached_cpp_graph(G, prepared=prepared, weight=weight, alpha=alpha)
    if graph is None:
        return None
    nat = _native()
    nodes = prepared.get("nodes", [])
    kw = dict(alpha=float(alpha), max_iterator=int(max_iter), threshold=float(eps), weight=weight)
    if _PR_PREFER_DICT and hasattr(nat, "cpp_gpu_pagerank"):
        out = nat.cpp_gpu_pagerank(graph, nodes=nodes, **kw)
        values = out.get("values") if isinstance(out, dict) else None
        ks = out.get("kernel_seconds") if isinstance(out, dict) else None
        declined = len(nodes) > 0 and isinstance(values, dict) and not values and not ks
        if isinstance(values, dict) and not declined:
            _record(out)
            return values, (float(ks) if ks is not None else None)
    out = nat.cpp_gpu_pagerank_dense(graph, **kw)
    if not isinstance(out, dict) or out.get("values_dense") is None:
        return None
    vals = _normalize_dense_array(out["values_dense"], len(nodes), dtype="float64", allow_leading_sentinel=True,
                                  result_name="PageRank values_dense")
    _record(out)
    ks = out.get("kernel_seconds")
    index = prepared.get("node_to_idx")
    keys = range(len(nodes)) if index is None and prepared.get("identity_labels") else nodes
    return (_DenseValueDict(keys, vals, node_to_idx=index, dtype=float), float(ks) if ks is not None else None)


def _record(out):
    if out.get("iterations") is not None:
        _last_solve.setdefault("pagerank", {})["iterations"] = int(out["iterations"])


def pagerank(G, alpha=0.85, max_iter=200, eps=1e-6, weight=None):
    """GPU PageRank (``gpu_eggpu_backend.py:1782-1821``): a read-only
    mapping node -> rank."""
    prepared = _graph_context(G)
    if prepared.get("bulk_csr") and weight is not None and not _bulk_weight_available(G, weight):
        raise NotImplementedError(
            f"weighted PageRank requires an EGGPU CSR artifact containing the {weight!r} edge attribute"
        )
    if not prepared["node_count"]:
        set_last_kernel_time("pagerank", 0.0)
        return {}
    key = ("pagerank", bool(prepared["directed"]), float(alpha), int(max_iter), float(eps), str(weight))
    hit = _result_cache_get(prepared, key)
    if hit is not None:
        set_last_kernel_time("pagerank", hit[1])
        return hit[0]
    solved = _cpp_gpu_pagerank(prepared, G, alpha=alpha, max_iter=max_iter, eps=eps, weight=weight)
    if solved is None:
        raise RuntimeError("native HIP PageRank backend unavailable")
    result, ks = solved
    set_last_kernel_time("pagerank", ks)
    _result_cache_put(prepared, key, result, ks)
    return result
