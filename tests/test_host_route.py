"""CPU tests of the host side of the dict-graph route: the C adjacency
reader (eggpu_amd/csrc/egx_host.c) against its Python form and against the
oracle's gen_CSR restatement, the C++-graph route of the reference's
@hybrid("cpp_pagerank") (decorators.py:140-163) and the dict-returning
native route switch (gpu_eggpu_backend.py:60-74,636-704)."""
import importlib
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import eggpu_amd
from eggpu_amd import gpu_backend
from eggpu_amd.classes import DiGraph, Graph, MultiDiGraph
from eggpu_amd.exception import EasyGraphNotImplemented
from oracle import oracle as O

pagerank_module = importlib.import_module("eggpu_amd.pagerank")

host = pytest.importorskip("eggpu_amd._egx_host")


def _random_graph(directed, labels, n=60, seed=3):
    rng = np.random.default_rng(seed)
    nodes = list(range(n))
    if labels == "shuffled":
        rng.shuffle(nodes)
    elif labels == "objects":
        nodes = [f"v{i}" if i % 3 else (i, "t") for i in range(n)]
    G = DiGraph() if directed else Graph()
    G.add_nodes_from(nodes)
    for _ in range(5 * n):
        u, v = nodes[int(rng.integers(n))], nodes[int(rng.integers(n))]  # self-loops included
        G.add_edge(u, v, weight=float(rng.uniform(0.1, 2.0)))
    return G


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("labels", ["identity", "shuffled", "objects"])
@pytest.mark.parametrize("weight", [None, "weight"])
def test_c_reader_matches_python_form_and_gen_csr(directed, labels, weight):
    G = _random_graph(directed, labels)
    prep = gpu_backend._prepare_graph_cache(G)
    assert gpu_backend._identity_labels(prep) == (labels == "identity")
    # adjacency order: the C reader equals the Python form
    V, E, W = gpu_backend._out_arcs(G, prep, weight, sort=False)
    index = None if gpu_backend._identity_labels(prep) else prep["node_to_idx"]
    Vp, Ep, Wp = gpu_backend._walk_py(G._adj, index, len(prep["nodes"]), not directed, weight)
    assert np.array_equal(V, Vp) and np.array_equal(E, Ep)
    assert (W is None and Wp is None) or np.array_equal(W, Wp)
    # sorted: Graph::gen_CSR order, as the oracle restates it
    V, E, W = gpu_backend.graph_out_csr(G, prep, weight)
    Vo, Eo, Wo, _ = O.csr_from_graph(G, weight)
    assert np.array_equal(V, Vo) and np.array_equal(E, Eo)
    if weight is not None:
        assert np.array_equal(W, Wo)


def test_sort_rows_long_rows_stable_with_weights():
    rng = np.random.default_rng(9)
    lens = [0, 1, 2, 33, 500, 31, 64]
    V = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    E = rng.integers(0, 40, V[-1]).astype(np.int32)  # repeated columns: stability matters
    W = np.arange(V[-1], dtype=np.float64)
    E0, W0 = E.copy(), W.copy()
    host.sort_rows(V, E, W)
    for r in range(len(lens)):
        b, e = V[r], V[r + 1]
        order = np.argsort(E0[b:e], kind="stable")
        assert np.array_equal(E[b:e], E0[b:e][order]) and np.array_equal(W[b:e], W0[b:e][order])


def test_c_reader_rejects_unknown_neighbours_and_short_buffers():
    G = DiGraph()
    G.add_edges_from([(0, 1), (1, 2)])
    G._adj[2]["ghost"] = {}
    prep = gpu_backend._prepare_graph_cache(G)
    with pytest.raises((KeyError, TypeError)):
        gpu_backend._out_arcs(G, prep, None, sort=False)
    V = np.empty(4, np.int32)
    with pytest.raises(ValueError):
        host.arcs({0: {1: {}}, 1: {0: {}}, 2: {}}, None, 3, False, V, np.empty(1, np.int32), None, None)


def test_stamp_sees_first_neighbour_rewiring():
    G = DiGraph()
    G.add_edges_from([(0, 1), (1, 2), (2, 0)])
    s0 = gpu_backend._graph_stamp(G)
    del G._adj[0][1]
    G._adj[0][2] = {}  # same degrees, node keys and order; another neighbour
    assert gpu_backend._graph_stamp(G) != s0


def test_cpp_graphs_take_the_cpp_route(monkeypatch):
    seen = {}

    def cpp_pagerank(G, *args, **kwargs):
        seen["args"], seen["kwargs"] = args, kwargs
        return {"cpp": True}

    monkeypatch.setattr(pagerank_module, "_CPP_PAGERANK", cpp_pagerank)
    G = DiGraph()
    G.add_edge(0, 1)
    G.cflag = 1
    assert eggpu_amd.pagerank(G, 0.9, max_iter=7, tol=1e-3) == {"cpp": True}
    assert seen == {"args": (0.9,), "kwargs": {"max_iterator": 7, "threshold": 1e-3}}
    # the caller's own names win; nothing is added that the caller did not pass
    eggpu_amd.pagerank(G, max_iterator=3)
    assert seen["kwargs"] == {"max_iterator": 3}
    # multigraphs are refused before any route
    M = MultiDiGraph()
    M.add_edge(0, 1)
    M.cflag = 1
    with pytest.raises(EasyGraphNotImplemented):
        eggpu_amd.pagerank(M)
    # python graphs never take it
    monkeypatch.setattr(pagerank_module, "_CPU_PAGERANK", lambda *a, **k: {"cpu": True})
    monkeypatch.delenv("EASYGRAPH_ENABLE_GPU", raising=False)
    H = DiGraph()
    H.add_edge(0, 1)
    assert eggpu_amd.pagerank(H) == {"cpu": True}


def test_dict_route_switch(monkeypatch):
    calls = []

    def dict_route(graph, nodes=None, **kw):
        calls.append("dict")
        return {"values": dict(zip(nodes, [0.25, 0.75])), "kernel_seconds": 0.5}

    def dense_route(graph, **kw):
        calls.append("dense")
        return {"values_dense": np.array([0.5, 0.5]), "kernel_seconds": 0.25}

    fake = SimpleNamespace(cpp_gpu_pagerank=dict_route, cpp_gpu_pagerank_dense=dense_route)
    monkeypatch.setitem(sys.modules, "eggpu_amd.native", fake)
    monkeypatch.setattr(gpu_backend, "_get_cached_cpp_graph", lambda *a, **k: object())
    prepared = {"nodes": ["a", "b"], "node_to_idx": {"a": 0, "b": 1}}
    monkeypatch.setattr(gpu_backend, "_PR_PREFER_DICT", False)
    vals, ks = gpu_backend._cpp_gpu_pagerank(prepared, object(), 0.85, 10, 1e-6, None)
    assert calls == ["dense"] and dict(vals) == {"a": 0.5, "b": 0.5} and ks == 0.25
    monkeypatch.setattr(gpu_backend, "_PR_PREFER_DICT", True)
    vals, ks = gpu_backend._cpp_gpu_pagerank(prepared, object(), 0.85, 10, 1e-6, None)
    assert calls[-1] == "dict" and vals == {"a": 0.25, "b": 0.75} and ks == 0.5
    # an empty answer without kernel time for a non-empty graph: the dense route
    fake.cpp_gpu_pagerank = lambda graph, nodes=None, **kw: {"values": {}, "kernel_seconds": 0.0}
    vals, _ = gpu_backend._cpp_gpu_pagerank(prepared, object(), 0.85, 10, 1e-6, None)
    assert calls[-1] == "dense" and dict(vals) == {"a": 0.5, "b": 0.5}
