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


def _int_adjacency(n, arcs_per_row, directed, seed=5):
    rng = np.random.default_rng(seed)
    G = DiGraph() if directed else Graph()
    G.add_nodes_from(range(n))
    deg = rng.integers(0, 2 * arcs_per_row, n)
    deg[::97] = 40 * arcs_per_row  # a few long rows to balance across threads
    src = np.repeat(np.arange(n), deg)
    dst = rng.integers(0, n, len(src))
    dst[::50] = src[::50]  # self-loops, which undirected graphs drop
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    return G


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_int_reader_matches_arcs_and_stamp(directed, threads):
    G = _int_adjacency(3000, 24, directed)
    adj, n = G._adj, len(G._adj)
    Vb, Eb, m, fp, slots = host.int_arcs(adj, n, not directed, threads)
    cap = sum(map(len, adj.values()))
    V, E = np.empty(n + 1, np.int32), np.empty(cap, np.int32)
    m0 = host.arcs(adj, None, n, not directed, V, E, None, None)
    assert cap >= 65536 and m == m0
    assert np.array_equal(np.frombuffer(Vb, np.int32), V)
    assert np.array_equal(np.frombuffer(Eb, np.int32)[:m], E[:m0])
    assert (fp, slots) == host.stamp(adj)
    # the route through the backend records that stamp on the new context
    prep = gpu_backend._prepare_graph_cache(G)
    assert prep["stamp"] is None
    V2, E2, _ = gpu_backend._out_arcs(G, prep, None, sort=False)
    assert np.array_equal(V2, V) and np.array_equal(E2, E[:m0])
    assert prep["stamp"] == gpu_backend._graph_stamp(G)


def _walk_in_child(conn):
    G = _int_adjacency(3000, 24, True)
    got = host.int_arcs(G._adj, len(G._adj), False, 8)
    conn.send(got is not None and got[2] == sum(map(len, G._adj.values())))


def test_threaded_int_reader_pool_survives_fork():
    import multiprocessing as mp

    G = _int_adjacency(3000, 24, True)
    assert host.int_arcs(G._adj, len(G._adj), False, 8) is not None  # the pool exists here
    ctx = mp.get_context("fork")
    a, b = ctx.Pipe()
    p = ctx.Process(target=_walk_in_child, args=(b,))
    p.start()
    p.join(60)
    alive = p.is_alive()
    if alive:
        p.kill()
    assert not alive and p.exitcode == 0 and a.poll(1) and a.recv()


def test_threaded_int_reader_declines_other_keys():
    G = _int_adjacency(200, 4, True)
    adj, n = G._adj, len(G._adj)
    assert host.int_arcs(adj, n, False, 4) is not None
    assert host.int_arcs(adj, n + 1, False, 4) is None  # row count differs
    adj[7]["x"] = {}  # a neighbour that is not an int
    assert host.int_arcs(adj, n, False, 4) is None
    del adj[7]["x"]
    adj[7][n + 5] = {}  # an int outside the node range
    assert host.int_arcs(adj, n, False, 4) is None
    del adj[7][n + 5]
    H = DiGraph()
    H.add_nodes_from([1, 0, 2])  # keys not in index order
    H.add_edges_from([(0, 1), (1, 2)])
    assert host.int_arcs(H._adj, 3, False, 2) is None


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
