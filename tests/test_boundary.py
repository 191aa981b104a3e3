"""CPU tests of the drop-in boundary: the C ABI surface, the host-side graph
conversion and the backend/dispatch contracts of the reference
(easygraph/tests/test_eggpu_runtime.py, test_eggpu_cache_contracts.py,
functions/centrality/tests/test_pagerank.py).  No kernel runs here."""
import json
import os
import pickle
import re
import sys
from collections.abc import Mapping
from types import SimpleNamespace

import networkx as nx
import numpy as np
import pytest

import eggpu_amd
from eggpu_amd import gpu_backend, gpu_runtime, native
import importlib

pagerank_module = importlib.import_module("eggpu_amd.pagerank")
from eggpu_amd.bulk import read_eggpu_csr, write_eggpu_csr
from eggpu_amd.classes import DiGraph, Graph, MultiDiGraph
from eggpu_amd.exception import EasyGraphNotImplemented
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# -- C ABI -------------------------------------------------------------------
def _header_functions():
    text = open(os.path.join(ROOT, "include", "egx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(egx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = native.lib()
    names = _header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(native.SIGNATURES) == names


def test_error_strings_match_reference():
    # Easy-Graph/gpu_easygraph/common/err.cpp:9-31
    expected = {
        0: "EasyGraph GPU: success",
        1: "EasyGraph GPU: failed to allocate host mem",
        2: "EasyGraph GPU: failed to allocate gpu mem",
        3: "EasyGraph GPU: gpu error occurred",
        4: "EasyGraph GPU: gpu unkonw error",
        5: "EasyGraph GPU: unsupported graph type",
        6: "EasyGraph GPU: negative weight cycle detected",
        99: "EasyGraph GPU: not a valid err_code",
    }
    for code, text in expected.items():
        assert native.err_code_detail(code) == text


def test_device_count_does_not_fail():
    assert native.device_count() >= 0


def test_malformed_offsets_rejected_before_any_device_call():
    with pytest.raises(native.EgxError) as exc:
        native.DeviceGraph.from_csr(np.array([0, 5, 3, 6]), np.array([1, 2, 0, 1, 2, 0]))
    assert exc.value.status == native.EGX_UNSUPPORTED_GRAPH
    assert "monotonic" in str(exc.value)
    with pytest.raises(native.EgxError):
        native.DeviceGraph.from_csr(np.array([0, 2, 3]), np.array([1, 0]))  # V[n] != m


def test_empty_graph_pagerank_is_success_without_device():
    out, it, ks = native.pagerank_csr(np.zeros(1, np.int32), np.zeros(0, np.int32))
    assert out.shape == (0,) and it == 0 and ks == 0.0


# -- host graph conversion (graph_convert.cpp / Graph::gen_CSR) ---------------
def _prep(G):
    return gpu_backend._prepare_graph_cache(G)


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("weight", [None, "weight"])
def test_graph_out_csr_matches_gen_csr(directed, weight):
    rng = np.random.default_rng(5)
    G = DiGraph() if directed else Graph()
    G.add_nodes_from([10, 3, "x", (1, 2), 7])
    nodes = list(G.nodes)
    for _ in range(25):
        u, v = nodes[rng.integers(len(nodes))], nodes[rng.integers(len(nodes))]
        G.add_edge(u, v, weight=float(rng.uniform(0.5, 2)))
    V, E, W = gpu_backend.graph_out_csr(G, _prep(G), weight)
    Vo, Eo, Wo, _ = O.csr_from_graph(G, weight)
    assert np.array_equal(V, Vo) and np.array_equal(E, Eo)
    if weight is None:
        assert W is None
    else:
        assert np.array_equal(W, Wo)


def test_graph_out_csr_keeps_directed_self_loops_drops_undirected_ones():
    D = DiGraph()
    D.add_edges_from([(0, 0), (0, 1), (1, 2)])
    V, E, _ = gpu_backend.graph_out_csr(D, _prep(D))
    assert list(E) == [0, 1, 2]
    U = Graph()
    U.add_edges_from([(0, 0), (0, 1), (1, 2)])
    V, E, _ = gpu_backend.graph_out_csr(U, _prep(U))
    assert list(V) == [0, 1, 3, 4] and list(E) == [1, 0, 2, 1]


@pytest.mark.parametrize("labels", ["identity", "shuffled", "floaty"])
@pytest.mark.parametrize("weight", [None, "weight"])
def test_graph_out_csr_integer_labels_and_unsorted_rows(labels, weight):
    # nodes 0..n-1 in order take the no-lookup path; neighbours are inserted
    # in descending order so every row needs sorting
    rng = np.random.default_rng(11)
    n = 40
    nodes = list(range(n))
    if labels == "shuffled":
        rng.shuffle(nodes)
    elif labels == "floaty":
        nodes[3] = 3.5  # not an int: the lookup path
    G = DiGraph()
    G.add_nodes_from(nodes)
    for u in nodes:
        for v in sorted(rng.choice(n, size=int(rng.integers(0, 6)), replace=False), reverse=True):
            G.add_edge(u, nodes[int(v)], weight=float(rng.uniform(0.5, 2)))
    prep = _prep(G)
    assert gpu_backend._identity_labels(prep) == (labels == "identity")
    V, E, W = gpu_backend.graph_out_csr(G, prep, weight)
    Vo, Eo, Wo, _ = O.csr_from_graph(G, weight)
    assert np.array_equal(V, Vo) and np.array_equal(E, Eo)
    if weight is not None:
        assert np.array_equal(W, Wo)


def test_graph_stamp_sees_changes_behind_the_generation_counter():
    G = DiGraph()
    G.add_edges_from([(0, 1), (1, 2), (2, 0), (2, 3)])
    s0 = gpu_backend._graph_stamp(G)
    H = DiGraph()
    H.add_edges_from([(0, 1), (1, 2), (2, 0), (2, 3)])
    assert gpu_backend._graph_stamp(H) == s0  # same graph, same stamp
    G._adj[1][3] = {}  # an arc added without touching the generation counter
    assert gpu_backend._graph_stamp(G) != s0
    K = DiGraph()
    K.add_edges_from([(0, 1), (1, 2), (2, 0), (2, 4)])  # same counts, different labels
    assert gpu_backend._graph_stamp(K) != s0
    P = DiGraph()
    P.add_nodes_from([1, 0, 2, 3])  # same nodes and degrees, another order
    P.add_edges_from([(0, 1), (1, 2), (2, 0), (2, 3)])
    assert gpu_backend._graph_stamp(P) != s0


def test_networkx_graph_is_accepted():
    G = nx.DiGraph()
    G.add_edges_from([(0, 1), (1, 2), (2, 0), (2, 3)])
    V, E, _ = gpu_backend.graph_out_csr(G, _prep(G))
    assert list(V) == [0, 1, 2, 4, 4] and list(E) == [1, 2, 0, 3]


# -- dense result contract (test_eggpu_cache_contracts.py:76-187) -------------
def test_dense_result_is_an_honest_read_only_mapping():
    values = gpu_backend._DenseValueDict(range(3), [0.25, 0.5, 0.25])
    expected = {0: 0.25, 1: 0.5, 2: 0.25}
    assert isinstance(values, Mapping)
    assert not isinstance(values, dict)
    assert values == expected
    assert dict(values) == expected
    assert pickle.loads(pickle.dumps(values)) == expected
    assert values.to_numpy(copy=False).flags.writeable is False
    with pytest.raises(TypeError):
        values[0] = 1.0
    with pytest.raises(TypeError):
        json.dumps(values)
    assert 3 not in values and values.get(3, "d") == "d"
    with pytest.raises(KeyError):
        values[5]


def test_dense_result_with_arbitrary_nodes():
    nodes = ["a", (1, 2), None]
    values = gpu_backend._DenseValueDict(nodes, [0.1, 0.2, 0.7], node_to_idx={"a": 0, (1, 2): 1, None: 2})
    assert values[(1, 2)] == 0.2 and list(values) == nodes


def test_dense_shape_contract_accepts_only_exact_or_explicit_sentinel():
    exact = gpu_backend._normalize_dense_array([1.0, 2.0, 3.0], 3, result_name="test vector")
    sentinel = gpu_backend._normalize_dense_array([-1.0, 1.0, 2.0, 3.0], 3, allow_leading_sentinel=True,
                                                  result_name="test vector")
    np.testing.assert_array_equal(exact, [1.0, 2.0, 3.0])
    np.testing.assert_array_equal(sentinel, [1.0, 2.0, 3.0])
    with pytest.raises(gpu_backend._DenseResultShapeError, match="expected \\(3,\\)"):
        gpu_backend._normalize_dense_array([-1.0, 1.0, 2.0, 3.0], 3, result_name="test vector")
    with pytest.raises(gpu_backend._DenseResultShapeError, match="dense node-value mapping.*expected \\(3,\\)"):
        gpu_backend._DenseValueDict(range(3), [1.0, 2.0])


def test_pagerank_wrong_dense_size_fails_closed(monkeypatch):
    """A native module returning the wrong length is refused (test_eggpu_cache_contracts.py:199-220)."""
    monkeypatch.setattr(gpu_backend, "_get_cached_cpp_graph", lambda *_a, **_k: object())
    fake = SimpleNamespace(cpp_gpu_pagerank_dense=lambda _g, **_k: {"values_dense": np.asarray([0.5, 0.5]),
                                                                      "kernel_seconds": 0.001})
    monkeypatch.setitem(sys.modules, "eggpu_amd.native", fake)
    prepared = {"nodes": [0, 1, 2], "node_to_idx": {0: 0, 1: 1, 2: 2}}
    with pytest.raises(gpu_backend._DenseResultShapeError, match="PageRank values_dense.*expected \\(3,\\) or \\(4,\\)"):
        gpu_backend._cpp_gpu_pagerank(prepared, object(), alpha=0.85, max_iter=100, eps=1e-6, weight=None)


def test_opt_in_result_cache_owns_independent_snapshots(monkeypatch):
    monkeypatch.setattr(gpu_backend, "_RESULT_CACHE_ENABLED", True)
    monkeypatch.setattr(gpu_backend, "_RESULT_CACHE_RETURN_COPY", True)
    prepared = {"nodes": [0], "result_cache": {}, "result_cache_order": []}
    original = {0: [1]}
    gpu_backend._result_cache_put(prepared, ("test",), original, 0.25)
    original[0].append(2)
    first, kernel = gpu_backend._result_cache_get(prepared, ("test",))
    first[0].append(3)
    second, _ = gpu_backend._result_cache_get(prepared, ("test",))
    assert kernel == 0.25 and first == {0: [1, 3]} and second == {0: [1]}


# -- runtime switches (test_eggpu_runtime.py) ---------------------------------
@pytest.mark.parametrize("value,expected", [("TRUE", True), ("1", True), (" on ", True), ("yes", True),
                                            ("FALSE", False), ("", False), ("2", False)])
def test_gpu_switch_parsing(monkeypatch, value, expected):
    monkeypatch.setenv("EASYGRAPH_ENABLE_GPU", value)
    assert gpu_runtime.gpu_runtime_enabled() is expected


def test_switches_default_off_and_independent(monkeypatch):
    monkeypatch.delenv("EASYGRAPH_ENABLE_GPU", raising=False)
    monkeypatch.delenv("EASYGRAPH_GPU_STRICT_ERRORS", raising=False)
    assert not gpu_runtime.gpu_runtime_enabled() and not gpu_runtime.gpu_strict_errors()
    monkeypatch.setenv("EASYGRAPH_GPU_STRICT_ERRORS", "TRUE")
    assert gpu_runtime.gpu_strict_errors() and not gpu_runtime.gpu_runtime_enabled()


# -- dispatch (pagerank.py:12-72,161-179) --------------------------------------
def _small_graph():
    G = DiGraph()
    G.add_edges_from([(0, 1), (1, 2), (2, 0)])
    return G


def test_multigraph_raises():
    G = MultiDiGraph()
    G.add_edges_from([(0, 1), (0, 1)])
    with pytest.raises(EasyGraphNotImplemented):
        eggpu_amd.pagerank(G)


def test_gpu_off_without_cpu_route_raises(monkeypatch):
    monkeypatch.delenv("EASYGRAPH_ENABLE_GPU", raising=False)
    monkeypatch.setattr(pagerank_module, "_cpu_route", lambda: None)
    with pytest.raises(RuntimeError, match="GPU route"):
        eggpu_amd.pagerank(_small_graph())
    assert eggpu_amd.pagerank(DiGraph()) == {}


def test_gpu_off_uses_registered_cpu_route(monkeypatch):
    monkeypatch.delenv("EASYGRAPH_ENABLE_GPU", raising=False)
    seen = {}

    def cpu(G, alpha, weight, max_iter, tol):
        seen.update(alpha=alpha, max_iter=max_iter, tol=tol)
        return {"cpu": True}

    monkeypatch.setattr(pagerank_module, "_CPU_PAGERANK", cpu)
    monkeypatch.setenv("EASYGRAPH_CPU_PR_MAX_ITER", "0")  # clamped to 1 (pagerank.py:46)
    monkeypatch.setenv("EASYGRAPH_CPU_PR_TOL", "-3")      # clamped to 0 (pagerank.py:50)
    assert eggpu_amd.pagerank(_small_graph(), alpha=0.7) == {"cpu": True}
    assert seen == {"alpha": 0.7, "max_iter": 1, "tol": 0.0}


def test_strict_mode_propagates_backend_errors(monkeypatch):
    monkeypatch.setenv("EASYGRAPH_ENABLE_GPU", "TRUE")
    monkeypatch.setenv("EASYGRAPH_GPU_STRICT_ERRORS", "TRUE")

    def boom(*a, **k):
        raise RuntimeError("device failure")

    monkeypatch.setattr(gpu_backend, "pagerank", boom)
    monkeypatch.setattr(pagerank_module, "_CPU_PAGERANK", lambda *a, **k: {"cpu": True})
    with pytest.raises(RuntimeError, match="device failure"):
        eggpu_amd.pagerank(_small_graph())
    monkeypatch.delenv("EASYGRAPH_GPU_STRICT_ERRORS")
    assert eggpu_amd.pagerank(_small_graph()) == {"cpu": True}


def test_missing_native_library_always_raises(monkeypatch):
    monkeypatch.setenv("EASYGRAPH_ENABLE_GPU", "TRUE")
    monkeypatch.delenv("EASYGRAPH_GPU_STRICT_ERRORS", raising=False)

    def missing(*a, **k):
        raise native.NativeLibraryMissing("libegx.so is not built")

    monkeypatch.setattr(gpu_backend, "pagerank", missing)
    monkeypatch.setattr(pagerank_module, "_CPU_PAGERANK", lambda *a, **k: {"cpu": True})
    with pytest.raises(native.NativeLibraryMissing):
        eggpu_amd.pagerank(_small_graph())


def test_backend_empty_graph_returns_empty_without_native_call(monkeypatch):
    monkeypatch.setattr(gpu_backend, "_cpp_gpu_pagerank", lambda *a, **k: pytest.fail("native call"))
    assert gpu_backend.pagerank(DiGraph()) == {}


# -- mutation and context validity (graph.py:17-78,1641-1644) ------------------
def test_edge_attribute_mutation_invalidates_graph_state():
    G = Graph()
    G.add_edge("u", "v", weight=1.0)
    G.cache["__eggpu_prepared_graph__"] = {"ctx_id": 1}
    gen = G._mutation_generation
    G["u"]["v"]["weight"] = 7.0
    assert G.cache == {} and G._mutation_generation == gen + 1 and G["u"]["v"]["weight"] == 7.0


def test_graph_context_binds_mutation_generation():
    G = Graph()
    G.add_edge(0, 1, weight=1.0)
    first = gpu_backend._graph_context(G)
    assert gpu_backend._graph_context(G) is first
    G[0][1]["weight"] = 2.0
    second = gpu_backend._graph_context(G)
    assert second is not first and second["mutation_generation"] == G._mutation_generation


def test_remove_nodes_invalidates():
    G = Graph()
    G.add_edges_from([(0, 1), (1, 2)])
    G.cache["x"] = 1
    G.remove_nodes_from([1])
    assert G.cache == {} and 1 not in G


# -- bulk manifests (eggpu_bulk_graph.py:170-334, graph_convert.cpp:447-526) ---
EDGES = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3)]


def _tiny_csr(directed):
    rows = [[] for _ in range(6)]
    for u, v in EDGES:
        rows[u].append(v)
        if not directed:
            rows[v].append(u)
    V = np.zeros(7, np.int32)
    V[1:] = np.cumsum([len(r) for r in rows])
    E = np.array([v for r in rows for v in r], np.int32)
    W = np.array([1.0 + (s * t) % 6 for s, r in enumerate(rows) for t in r])
    return V, E, W


def test_bulk_round_trip(tmp_path):
    V, E, W = _tiny_csr(directed=False)
    mp = write_eggpu_csr(tmp_path / "tiny", V, E, W, directed=False)
    G = read_eggpu_csr(mp)
    assert len(G) == 6 and list(G.nodes) == list(range(6)) and not G.is_directed()
    V2, E2, W2 = G.csr_arrays("weight")
    assert np.array_equal(V2, V) and np.array_equal(E2, E) and np.array_equal(W2, W)
    with pytest.raises(NotImplementedError, match="weighted PageRank requires an EGGPU CSR artifact"):
        G.csr_arrays("cost")


def test_bulk_weighted_on_topology_only_raises(tmp_path, monkeypatch):
    V, E, _ = _tiny_csr(directed=False)
    G = read_eggpu_csr(write_eggpu_csr(tmp_path / "tiny", V, E))
    monkeypatch.setenv("EASYGRAPH_ENABLE_GPU", "TRUE")
    monkeypatch.delenv("EASYGRAPH_GPU_STRICT_ERRORS", raising=False)
    with pytest.raises(NotImplementedError, match="weighted PageRank requires an EGGPU CSR artifact"):
        eggpu_amd.pagerank(G, weight="weight")


def test_bulk_validation(tmp_path):
    V, E, W = _tiny_csr(directed=True)
    mp = write_eggpu_csr(tmp_path / "g", V, E, W)
    meta = json.loads(mp.read_text())
    bad = dict(meta, format="other")
    (tmp_path / "bad.json").write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="unsupported EGGPU CSR format"):
        read_eggpu_csr(tmp_path / "bad.json")
    E_bad = E.copy()
    E_bad[0] = 9
    with pytest.raises(ValueError, match="outside"):
        read_eggpu_csr(write_eggpu_csr(tmp_path / "h", V, E_bad))
    V_bad = V.copy()
    V_bad[2], V_bad[3] = V_bad[3], V_bad[2]
    with pytest.raises(ValueError, match="monotonic"):
        read_eggpu_csr(write_eggpu_csr(tmp_path / "i", V_bad, E))
    W_bad = W.copy()
    W_bad[1] = np.nan
    with pytest.raises(ValueError, match="finite"):
        read_eggpu_csr(write_eggpu_csr(tmp_path / "j", V, E, W_bad))
    short = dict(meta, num_entries=len(E) - 1)
    (tmp_path / "short.json").write_text(json.dumps(short))
    with pytest.raises(ValueError, match="size mismatch"):
        read_eggpu_csr(tmp_path / "short.json")
    G = read_eggpu_csr(mp)
    assert G.simple_sorted_topology is True
