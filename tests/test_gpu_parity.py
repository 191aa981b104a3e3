"""GPU parity tests: the HIP path (through libegx.so's C ABI) against the
pinned CPU oracle and the golden fixtures.

Bars: integer graph construction bit-exact; PageRank within
max-abs 1e-6 per node and L1 1e-5 of the FP64 oracle with the identical
iteration count; FP64 iterates within 1e-12.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import eggpu_amd
from eggpu_amd import gpu_backend, native
from eggpu_amd.bulk import read_eggpu_csr, write_eggpu_csr
from eggpu_amd.classes import DiGraph, Graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MAX_ABS = 1e-6   # per node, FP32 iterates with FP64 reductions
MAX_L1 = 1e-5
FP64_TOL = 1e-12


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _assert_close(ours, ref, tag, tol=MAX_ABS, l1=MAX_L1):
    err = np.abs(np.asarray(ours) - np.asarray(ref))
    _say(f"{tag}: max-abs {err.max():.3e}, L1 {err.sum():.3e}")
    assert err.max() <= tol, tag
    assert err.sum() <= l1, tag


# -- generators: bit-exact against the oracle's C definition ------------------
@pytest.mark.parametrize("scale,weighted,seed", [(8, False, 1), (10, True, 3), (12, True, 7), (14, False, 1)])
def test_rmat_generator_bit_exact(gpu_ready, scale, weighted, seed):
    _say(f"rmat generator scale={scale} weighted={weighted} seed={seed}")
    g = native.DeviceGraph.rmat(scale, 16, seed=seed, weighted=weighted, wseed=seed + 100, keep_csr=True)
    V, E, W = g.download_csr()
    Vo, Eo = O.rmat_csr(scale, 16, seed=seed)
    assert np.array_equal(V, Vo) and np.array_equal(E, Eo)
    if weighted:
        assert np.array_equal(W, O.edge_weights(Vo, Eo, seed + 100))


def test_er_generator_bit_exact(gpu_ready):
    _say("er generator n=5000 m=40000")
    g = native.DeviceGraph.er(5000, 40000, seed=17, keep_csr=True)
    V, E, _ = g.download_csr()
    Vo, Eo = O.er_csr(5000, 40000, 17)
    assert np.array_equal(V, Vo) and np.array_equal(E, Eo)


# -- pull-CSC build: bit-exact against build_inbound_from_csr ------------------
def _random_csr(rng, n, m, invalid=False, bad=False):
    src = np.sort(rng.integers(0, n, m))
    dst = rng.integers(0, n, m)
    if invalid:
        idx = rng.choice(m, m // 10, replace=False)
        dst[idx] = rng.choice([-1, -7, n, n + 5], len(idx))
    V = np.zeros(n + 1, np.int32)
    V[1:] = np.cumsum(np.bincount(src, minlength=n))
    W = rng.uniform(0.1, 2.0, m)
    if bad:
        idx = rng.choice(m, m // 8, replace=False)
        W[idx] = rng.choice([-1.0, np.nan, np.inf, -np.inf, 0.0], len(idx))
    return V, dst.astype(np.int32), W


CSC_CASES = [
    ("tiny", lambda r: (np.array([0, 2, 3, 4, 4], np.int32), np.array([1, 2, 2, 0], np.int32), None)),
    ("random_unweighted_dups", lambda r: _random_csr(r, 300, 4000)[:2] + (None,)),
    ("random_weighted", lambda r: _random_csr(r, 300, 4000)),
    ("invalid_endpoints", lambda r: _random_csr(r, 500, 6000, invalid=True)),
    ("bad_weights", lambda r: _random_csr(r, 500, 6000, bad=True)),
    ("invalid_and_bad", lambda r: _random_csr(r, 800, 9000, invalid=True, bad=True)),
    ("invalid_unweighted", lambda r: _random_csr(r, 800, 9000, invalid=True)[:2] + (None,)),
    ("all_zero_weights", lambda r: (np.array([0, 2, 3, 3], np.int32), np.array([1, 2, 0], np.int32),
                                    np.zeros(3))),
    ("single_self_loop", lambda r: (np.array([0, 1], np.int32), np.array([0], np.int32), None)),
    ("rmat12", lambda r: O.rmat_csr(12, 16, seed=5) + (None,)),
    ("rmat12_weighted", lambda r: (lambda V, E: (V, E, O.edge_weights(V, E, 9)))(*O.rmat_csr(12, 16, seed=5))),
]


@pytest.mark.parametrize("name,make", CSC_CASES, ids=[c[0] for c in CSC_CASES])
@pytest.mark.parametrize("alpha", [0.85, 0.75])
def test_csc_build_bit_exact(gpu_ready, name, make, alpha):
    _say(f"csc build {name} alpha={alpha}")
    V, E, W = make(np.random.default_rng(abs(hash(name)) % 1000))
    g = native.DeviceGraph.from_csr(V, E, W, alpha=alpha)
    in_row, in_col, in_val, dang = g.download_csc()
    r_row, r_col, r_val, r_dang, _ = O.build_inbound(V, E, W, alpha)
    assert np.array_equal(in_row, r_row)
    assert np.array_equal(in_col, r_col)
    assert np.array_equal(in_val.view(np.uint32), r_val.view(np.uint32))
    assert np.array_equal(dang, r_dang)


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("weight", [None, "weight"])
def test_dict_graph_device_csc_bit_exact(gpu_ready, directed, weight):
    # the backend hands unweighted graphs over in adjacency order (no per-row
    # column sort): the pull CSC must still be the one built from the
    # gen_CSR-ordered out-CSR, bit for bit; weighted graphs go over sorted
    rng = np.random.default_rng(21)
    n = 3000
    nodes = list(range(n))
    rng.shuffle(nodes)
    G = DiGraph() if directed else Graph()
    G.add_nodes_from(nodes)
    for _ in range(12 * n):
        G.add_edge(nodes[int(rng.integers(n))], nodes[int(rng.integers(n))], weight=float(rng.uniform(0.1, 2)))
    prep = gpu_backend._prepare_graph_cache(G)
    Vu, Eu, _ = gpu_backend._out_arcs(G, prep, None, sort=False)
    Vs, Es, _ = gpu_backend.graph_out_csr(G, prep, None)
    assert not np.array_equal(Eu, Es)  # the case: rows really are out of order
    _say(f"dict graph directed={directed} weight={weight} n={n} m={len(Es)}")
    g = gpu_backend._get_cached_cpp_graph(G, prepared=prep, weight=weight, alpha=0.85)
    in_row, in_col, in_val, dang = g.download_csc()
    V, E, W = gpu_backend.graph_out_csr(G, prep, weight)
    r_row, r_col, r_val, r_dang, _ = O.build_inbound(V, E, W, 0.85)
    assert np.array_equal(in_row, r_row)
    assert np.array_equal(in_col, r_col)
    assert np.array_equal(in_val.view(np.uint32), r_val.view(np.uint32))
    assert np.array_equal(dang, r_dang)


def test_all_ones_weights_take_the_unweighted_path(gpu_ready):
    # the reference's C++ route hands unweighted_W (m ones, graph.cpp:881) as W
    V, E = O.rmat_csr(13, 16, seed=12)
    ones = np.ones(len(E))
    g1 = native.DeviceGraph.from_csr(V, E, ones)
    g0 = native.DeviceGraph.from_csr(V, E, None)
    assert g1.info()["weighted"] == 0
    in1, in0 = g1.download_csc(), g0.download_csc()
    r = O.build_inbound(V, E, ones, 0.85)
    for got in (in1, in0):
        assert np.array_equal(got[0], r[0]) and np.array_equal(got[1], r[1]) and np.array_equal(got[3], r[3])
        assert np.array_equal(got[2].view(np.uint32), r[2].view(np.uint32))
    p1, i1, _ = g1.pagerank(0.85, 100, 1e-6)
    p0, i0, _ = g0.pagerank(0.85, 100, 1e-6)
    assert i1 == i0 and np.array_equal(p1.view(np.uint64), p0.view(np.uint64))
    # one weight off 1: weighted
    w = ones.copy()
    w[7] = 0.5
    assert native.DeviceGraph.from_csr(V, E, w).info()["weighted"] == 1
    # through the raw C++-route entry as well
    out, it, _ = native.pagerank_csr(V, E, ones, 0.85, 100, 1e-6, graph_id=77)
    assert it == i0 and np.array_equal(out.view(np.uint64), p0.view(np.uint64))
    g1.close()
    g0.close()


def test_alpha_change_refreshes_values_bit_exact(gpu_ready):
    V, E = O.rmat_csr(10, 16, seed=2)
    W = O.edge_weights(V, E, 4)
    g = native.DeviceGraph.from_csr(V, E, W, alpha=0.85)
    g.pagerank(alpha=0.9, max_iter=3, tol=0.0)
    _, _, in_val, _ = g.download_csc()
    _, _, r_val, _, _ = O.build_inbound(V, E, W, 0.9)
    assert np.array_equal(in_val.view(np.uint32), r_val.view(np.uint32))


# -- PageRank parity on the golden fixtures -------------------------------------
def _graph_from_case(case):
    G = DiGraph() if case["directed"] else Graph()
    if case["nodes"] is not None:
        G.add_nodes_from(case["nodes"])
    for e in case["edges"]:
        if len(e) == 3:
            G.add_edge(e[0], e[1], **{case["weight"]: e[2]})
        else:
            G.add_edge(e[0], e[1])
    return G


with open(os.path.join(GOLDEN, "golden_small.json")) as _fh:
    SMALL = json.load(_fh)


@pytest.mark.parametrize("case", SMALL["cases"], ids=[c["name"] for c in SMALL["cases"]])
def test_pagerank_golden_cases(gpu_ready, gpu_env, case):
    _say(f"golden {case['name']}")
    G = _graph_from_case(case)
    res = eggpu_amd.pagerank(G, alpha=case["alpha"], weight=case["weight"], max_iter=case["max_iter"],
                             tol=case["tol"])
    nxv = np.asarray(case["nx_values"])
    nxv = nxv / nxv.sum()
    ours = np.asarray([res[v] for v in case["nodes"]])
    _assert_close(ours, nxv, case["name"])
    assert gpu_backend.get_last_iterations("pagerank") == case["nx_iterations"]
    oracle_vals, oracle_it = O.pagerank_graph(G, alpha=case["alpha"], weight=case["weight"],
                                              max_iter=case["max_iter"], tol=case["tol"])
    assert oracle_it == case["nx_iterations"]


def test_readme_converged_known_answer(gpu_ready, gpu_env):
    G = Graph()
    G.add_edges_from([tuple(e) for e in SMALL["readme_edges"]])
    res = eggpu_amd.pagerank(G, max_iter=10000, tol=1e-15)
    kat = {int(k): v for k, v in SMALL["readme_kat"].items()}
    _assert_close([res[k] for k in kat], list(kat.values()), "readme KAT")


def test_rmat14_c1_parity(gpu_ready):
    """C1: R-MAT scale 14, alpha 0.85, tol 1e-6, max_iter 100."""
    z = np.load(os.path.join(GOLDEN, "rmat14.npz"))
    g = native.DeviceGraph.rmat(14, 16, seed=1, keep_csr=True)
    V, E, _ = g.download_csr()
    assert hashlib.sha256(E.tobytes()).hexdigest() == str(z["E_sha256"])
    pr, it, ks = g.pagerank(0.85, 100, 1e-6)
    nxv = z["nx_values"] / z["nx_values"].sum()
    _assert_close(pr, nxv, "C1 rmat14 fp32")
    assert it == int(z["nx_iterations"]) == 5
    ref, it_ref, _ = O.pagerank_csr(V, E, None, 0.85, 100, 1e-6)
    _assert_close(pr, ref, "C1 rmat14 vs oracle")
    assert it == it_ref
    pr64, it64, _ = g.pagerank(0.85, 100, 1e-6, precision=native.PRECISION_FP64)
    _assert_close(pr64, ref, "C1 rmat14 fp64", tol=FP64_TOL, l1=1e-10)
    assert it64 == it_ref


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("tol,max_iter", [(1e-6, 100), (0.0, 30), (1e-9, 200)])
def test_rmat_parity_fixed_and_converged(gpu_ready, weighted, tol, max_iter):
    V, E = O.rmat_csr(13, 16, seed=21)
    W = O.edge_weights(V, E, 3) if weighted else None
    g = native.DeviceGraph.from_csr(V, E, W)
    pr, it, _ = g.pagerank(0.85, max_iter, tol)
    ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, max_iter, tol)
    _assert_close(pr, ref, f"rmat13 weighted={weighted} tol={tol}")
    assert it == it_ref


@pytest.mark.parametrize("bins", [
    dict(thread_row_max=0, wave_row_max=64, hub_segment=256, tile_items=512),
    dict(thread_row_max=1000, wave_row_max=4096, hub_segment=8192, tile_items=3840),
    dict(thread_row_max=8, wave_row_max=128, hub_segment=1024, tile_items=7000, reorder=0),
    dict(thread_row_max=16, wave_row_max=7000, hub_segment=100000, tile_items=900),
])
def test_bin_settings_do_not_change_results(gpu_ready, bins):
    V, E = O.rmat_csr(14, 16, seed=1)
    g = native.DeviceGraph.from_csr(V, E)
    g.set_bins(**bins)
    pr, it, _ = g.pagerank(0.85, 100, 1e-6)
    ref, it_ref, _ = O.pagerank_csr(V, E, None, 0.85, 100, 1e-6)
    _assert_close(pr, ref, f"bins {bins}")
    assert it == it_ref
    pr64, _, _ = g.pagerank(0.85, 20, 0.0, precision=native.PRECISION_FP64)
    ref64, _, _ = O.pagerank_csr(V, E, None, 0.85, 20, 0.0)
    _assert_close(pr64, ref64, f"bins fp64 {bins}", tol=FP64_TOL, l1=1e-10)


@pytest.mark.parametrize("weighted", [False, True])
def test_source_sorted_tiles_are_bit_identical(gpu_ready, weighted):
    """Gathering tile arcs in source order writes the same terms into the same
    LDS slots, and non-temporal loads of cold sources read the same values:
    every result bit is unchanged, also after an alpha change (which
    refreshes the sorted weights)."""
    V, E = O.rmat_csr(15, 16, seed=9)
    W = O.edge_weights(V, E, 4) if weighted else None
    g = native.DeviceGraph.from_csr(V, E, W)
    runs = {}
    for ts, hot in ((0, -1), (1, -1), (1, 1000), (0, 5000)):
        g.set_bins(tile_sort=ts, tile_items=2944, hot_sources=hot)
        assert g.info()["tile_sort"] == ts and g.info()["hot_sources"] == max(hot, 0)
        runs[ts, hot] = [g.pagerank(0.85, 100, 1e-6), g.pagerank(0.9, 30, 0.0),
                         g.pagerank(0.85, 20, 0.0, precision=native.PRECISION_FP64)]
    for key in runs:
        for (a, ia, _), (b, ib, _) in zip(runs[0, -1], runs[key]):
            assert ia == ib and np.array_equal(a, b), key
    runs[1] = runs[1, 1000]
    ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 100, 1e-6)
    _assert_close(runs[1][0][0], ref, f"sorted tiles weighted={weighted}")
    assert runs[1][0][1] == it_ref


def test_solves_are_deterministic(gpu_ready):
    g = native.DeviceGraph.rmat(16, 16, seed=4)
    a, ia, _ = g.pagerank(0.85, 40, 0.0)
    b, ib, _ = g.pagerank(0.85, 40, 0.0)
    assert ia == ib and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_edge_cases(gpu_ready, gpu_env):
    assert eggpu_amd.pagerank(DiGraph()) == {}
    G = DiGraph()
    G.add_node("solo")
    assert dict(eggpu_amd.pagerank(G)) == {"solo": 1.0}
    G = DiGraph()
    G.add_nodes_from(range(4))
    assert dict(eggpu_amd.pagerank(G)) == {i: 0.25 for i in range(4)}
    G = DiGraph()
    G.add_edge(0, 0)
    assert dict(eggpu_amd.pagerank(G)) == {0: 1.0}
    G = DiGraph()
    G.add_edges_from([(1, 2), (2, 3), ("String", "Bool"), (2, 1), (0, 0), ((None, None), (None, None))])
    res = eggpu_amd.pagerank(G)
    ref, _ = O.pagerank_graph(G)
    assert set(res.keys()) == set(G.nodes)
    _assert_close([res[k] for k in ref], list(ref.values()), "mixed node types")


# -- the reference's cache / mutation contracts ---------------------------------
def test_pagerank_cache_invalidates_when_alpha_changes(gpu_ready, gpu_env):
    """test_eggpu_bulk_csr.py:339-378."""
    G = DiGraph()
    G.add_nodes(range(7))
    G.add_edges([(0, 1), (0, 2), (1, 2), (2, 0), (2, 3), (3, 2), (3, 4), (4, 5), (5, 4), (5, 6)])
    eggpu_amd.pagerank(G, alpha=0.75, max_iter=200, tol=1.0e-7)
    actual = eggpu_amd.pagerank(G, alpha=0.90, max_iter=200, tol=1.0e-7)
    expected, _ = O.pagerank_graph(G, alpha=0.90, max_iter=200, tol=1.0e-7)
    _assert_close([actual[i] for i in range(7)], [expected[i] for i in range(7)], "alpha change")


def test_weight_mutation_rebuilds(gpu_ready, gpu_env):
    """run_eggpu_gpu_contract_regression.py:83-118."""
    G = DiGraph()
    G.add_edge(0, 1, weight=8.0)
    G.add_edge(0, 2, weight=1.0)
    G.add_edge(1, 2, weight=2.0)
    G.add_edge(2, 0, weight=1.0)
    first = dict(eggpu_amd.pagerank(G, weight="weight"))
    G[0][1]["weight"] = 0.25
    second = eggpu_amd.pagerank(G, weight="weight")
    fresh = DiGraph()
    for u, v, w in [(0, 1, 0.25), (0, 2, 1.0), (1, 2, 2.0), (2, 0, 1.0)]:
        fresh.add_edge(u, v, weight=w)
    ref, _ = O.pagerank_graph(fresh, weight="weight")
    _assert_close([second[k] for k in ref], list(ref.values()), "weight mutation")
    assert any(abs(first[k] - second[k]) > 1e-4 for k in ref)


def test_device_graph_is_reused_across_calls(gpu_ready, gpu_env):
    G = DiGraph()
    G.add_edges_from([(i, (i * 7 + 3) % 50) for i in range(50)] + [(i, (i + 1) % 50) for i in range(50)])
    eggpu_amd.pagerank(G)
    key = next(k for k in G.cache if k.startswith("__egx_device_graph__"))
    dg = G.cache[key][1]
    eggpu_amd.pagerank(G, alpha=0.8)
    assert G.cache[key][1] is dg


def test_raw_csr_entry_point_cache(gpu_ready):
    """egx_pagerank: per-thread prepared graph keyed like pagerank.cu:140-167."""
    native.cache_reset()
    V, E = O.rmat_csr(11, 16, seed=8)
    W = O.edge_weights(V, E, 1)
    Vc, Ec, Wc = (np.ascontiguousarray(a) for a in (V, E, W))
    r1, it1, _ = native.pagerank_csr(Vc, Ec, Wc, 0.85, 100, 1e-6, graph_id=42)
    r2, _, _ = native.pagerank_csr(Vc, Ec, Wc, 0.85, 100, 1e-6, graph_id=42)
    r3, _, _ = native.pagerank_csr(Vc, Ec, Wc, 0.9, 100, 1e-6, graph_id=42)
    st = native.cache_stats()
    assert st["builds"] == 1 and st["hits"] == 2 and st["value_refreshes"] == 1
    ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 100, 1e-6)
    ref9, _, _ = O.pagerank_csr(V, E, W, 0.9, 100, 1e-6)
    _assert_close(r1, ref, "raw csr")
    assert it1 == it_ref and np.array_equal(r1, r2)
    _assert_close(r3, ref9, "raw csr alpha 0.9")
    native.pagerank_csr(Vc, Ec, Wc, 0.85, 100, 1e-6, graph_id=43)
    st = native.cache_stats()
    assert st["builds"] == 2 and st["entries"] == 2 and st["resident_bytes"] > 0
    native.cache_reset()


def test_raw_csr_cache_is_lru_with_limits(gpu_ready, monkeypatch):
    """Entry limit and eviction order after device_graph_cache.cu:53-168."""
    native.cache_reset()
    V, E = O.rmat_csr(10, 16, seed=2)
    Vc, Ec = np.ascontiguousarray(V), np.ascontiguousarray(E)
    monkeypatch.setenv("EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_ENTRIES", "2")
    for gid in (1, 2):
        native.pagerank_csr(Vc, Ec, None, 0.85, 50, 1e-6, graph_id=gid)
    native.pagerank_csr(Vc, Ec, None, 0.85, 50, 1e-6, graph_id=1)  # 1 is now the most recent
    native.pagerank_csr(Vc, Ec, None, 0.85, 50, 1e-6, graph_id=3)  # evicts 2
    st = native.cache_stats()
    assert st["entries"] == 2 and st["evictions"] == 1 and st["builds"] == 3 and st["hits"] == 1
    native.pagerank_csr(Vc, Ec, None, 0.85, 50, 1e-6, graph_id=1)
    assert native.cache_stats()["hits"] == 2
    native.pagerank_csr(Vc, Ec, None, 0.85, 50, 1e-6, graph_id=2)  # rebuilt
    assert native.cache_stats()["builds"] == 4
    monkeypatch.setenv("EASYGRAPH_GPU_DISABLE_DEVICE_CSR_CACHE", "1")
    before = native.cache_stats()
    r, _, _ = native.pagerank_csr(Vc, Ec, None, 0.85, 50, 1e-6, graph_id=1)
    after = native.cache_stats()
    assert after["builds"] == before["builds"] + 1 and after["hits"] == before["hits"]
    ref, _, _ = O.pagerank_csr(V, E, None, 0.85, 50, 1e-6)
    _assert_close(r, ref, "uncached raw csr")
    native.cache_reset()


def test_bulk_graph_matches_regular_graph(gpu_ready, gpu_env, tmp_path):
    """test_eggpu_bulk_csr.py:296-310 (undirected, alpha 0.75)."""
    edges = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3)]
    regular = Graph()
    regular.add_nodes(range(6))
    regular.add_edges(edges)
    rows = [[] for _ in range(6)]
    for u, v in edges:
        rows[u].append(v)
        rows[v].append(u)
    V = np.zeros(7, np.int32)
    V[1:] = np.cumsum([len(r) for r in rows])
    E = np.array([v for r in rows for v in r], np.int32)
    bulk = read_eggpu_csr(write_eggpu_csr(tmp_path / "tiny", V, E, directed=False))
    a = eggpu_amd.pagerank(regular, alpha=0.75, max_iter=200, tol=1e-6)
    b = eggpu_amd.pagerank(bulk, alpha=0.75, max_iter=200, tol=1e-6)
    assert np.allclose([a[i] for i in range(6)], [b[i] for i in range(6)], rtol=1e-6, atol=1e-8)


# -- full-size properties ---------------------------------------------------------
def test_scale22_csc_invariants(gpu_ready):
    """C3-sized graph: the pull CSC is a stable transpose of the out-CSR
    (in-degree histogram, arc multiset checksum, rows sorted)."""
    _say("scale 22 build")
    g = native.DeviceGraph.rmat(22, 16, seed=1, keep_csr=True)
    V, E, _ = g.download_csr()
    in_row, in_col, _, dang = g.download_csc()
    n = len(V) - 1
    assert in_row[0] == 0 and in_row[-1] == len(E) == len(in_col)
    assert np.array_equal(np.diff(in_row), np.bincount(E, minlength=n))
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(V))
    dst_of_slot = np.repeat(np.arange(n, dtype=np.int64), np.diff(in_row))
    k1 = np.sort(src * n + E.astype(np.int64))
    k2 = np.sort(in_col.astype(np.int64) * n + dst_of_slot)
    assert np.array_equal(k1, k2)
    # sources ascending inside each row
    same = dst_of_slot[1:] == dst_of_slot[:-1]
    assert (in_col[1:][same] > in_col[:-1][same]).all()
    assert np.array_equal(dang, (np.diff(V) == 0).astype(np.int32))


@pytest.mark.parametrize("blocked", [1, 0])
def test_scale24_solve_properties(gpu_ready, blocked):
    """C2 size: 1 iteration at tol=1e-6 (err < tol*n for n > 2e6), ranks sum
    to 1, and FP32 iterates stay within 1e-9 of FP64 ones after 20 iterations,
    on the blocked iteration (the default at this size) and on the pull SpMV."""
    _say(f"scale 24 build, blocked={blocked}")
    g = native.DeviceGraph.rmat(24, 16, seed=1)
    g.set_bins(blocked=blocked)
    assert g.info()["blocked"] == blocked
    pr, it, _ = g.pagerank(0.85, 100, 1e-6)
    assert it == 1 and abs(pr.sum() - 1.0) < 1e-9
    p32, _, _ = g.pagerank(0.85, 20, 0.0)
    p64, _, _ = g.pagerank(0.85, 20, 0.0, precision=native.PRECISION_FP64)
    err = np.abs(p32 - p64)
    _say(f"scale 24 fp32 vs fp64 after 20 iterations: max-abs {err.max():.3e} L1 {err.sum():.3e}")
    assert err.max() < 1e-9 and err.sum() < 1e-5


@pytest.mark.parametrize("blocked", [1, 0])
def test_scale20_fixed_k_vs_openmp_port(gpu_ready, blocked):
    V, E = O.rmat_csr(20, 16, seed=1)
    g = native.DeviceGraph.from_csr(V, E)
    g.set_bins(blocked=blocked)
    assert g.info()["blocked"] == blocked
    pr, it, _ = g.pagerank(0.85, 20, 0.0)
    omp, it_omp = O.pagerank_omp(V, E, None, 0.85, 20, 0.0, threads=16)
    assert it == it_omp == 20
    _assert_close(pr, omp / omp.sum(), "scale 20 K=20 vs OpenMP port")
