"""Pins the CPU oracle (oracle/egx_oracle.c) before anything is checked against it.

Each oracle function is checked against something that does not share its
code: networkx 3.4.2 (golden fixtures made by tests/golden/make_golden.py),
the EasyGraph README known-answer vector, closed-form linear solves of the
PageRank fixed point, scipy's CSR transpose, and invariants.  A perturbed
oracle call shows the checks are tight enough to catch a wrong constant.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from eggpu_amd.classes import DiGraph, Graph
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load_small():
    with open(os.path.join(GOLDEN, "golden_small.json")) as fh:
        return json.load(fh)


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


def _fixed_point(n, src, dst, w, alpha, p=None):
    """Exact solution of x = alpha*(P^T x + dang(x) p) + (1-alpha) p, sum x = 1,
    by a dense linear solve (P row-normalised, dangling rows spread by p)."""
    p = np.full(n, 1.0 / n) if p is None else np.asarray(p, float)
    A = np.zeros((n, n))
    wts = np.ones(len(src)) if w is None else np.asarray(w, float)
    for s_, d_, w_ in zip(src, dst, wts):
        A[s_, d_] += w_
    out = A.sum(axis=1)
    dang = out <= 1e-15
    P = np.zeros_like(A)
    P[~dang] = A[~dang] / out[~dang, None]
    # G x = x with G = alpha P^T + alpha p d^T + (1-alpha) p 1^T (uses sum x = 1)
    Gm = alpha * P.T + alpha * np.outer(p, dang.astype(float)) + (1 - alpha) * np.outer(p, np.ones(n))
    M = np.eye(n) - Gm
    M[-1, :] = 1.0
    rhs = np.zeros(n)
    rhs[-1] = 1.0
    return np.linalg.solve(M, rhs)


# -- pagerank.py:75-139 ------------------------------------------------------
def test_readme_known_answer_vector():
    """Easy-Graph/README.md:168-173 holds the converged fixed point."""
    d = _load_small()
    G = Graph()
    G.add_edges_from([tuple(e) for e in d["readme_edges"]])
    vals, it = O.pagerank_graph(G, alpha=0.85, max_iter=10000, tol=1e-15)
    kat = {int(k): v for k, v in d["readme_kat"].items()}
    assert max(abs(vals[k] - kat[k]) for k in kat) < 1e-12
    # the default tol=1e-6 result is the unconverged iterate, 23 iterations
    vals6, it6 = O.pagerank_graph(G)
    assert it6 == 23
    assert max(abs(vals6[k] - kat[k]) for k in kat) < 2e-6


def test_readme_planted_mistake_is_caught():
    """A wrong damping factor (0.8499 for 0.85) must fail the KAT check."""
    d = _load_small()
    G = Graph()
    G.add_edges_from([tuple(e) for e in d["readme_edges"]])
    vals, _ = O.pagerank_graph(G, alpha=0.8499, max_iter=10000, tol=1e-15)
    kat = {int(k): v for k, v in d["readme_kat"].items()}
    assert max(abs(vals[k] - kat[k]) for k in kat) > 1e-7


@pytest.mark.parametrize("case_idx", range(11))
def test_networkx_golden_cases(case_idx):
    case = _load_small()["cases"][case_idx]
    G = _graph_from_case(case)
    vals, it = O.pagerank_graph(G, alpha=case["alpha"], weight=case["weight"], max_iter=case["max_iter"],
                                tol=case["tol"])
    nxv = np.asarray(case["nx_values"])
    nxv = nxv / nxv.sum()  # networkx skips the final sum-normalisation
    ours = np.asarray([vals[v] for v in case["nodes"]])
    assert np.abs(ours - nxv).max() < 1e-12, case["name"]
    assert it == case["nx_iterations"], case["name"]


@pytest.mark.parametrize("seed,weighted", [(1, False), (2, True), (3, False), (4, True)])
def test_fixed_point_closed_form(seed, weighted):
    rng = np.random.default_rng(seed)
    n = 40
    m = 150
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    pairs = np.unique(np.stack([src[keep], dst[keep]], 1), axis=0)
    src, dst = pairs[:, 0].astype(np.int32), pairs[:, 1].astype(np.int32)
    w = rng.uniform(0.1, 3.0, len(src)) if weighted else None
    ours, it, _ = O.pagerank_rows(n, src, dst, w, alpha=0.85, max_iter=100000, tol=1e-16)
    exact = _fixed_point(n, src, dst, w, 0.85)
    assert np.abs(ours - exact).max() < 1e-12


def test_special_cases():
    assert O.pagerank_graph(DiGraph())[0] == {}
    G = DiGraph()
    G.add_node("solo")
    assert O.pagerank_graph(G)[0] == {"solo": 1.0}  # test_pagerank.py:75-77
    G = DiGraph()
    G.add_nodes_from(range(5))
    vals, it = O.pagerank_graph(G)
    assert all(v == 0.2 for v in vals.values()) and it == 0  # pagerank.py:93-95
    G = DiGraph()
    G.add_edges_from([(i, (i + 1) % 7) for i in range(7)])  # directed cycle: uniform, one iteration
    vals, it = O.pagerank_graph(G)
    assert it == 1 and max(abs(v - 1 / 7) for v in vals.values()) < 1e-15
    G = DiGraph()
    G.add_edge(0, 0)
    assert O.pagerank_graph(G)[0] == {0: 1.0}


def test_mixed_node_types_follow_insertion_order():
    # test_pagerank.py:10-18 node types; the oracle keys results by node
    G = DiGraph()
    G.add_edges_from([(1, 2), (2, 3), ("String", "Bool"), (2, 1), (0, 0), ((None, None), (None, None))])
    vals, _ = O.pagerank_graph(G)
    assert list(vals.keys()) == list(G.nodes.keys())
    assert abs(sum(vals.values()) - 1.0) < 1e-12


# -- generators ----------------------------------------------------------------
def test_rmat14_reproduces_fixture_and_networkx():
    z = np.load(os.path.join(GOLDEN, "rmat14.npz"))
    V, E = O.rmat_csr(14, 16, seed=1)
    assert len(E) == int(z["m"])
    assert hashlib.sha256(V.tobytes()).hexdigest() == str(z["V_sha256"])
    assert hashlib.sha256(E.tobytes()).hexdigest() == str(z["E_sha256"])
    pr, it, _ = O.pagerank_csr(V, E, None, 0.85, 100, 1e-6)
    nxv = z["nx_values"] / z["nx_values"].sum()
    assert np.abs(pr - nxv).max() < 1e-12
    assert it == int(z["nx_iterations"]) == 5


def test_rmat_quadrant_statistics():
    """Top-bit pair frequencies of the raw draws follow (a, b, c, d)."""
    src, dst = O.rmat_coo(16, 400000, seed=9)
    q = (src >> 15) * 2 + (dst >> 15)
    freq = np.bincount(q, minlength=4) / len(q)
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)


def test_rmat_csr_is_simple_and_sorted():
    V, E = O.rmat_csr(10, 16, seed=3)
    src, dst = O.csr_rows(V, E)
    assert (src != dst).all()
    key = src.astype(np.int64) * 1024 + dst
    assert (np.diff(key) > 0).all()


def test_edge_weights_range_and_determinism():
    V, E = O.rmat_csr(10, 16, seed=3)
    w1 = O.edge_weights(V, E, wseed=5)
    w2 = O.edge_weights(V, E, wseed=5)
    assert np.array_equal(w1, w2)
    assert (w1 >= 0.1).all() and (w1 <= 1.0).all()
    assert np.array_equal(w1, w1.astype(np.float32).astype(np.float64))


# -- pagerank.cu:286-347 -------------------------------------------------------
def _random_csr(rng, n, m, invalid=False, bad_weights=False, dups=True):
    src = np.sort(rng.integers(0, n, m))
    dst = rng.integers(0, n, m)
    if not dups:
        pairs = np.unique(np.stack([src, dst], 1), axis=0)
        src, dst = pairs[:, 0], pairs[:, 1]
    if invalid and len(dst):
        idx = rng.choice(len(dst), max(1, len(dst) // 10), replace=False)
        dst[idx] = rng.choice([-1, -5, n, n + 3], len(idx))
    V = np.zeros(n + 1, np.int32)
    V[1:] = np.cumsum(np.bincount(src, minlength=n))
    W = rng.uniform(0.1, 2.0, len(dst))
    if bad_weights and len(W):
        W[rng.choice(len(W), max(1, len(W) // 8), replace=False)] = rng.choice([-1.0, np.nan, np.inf, 0.0],
                                                                                 max(1, len(W) // 8))
    return V, dst.astype(np.int32), W


@pytest.mark.parametrize("seed", range(6))
def test_build_inbound_matches_scipy_transpose(seed):
    """in_row/in_col equal scipy's CSR->CSC conversion of the valid arcs
    (a stable counting sort: sources ascending, duplicates in CSR order);
    in_val sums to alpha over each non-dangling source."""
    rng = np.random.default_rng(seed)
    n, m = 50, 400
    V, E, W = _random_csr(rng, n, m, invalid=(seed % 2 == 1), bad_weights=(seed >= 3))
    in_row, in_col, in_val, dang, outw = O.build_inbound(V, E, W, alpha=0.85)
    src, dst = O.csr_rows(V, E)
    ok = (dst >= 0) & (dst < n)
    w = np.where(np.isfinite(W) & (W >= 0), W, 0.0)[ok]
    A = sp.csr_matrix((np.ones(ok.sum()), (src[ok], dst[ok])), shape=(n, n))
    A.sum_duplicates = lambda: None  # keep duplicate arcs as separate slots
    T = sp.csr_matrix((np.arange(ok.sum(), dtype=float) + 1, E[ok], V_valid(V, ok)), shape=(n, n)).T.tocsr()
    assert np.array_equal(in_row, T.indptr.astype(np.int32))
    assert np.array_equal(in_col, T.indices.astype(np.int32))
    # values: the slot of each CSC entry is recovered from the stored index
    slot = T.data.astype(np.int64) - 1
    vsrc = src[ok][slot]
    wsum = np.bincount(src[ok], weights=w, minlength=n)
    assert np.array_equal(dang, (wsum <= 0).astype(np.int32))
    per_src = np.bincount(vsrc, weights=in_val.astype(np.float64), minlength=n)
    live = wsum > 0
    assert np.allclose(per_src[live], 0.85, rtol=1e-5)
    assert np.abs(outw - wsum).max() < 1e-12
    assert in_val.dtype == np.float32


def V_valid(V, ok):
    """Row pointer of the CSR restricted to the arcs flagged ok."""
    n = len(V) - 1
    src = np.repeat(np.arange(n), np.diff(V))
    Vv = np.zeros(n + 1, np.int64)
    Vv[1:] = np.cumsum(np.bincount(src[ok], minlength=n))
    return Vv


def test_build_inbound_clamps_malformed_offsets():
    """pagerank.cu:302-306 clamps each row's [V[u], V[u+1]) into [0, m]."""
    V = np.array([0, 5, 3, 6], np.int32)
    E = np.array([1, 2, 0, 1, 2, 0], np.int32)
    in_row, in_col, in_val, dang, _ = O.build_inbound(V, E, None, 0.85)
    # row 0 covers slots 0..4, row 1 is empty (5 > 3), row 2 covers 3..5
    arcs = sorted([(0, e) for e in E[0:5]] + [(2, e) for e in E[3:6]], key=lambda a: (a[1], a[0]))
    assert in_row[-1] == 8
    assert list(in_col) == [a[0] for a in arcs]
    assert list(dang) == [0, 1, 0]


# -- personalised PageRank -----------------------------------------------------
def test_ppr_matches_networkx_fixture():
    z = np.load(os.path.join(GOLDEN, "ppr_rmat10.npz"))
    V, E, W = z["V"], z["E"], z["W"]
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    for j, S in enumerate(z["seed_sets"]):
        S = [int(s) for s in S if s >= 0]
        p = np.zeros(n)
        p[S] = 1.0 / len(S)
        ours, it, _ = O.ppr_rows(n, src, dst, W, p, 0.85, 200, 1e-6)
        nxv = z["nx_values"][j] / z["nx_values"][j].sum()
        assert np.abs(ours - nxv).max() < 1e-12
        assert it == int(z["nx_iterations"][j])


def test_ppr_closed_form():
    rng = np.random.default_rng(11)
    n = 30
    src = rng.integers(0, n, 120)
    dst = rng.integers(0, n, 120)
    keep = src != dst
    pairs = np.unique(np.stack([src[keep], dst[keep]], 1), axis=0)
    src, dst = pairs[:, 0].astype(np.int32), pairs[:, 1].astype(np.int32)
    p = np.zeros(n)
    p[[3, 7]] = 0.5
    ours, _, _ = O.ppr_rows(n, src, dst, None, p, 0.9, 100000, 1e-16)
    assert np.abs(ours - _fixed_point(n, src, dst, None, 0.9, p)).max() < 1e-12


# -- pagerank.cpp:21-122 (CPU baseline) ---------------------------------------
@pytest.mark.parametrize("weighted", [False, True])
def test_openmp_baseline_agrees_with_oracle(weighted):
    V, E = O.rmat_csr(12, 16, seed=4)
    W = O.edge_weights(V, E, 2) if weighted else None
    ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 100, 1e-6)
    omp, it = O.pagerank_omp(V, E, W, 0.85, 100, 1e-6, threads=4)
    assert it == it_ref
    assert np.abs(omp / omp.sum() - ref).max() < 1e-12


def test_rmat_n_is_rmat_at_powers_of_two_and_rejects_out_of_range():
    V, E = O.rmat_n_csr(1 << 10, 16 << 10, seed=5)
    Vs, Es = O.rmat_csr(10, 16, seed=5)
    assert np.array_equal(V, Vs) and np.array_equal(E, Es)
    # n = 1000: the draws of the 1024-node graph, kept when both ends are < 1000
    V2, E2 = O.rmat_n_csr(1000, 16 << 10, seed=5)
    src, dst = O.csr_rows(Vs, Es)
    keep = (src < 1000) & (dst < 1000)
    assert len(V2) == 1001 and V2[-1] == int(keep.sum())
    assert np.array_equal(E2, dst[keep])


def test_hetero_workload_specs():
    from eggpu_amd import workloads

    a = workloads.hetero_specs(512)
    assert a == workloads.hetero_specs(512)
    assert len(a) == 512 and sum(s.kind == "er" for s in a) == 256
    ns = np.array([s.n for s in a])
    deg = np.array([s.nedges / s.n for s in a])
    assert ns.min() >= 1000 and ns.max() <= 1_000_000
    assert 4.0 - 1e-3 <= deg.min() and deg.max() <= 32.0 + 1e-3
    # log-uniform: about a third of the graphs per decade
    dec = np.bincount(np.minimum(np.log10(ns).astype(int) - 3, 2), minlength=3) / 512
    assert np.all(np.abs(dec - 1 / 3) < 0.08)
    pool = workloads.serving_pool_specs(32)
    assert len(pool) == 32 and any(s.weighted for s in pool) and not all(s.weighted for s in pool)
    assert min(s.nedges for s in pool) >= 1e4 and max(s.nedges for s in pool) <= 5e7


# -- Katz centrality (katz_centrality.py:7-110) --------------------------------
@pytest.mark.parametrize("directed", [False, True])
def test_katz_matches_networkx_at_convergence(directed):
    # networkx sums in-arcs (A^T x), starts from 0 and keeps the last update;
    # at a tight tolerance both reach the fixed point, so the reference's
    # out-arc sum equals networkx on the reversed graph.  (The docstring
    # example of katz_centrality.py:64-70, 0.370/0.447, is not unit-norm and
    # is not what the function computes: 0.4877/0.5120 here and in networkx.)
    import networkx as nx

    V, E = O.rmat_csr(7, 6, seed=3)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    H = nx.DiGraph() if directed else nx.Graph()
    H.add_nodes_from(range(n))
    H.add_edges_from(zip(src.tolist(), dst.tolist()))
    G = DiGraph() if directed else Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    x, _ = O.katz_graph(G, alpha=0.02, tol=1e-13)
    ref = nx.katz_centrality(H.reverse() if directed else H, alpha=0.02, tol=1e-15, max_iter=5000)
    assert max(abs(x[v] - ref[v]) for v in range(n)) < 1e-11
    path = Graph()
    path.add_edges_from([(0, 1), (1, 2), (2, 3)])
    xp, _ = O.katz_graph(path, alpha=0.05, tol=1e-13)
    rp = nx.katz_centrality(nx.path_graph(4), alpha=0.05, tol=1e-15, max_iter=5000)
    assert max(abs(xp[v] - rp[v]) for v in range(4)) < 1e-11 and abs(xp[0] - 0.48766) < 1e-5


def test_katz_keeps_the_iterate_before_the_passing_update_and_reads_out_arcs():
    # directed star 0 -> 1..5, alpha 0.1, beta 2: x1 = alpha*outdeg + beta =
    # (2.5, 2, ..., 2), |x1 - x0|_1 = 6.5; x2 = (3.0, 2, ..., 2), |x2 - x1|_1 = 0.5.
    # tol 1.0 passes on the second update and returns x1, not x2; in-arcs
    # would give (2, 2.2, ...) instead.
    G = DiGraph()
    G.add_edges_from([(0, k) for k in range(1, 6)])
    x, it = O.katz_graph(G, alpha=0.1, beta=2.0, tol=1.0, normalized=False)
    assert it == 2
    assert [x[v] for v in range(6)] == [2.5, 2.0, 2.0, 2.0, 2.0, 2.0]
    # a test that passes at once returns x0 = 1, normalised to 1/sqrt(n)
    x, it = O.katz_graph(G, alpha=0.1, beta=2.0, tol=1e9)
    assert it == 1 and np.allclose([x[v] for v in range(6)], 1.0 / np.sqrt(6.0), rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", [1, 2])
def test_katz_fixed_point_closed_form(seed):
    # converged Katz solves (I - alpha A) x = b; beta varies per node
    V, E = O.rmat_csr(8, 8, seed=seed)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    A = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    alpha = 0.5 / max(1.0, float(abs(sp.linalg.eigs(A.astype(float), k=1, return_eigenvectors=False)[0])))
    b = 1.0 + np.random.default_rng(seed).random(n)
    x, it = O.katz_rows(n, src, dst, alpha, b, max_iter=2000, tol=1e-13, normalized=False)
    exact = sp.linalg.spsolve((sp.identity(n) - alpha * A).tocsc(), b)
    assert it > 3 and np.max(np.abs(x - exact)) < 1e-11


def test_katz_not_converging_is_reported():
    # a directed cycle with alpha 1 grows without bound; max_iter 0 never tests
    G = DiGraph()
    G.add_edges_from([(0, 1), (1, 2), (2, 0)])
    assert O.katz_graph(G, alpha=1.0, max_iter=50)[0] is None
    assert O.katz_graph(G, alpha=0.1, max_iter=0)[0] is None


# -- Eigenvector centrality (eigenvector.py:62-154, eigenvector.cpp:33-97) -----
_EIG_ROUTES = {
    "py": dict(delta=2, start_norm=1, norm_floor=1e-10),
    "cpp": dict(delta=1, start_norm=2, norm_floor=1e-12),
}


def _dense_adj(n, src, dst, w=None):
    A = np.zeros((n, n))
    for k, (u, v) in enumerate(zip(src.tolist(), dst.tolist())):
        A[v, u] += 1.0 if w is None else w[k]
    return A


def _eig_case(seed):
    V, E = O.rmat_csr(6, 6, seed=seed)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    rng = np.random.default_rng(seed)
    return n, src, dst, rng.uniform(0.1, 1.0, len(src)), rng.uniform(0.5, 2.0, n)


@pytest.mark.parametrize("route", ["py", "cpp"])
@pytest.mark.parametrize("shift", [0.0, 1.0])
def test_eigenvector_iterates_are_normalised_matrix_powers(route, shift):
    # x_k = (A + s I)^k x_0 / ||.||_2 whatever the start scaling: the loop run
    # with tol 0 for k iterations must land on the dense closed form (a
    # transposed A, a dropped weight or a missed shift all fail it)
    n, src, dst, w, x0 = _eig_case(5)
    M = _dense_adj(n, src, dst, w) + shift * np.eye(n)
    for k in (1, 2, 7, 20):
        x, it, stop = O.eigenvector_rows(n, src, dst, w, x0, max_iter=k, tol=0.0, shift=shift, **_EIG_ROUTES[route])
        ref = np.linalg.matrix_power(M, k) @ x0
        ref /= np.linalg.norm(ref)
        assert (it, stop) == (k, 2)
        assert np.max(np.abs(x - ref)) < 1e-12


@pytest.mark.parametrize("route", ["py", "cpp"])
def test_eigenvector_stops_at_the_first_small_delta(route):
    # the stop iteration from the dense sequence: the first k with
    # delta(x_k, x_{k-1}) < tol, L2 (Python route) or L1 (C++ twin), where
    # x_0 is the start scaled by sum|x0| or ||x0||_2
    n, src, dst, w, x0 = _eig_case(9)
    M = _dense_adj(n, src, dst, w)
    kw = _EIG_ROUTES[route]
    xs = [x0 / np.sum(np.abs(x0)) if kw["start_norm"] == 1 else x0 / np.linalg.norm(x0)]
    for _ in range(400):
        y = M @ xs[-1]
        xs.append(y / np.linalg.norm(y))
    d = np.array([(np.sum(np.abs(xs[k] - xs[k - 1])) if kw["delta"] == 1 else np.linalg.norm(xs[k] - xs[k - 1]))
                  for k in range(1, len(xs))])
    for tol in (1e-3, 1e-6, 1e-9):
        k_ref = int(np.argmax(d < tol)) + 1
        assert d[k_ref - 1] < tol and np.min(np.abs(d / tol - 1.0)) > 1e-6
        x, it, stop = O.eigenvector_rows(n, src, dst, w, x0, max_iter=1000, tol=tol, **kw)
        assert (it, stop) == (k_ref, 1)
        assert np.max(np.abs(x - xs[k_ref])) < 1e-12


def _karate(weighted):
    import networkx as nx

    H = nx.karate_club_graph()
    G = Graph()
    G.add_nodes_from(H.nodes)
    for u, v, a in H.edges(data=True):
        if weighted:
            G.add_edge(u, v, weight=float(a["weight"]))
        else:
            G.add_edge(u, v)
    return H, G


@pytest.mark.parametrize("weighted", [False, True])
def test_eigenvector_matches_networkx_at_convergence(weighted):
    # networkx iterates A^T + I with an L1 test; ARPACK (the reference's
    # default), the nstart power iteration and networkx all reach the
    # principal eigenvector of the (connected, non-bipartite) karate graph
    import networkx as nx

    H, G = _karate(weighted)
    wk = "weight" if weighted else None
    ref = nx.eigenvector_centrality(H, max_iter=100000, tol=1e-15, weight=wk)
    arpack = O.eigenvector_graph(G, tol=1e-14, max_iter=1000, weight=wk)
    power = O.eigenvector_graph(G, nstart={v: 1.0 for v in G.nodes}, tol=1e-13, max_iter=100000, weight=wk)
    for got in (arpack, power):
        assert max(abs(got[v] - ref[v]) for v in H.nodes) < 1e-9


def test_eigenvector_closed_forms():
    # complete graph: uniform; star K_{1,4} (bipartite): ARPACK gives centre
    # 1/sqrt(2) and leaves 1/(2 sqrt(2)) while the plain power iteration
    # oscillates and raises; on a directed cycle the uniform direction is
    # fixed: the C++ twin's L2-scaled start passes at the first iteration,
    # the Python route's L1-scaled start (1/3 each) only at the second
    K = Graph()
    K.add_edges_from([(i, j) for i in range(5) for j in range(i + 1, 5)])
    for x in (O.eigenvector_graph(K), O.eigenvector_graph(K, nstart={})):
        assert max(abs(x[v] - 1 / np.sqrt(5)) for v in range(5)) < 1e-12
    S = Graph()
    S.add_edges_from([(0, k) for k in range(1, 5)])
    x = O.eigenvector_graph(S, tol=1e-12)
    assert abs(x[0] - 1 / np.sqrt(2)) < 1e-10 and max(abs(x[k] - 1 / np.sqrt(8)) for k in range(1, 5)) < 1e-10
    with pytest.raises(RuntimeError):
        O.eigenvector_graph(S, nstart={})
    n, src, dst = 3, np.array([0, 1, 2]), np.array([1, 2, 0])
    for route, k in (("cpp", 1), ("py", 2)):
        x, it, stop = O.eigenvector_rows(n, src, dst, max_iter=100, tol=1e-6, **_EIG_ROUTES[route])
        assert (it, stop) == (k, 1) and np.allclose(x, 1 / np.sqrt(3), rtol=0, atol=1e-15)


def test_eigenvector_norm_stop_and_isolated_nodes():
    # directed path 0 -> 1 -> 2 is nilpotent: x_1 = (0, 1, 1)/sqrt(2),
    # x_2 = (0, 0, 1), A x_2 = 0, so the third iteration's norm test keeps x_2
    src, dst = np.array([0, 1]), np.array([1, 2])
    for route in ("py", "cpp"):
        x, it, stop = O.eigenvector_rows(3, src, dst, max_iter=100, tol=1e-6, **_EIG_ROUTES[route])
        assert (it, stop) == (3, 3) and np.array_equal(x, [0.0, 0.0, 1.0])
    # a triangle plus an isolated node: the node gets a 1e-4 self-loop and
    # then 0; the triangle stays uniform
    G = Graph()
    G.add_nodes_from(range(4))
    G.add_edges_from([(0, 1), (1, 2), (0, 2)])
    for x in (O.eigenvector_graph(G, nstart={}), O.eigenvector_graph(G)):
        assert x[3] == 0.0 and max(abs(x[v] - 1 / np.sqrt(3)) for v in range(3)) < 1e-9


def test_isolated_loops_host_rewrite():
    # the product's CSR rewrite against a dense construction: A plus 1e-4 on
    # the diagonal of every node with zero in- and out-weight (zero-weight
    # arcs included), existing arcs unchanged
    from eggpu_amd.eigenvector import with_isolated_loops

    rng = np.random.default_rng(3)
    n = 40
    V, E = O.er_csr(n, 50, seed=4)
    src, dst = O.csr_rows(V, E)
    W = rng.uniform(0.1, 1.0, len(E))
    W[src == 7] = 0.0
    W[dst == 7] = 0.0
    V2, E2, W2, iso = with_isolated_loops(np.asarray(V, dtype=np.int64), np.asarray(E, dtype=np.int32), W)
    A = _dense_adj(n, src, dst, W)
    expect = np.flatnonzero((A.sum(axis=0) == 0) & (A.sum(axis=1) == 0))
    assert 7 in expect and np.array_equal(iso, expect)
    s2, d2 = O.csr_rows(V2, E2)
    B = _dense_adj(n, s2, d2, W2)
    assert np.array_equal(B, A + np.diag(np.where(np.isin(np.arange(n), expect), 1e-4, 0.0)))
    # unweighted with nothing isolated: returned as given
    Vc = np.array([0, 1, 2], dtype=np.int64)
    Ec = np.array([1, 0], dtype=np.int32)
    assert with_isolated_loops(Vc, Ec, None)[2] is None


@pytest.mark.parametrize("alpha", [0.85, 0.7, 0.95])
def test_build_inbound_values_pinned_bit_for_bit(alpha):
    """in_val = fl32(alpha * (w / outW)) (pagerank.cu:344), outW the sum of
    the source's sanitised weights in CSR order (pagerank.cu:308-316).

    Every source here has two arcs, weights w and 1.0, with w chosen so that
    alpha * (w / (w + 1)) lands within a few double ulps of a midpoint
    between two FP32 numbers: there the FP32 bits depend on the order of the
    double operations.  The expected value is formed from Python's IEEE
    doubles in the reference's order and rounded to FP32; the arcs are kept
    where (alpha * w) / outW, alpha * w * (1 / outW) or an FP32 division
    give other bits, so an oracle with any of those orders fails."""
    rng = np.random.default_rng(int(alpha * 100))
    ws, seps = [], [0, 0, 0]
    while len(ws) < 200:
        f = np.float32(rng.uniform(0.05, 0.6) * alpha)
        mid = (float(f) + float(np.nextafter(f, np.float32(1.0)))) / 2.0  # an FP32 rounding midpoint
        w0 = mid / (alpha - mid)  # alpha * w / (w + 1) = mid
        for k in range(-40, 41):
            w = float(np.nextafter(w0, np.inf) if k > 0 else w0)
            for _ in range(abs(k) - 1):
                w = float(np.nextafter(w, np.inf if k > 0 else -np.inf))
            s_ = w + 1.0
            v = np.float32(alpha * (w / s_))
            alts = (np.float32((alpha * w) / s_), np.float32(alpha * w * (1.0 / s_)),
                    np.float32(np.float32(alpha) * (np.float32(w) / np.float32(s_))))
            if any(a != v for a in alts):
                for j, a in enumerate(alts):
                    seps[j] += int(a != v)
                ws.append((w, v))
                break
    assert min(seps) >= 10, seps  # each alternative order is caught on some arcs
    n = len(ws) + 2
    V = np.arange(0, 2 * len(ws) + 1, 2, dtype=np.int32)
    V = np.concatenate([V, [V[-1], V[-1]]]).astype(np.int32)  # two sinks
    E = np.tile(np.array([n - 2, n - 1], np.int32), len(ws))
    W = np.empty(2 * len(ws))
    W[0::2] = [w for w, _ in ws]
    W[1::2] = 1.0
    in_row, in_col, in_val, dang, outw = O.build_inbound(V, E, W, alpha=alpha)
    row = in_val[in_row[n - 2]:in_row[n - 1]]  # the w arcs, sources ascending
    assert np.array_equal(in_col[in_row[n - 2]:in_row[n - 1]], np.arange(len(ws)))
    assert np.array_equal(row.view(np.uint32), np.array([v for _, v in ws], np.float32).view(np.uint32))
