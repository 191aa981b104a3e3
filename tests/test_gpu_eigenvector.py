"""GPU tests of eigenvector centrality on the SpMV engine
(egx_eigenvector_csr, eggpu_amd.eigenvector_centrality) against the oracle's
restatement of eigenvector.py:62-154 and eigenvector.cpp:33-97.

Iterates and sums are FP64 as in the reference.  Weights are gathered as FP32
values, so the oracle is given the same fl32-rounded weights; the results then
differ only in summation order and are held to max-abs 1e-12 on unit-norm
vectors, with identical iteration counts and stop reasons.  The default route
(ARPACK in the reference, the power iteration on A + I here) is held to the
ARPACK vector at a tight tolerance, 1e-7.
"""
import numpy as np
import pytest

import eggpu_amd
from eggpu_amd import native
from eggpu_amd.classes import DiGraph, Graph, MultiGraph
from eggpu_amd.exception import EasyGraphError, EasyGraphNotImplemented, EasyGraphPointlessConcept
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROUTES = {
    "py": dict(delta=2, start_norm=1, norm_floor=1e-10),
    "cpp": dict(delta=1, start_norm=2, norm_floor=1e-12),
}


def _say(msg):
    print(f"[gpu-eig] {msg}", flush=True)


def _fl32(w):
    return None if w is None else np.asarray(w, dtype=np.float32).astype(np.float64)


def _graph(V, E, directed, W=None):
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    G = DiGraph() if directed else Graph()
    G.add_nodes_from(range(n))
    if W is None:
        G.add_edges_from(zip(src.tolist(), dst.tolist()))
    else:
        for u, v, w in zip(src.tolist(), dst.tolist(), W.tolist()):
            G.add_edge(u, v, weight=w)
    return G


def _check(V, E, W, x0, label, **kw):
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    _say(f"{label}: n={n} m={len(E)} {kw}")
    x, it, stop, _ = native.eigenvector_csr(V, E, W, x0, **kw)
    ref, it_ref, stop_ref = O.eigenvector_rows(n, src, dst, _fl32(W), x0, **kw)
    err = float(np.max(np.abs(x - ref))) if n else 0.0
    _say(f"  max-abs {err:.3e}, iterations {it} stop {stop} (oracle {it_ref} stop {stop_ref})")
    assert (it, stop) == (it_ref, stop_ref)
    assert err <= 1e-12
    return x, it, stop


@pytest.mark.parametrize("route", ["py", "cpp"])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("directed", [False, True])
def test_eigenvector_csr_matches_oracle(gpu_ready, route, weighted, directed):
    V, E = O.rmat_csr(11, 8, seed=21)
    if not directed:
        src, dst = O.csr_rows(V, E)
        n = len(V) - 1
        V, E = O.simple_csr_from_coo(n, np.concatenate([src, dst]), np.concatenate([dst, src]))
    W = O.edge_weights(V, E, 5) if weighted else None
    x0 = np.random.default_rng(3).uniform(0.5, 2.0, len(V) - 1)
    for tol in (1e-6, 1e-10):
        _check(V, E, W, x0, f"{route} weighted={weighted} directed={directed}", max_iter=300, tol=tol,
               **ROUTES[route])


@pytest.mark.parametrize("shift", [0.0, 1.0])
def test_eigenvector_csr_fixed_iterations_and_shift(gpu_ready, shift):
    # tol 0: every iteration runs; the last iterate must agree
    V, E = O.rmat_csr(12, 8, seed=4)
    x0 = np.ones(len(V) - 1)
    for k in (1, 2, 3, 17):
        _check(V, E, None, x0, f"K={k} shift={shift}", max_iter=k, tol=0.0, shift=shift, **ROUTES["py"])


def test_eigenvector_csr_large(gpu_ready):
    # R-MAT scale 18: the oracle's sparse product still finishes in seconds
    V, E = O.rmat_csr(18, 16, seed=2)
    x, it, stop = _check(V, E, None, np.ones(len(V) - 1), "scale 18", max_iter=60, tol=1e-9, **ROUTES["cpp"])
    assert abs(np.linalg.norm(x) - 1.0) < 1e-12


def test_eigenvector_csr_edge_cases(gpu_ready):
    # nilpotent path: the norm test keeps x_2 at the third iteration
    V = np.array([0, 1, 2, 2], dtype=np.int32)
    E = np.array([1, 2], dtype=np.int32)
    for route in ROUTES:
        x, it, stop = _check(V, E, None, np.ones(3), f"path {route}", max_iter=50, tol=1e-6, **ROUTES[route])
        assert (it, stop) == (3, native.EIG_NORM_FLOOR) and np.array_equal(x, [0.0, 0.0, 1.0])
    # directed 3-cycle: the C++ twin's start is already the fixed point
    Vc = np.array([0, 1, 2, 3], dtype=np.int32)
    Ec = np.array([1, 2, 0], dtype=np.int32)
    x, it, stop = _check(Vc, Ec, None, np.ones(3), "cycle", max_iter=50, tol=1e-6, **ROUTES["cpp"])
    assert (it, stop) == (1, native.EIG_CONVERGED)
    # max_iter 0 returns the scaled start; no arcs: the first norm test stops
    x, it, stop, _ = native.eigenvector_csr(Vc, Ec, None, np.array([1.0, 2.0, 1.0]), max_iter=0)
    assert (it, stop) == (0, native.EIG_MAX_ITER) and np.allclose(x, [0.25, 0.5, 0.25], rtol=0, atol=1e-15)
    V0 = np.zeros(4, dtype=np.int32)
    E0 = np.zeros(0, dtype=np.int32)
    x, it, stop, _ = native.eigenvector_csr(V0, E0, None, np.array([1.0, 2.0, 1.0]), max_iter=10)
    ref, it_ref, stop_ref = O.eigenvector_rows(3, np.zeros(0, int), np.zeros(0, int), None, [1.0, 2.0, 1.0],
                                               max_iter=10)
    assert (it, stop) == (it_ref, stop_ref) == (1, native.EIG_NORM_FLOOR) and np.array_equal(x, ref)
    # no arcs with the shift: x_1 = x_0 / ||x_0||, then x_2 = x_1
    x, it, stop, _ = native.eigenvector_csr(V0, E0, None, np.array([1.0, 2.0, 1.0]), max_iter=10, shift=1.0)
    ref, it_ref, stop_ref = O.eigenvector_rows(3, np.zeros(0, int), np.zeros(0, int), None, [1.0, 2.0, 1.0],
                                               max_iter=10, shift=1.0)
    assert (it, stop) == (it_ref, stop_ref) and np.max(np.abs(x - ref)) < 1e-15
    # negative weights and bad options are refused; n = 0 writes nothing
    with pytest.raises(Exception):
        native.eigenvector_csr(Vc, Ec, np.array([1.0, -1.0, 1.0]), None)
    with pytest.raises(Exception):
        native.eigenvector_csr(Vc, Ec, None, None, norm_floor=0.0)
    x, it, stop, _ = native.eigenvector_csr(np.zeros(1, dtype=np.int32), E0, None, None)
    assert len(x) == 0 and (it, stop) == (0, 0)


@pytest.mark.parametrize("directed", [False, True])
def test_eigenvector_api_nstart_matches_oracle(gpu_ready, gpu_env, directed):
    # the Python route with nstart, weights and isolated nodes
    V, E = O.rmat_csr(10, 8, seed=11)
    W = _fl32(O.edge_weights(V, E, 2))
    G = _graph(V, E, directed, W)
    G.add_nodes_from(range(len(G), len(G) + 3))  # isolated: 1e-4 self-loop, then 0
    rng = np.random.default_rng(6)
    nstart = {v: float(rng.uniform(0.5, 2.0)) for v in range(0, len(G), 2)}
    _say(f"api nstart directed={directed} n={len(G)}")
    try:
        ref = O.eigenvector_graph(G, max_iter=500, tol=1e-9, nstart=nstart, weight="weight")
    except RuntimeError:
        with pytest.raises(EasyGraphError):
            eggpu_amd.eigenvector_centrality(G, max_iter=500, tol=1e-9, nstart=nstart, weight="weight")
        _say("  both raise (no convergence)")
        return
    got = eggpu_amd.eigenvector_centrality(G, max_iter=500, tol=1e-9, nstart=nstart, weight="weight")
    err = max(abs(got[v] - ref[v]) for v in ref)
    _say(f"  max-abs {err:.3e}")
    assert set(got) == set(ref) and err <= 1e-12
    assert all(got[v] == 0.0 for v in range(len(G) - 3, len(G)))


@pytest.mark.parametrize("weighted", [False, True])
def test_eigenvector_default_route_matches_arpack(gpu_ready, gpu_env, weighted):
    # undirected R-MAT with isolated nodes: the shifted power iteration and
    # ARPACK reach the same principal eigenvector
    V, E = O.rmat_csr(11, 8, seed=8)
    src, dst = O.csr_rows(V, E)
    n = len(V) - 1
    V, E = O.simple_csr_from_coo(n, np.concatenate([src, dst]), np.concatenate([dst, src]))
    W = _fl32(O.edge_weights(V, E, 3)) if weighted else None
    G = _graph(V, E, False, W)
    wk = "weight" if weighted else None
    ref = O.eigenvector_graph(G, max_iter=2000, tol=1e-13, weight=wk)
    got = eggpu_amd.eigenvector_centrality(G, max_iter=200, tol=1e-11, weight=wk)
    err = max(abs(got[v] - ref[v]) for v in ref)
    loose = eggpu_amd.eigenvector_centrality(G, weight=wk)
    err_default = max(abs(loose[v] - ref[v]) for v in ref)
    _say(f"default route weighted={weighted}: max-abs {err:.3e} at tol 1e-11, {err_default:.3e} at tol 1e-6")
    assert err <= 1e-7


def test_eigenvector_api_closed_forms_and_errors(gpu_ready, gpu_env):
    K = Graph()
    K.add_edges_from([(i, j) for i in range(5) for j in range(i + 1, 5)])
    for x in (eggpu_amd.eigenvector_centrality(K), eggpu_amd.eigenvector_centrality(K, nstart={})):
        assert max(abs(x[v] - 1 / np.sqrt(5)) for v in range(5)) < 1e-12
    # star K_{1,4}: the default route converges to the closed form; the
    # nstart power iteration oscillates and raises, as the reference's does
    S = Graph()
    S.add_edges_from([(0, k) for k in range(1, 5)])
    x = eggpu_amd.eigenvector_centrality(S, tol=1e-12)
    assert abs(x[0] - 1 / np.sqrt(2)) < 1e-10 and max(abs(x[k] - 1 / np.sqrt(8)) for k in range(1, 5)) < 1e-10
    with pytest.raises(EasyGraphError):
        eggpu_amd.eigenvector_centrality(S, nstart={})
    with pytest.raises(EasyGraphError):
        eggpu_amd.eigenvector_centrality(K, nstart={}, max_iter=0)
    # triangle plus an isolated node
    T = Graph()
    T.add_nodes_from(range(4))
    T.add_edges_from([(0, 1), (1, 2), (0, 2)])
    for x in (eggpu_amd.eigenvector_centrality(T), eggpu_amd.eigenvector_centrality(T, nstart={})):
        assert x[3] == 0.0 and max(abs(x[v] - 1 / np.sqrt(3)) for v in range(3)) < 1e-9
    with pytest.raises(EasyGraphPointlessConcept):
        eggpu_amd.eigenvector_centrality(Graph())
    one = Graph()
    one.add_node(0)
    with pytest.raises(EasyGraphPointlessConcept):
        eggpu_amd.eigenvector_centrality(one)
    with pytest.raises(EasyGraphNotImplemented):
        eggpu_amd.eigenvector_centrality(MultiGraph())
