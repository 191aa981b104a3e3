"""GPU tests of Katz centrality on the SpMV engine (egx_katz_csr,
eggpu_amd.katz_centrality) against the oracle's restatement of
katz_centrality.py:7-110.

Iterates, sums and alpha * (row sum) are FP64 as in the reference, which
differs only in summation order, so values are held to max-abs 1e-12 on
unit-norm vectors (1e-12 relative unnormalised) and the iteration count
must be identical.  The closed-form cases (tol passing at
the first or second update) must match exactly up to FP64 rounding.
"""
import numpy as np
import pytest

import eggpu_amd
from eggpu_amd import native
from eggpu_amd.classes import DiGraph, Graph, MultiGraph
from eggpu_amd.exception import EasyGraphNotImplemented
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-katz] {msg}", flush=True)


def _graph(V, E, directed):
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    G = DiGraph() if directed else Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    return G


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("scale,alpha", [(9, 0.01), (12, 0.004)])
def test_katz_matches_oracle(gpu_ready, gpu_env, directed, scale, alpha):
    V, E = O.rmat_csr(scale, 8, seed=scale + 2)
    G = _graph(V, E, directed)
    _say(f"scale {scale} directed={directed} alpha={alpha}")
    got = eggpu_amd.katz_centrality(G, alpha=alpha)
    ref, it_ref = O.katz_graph(G, alpha=alpha)
    assert ref is not None
    err = max(abs(got[v] - ref[v]) for v in ref)
    # the C ABI reports the iteration count
    from eggpu_amd import gpu_backend

    Va, Ea = gpu_backend.adjacency_csr(G, gpu_backend._prepare_graph_cache(G))
    x, it, conv, _ = native.katz_csr(Va, Ea, alpha=alpha)
    _say(f"  max-abs {err:.3e}, iterations {it} (oracle {it_ref})")
    assert err <= 1e-12 and conv and it == it_ref


def test_katz_beta_dict_and_unnormalised(gpu_ready, gpu_env):
    V, E = O.rmat_csr(10, 8, seed=7)
    G = _graph(V, E, True)
    rng = np.random.default_rng(4)
    beta = {v: float(rng.uniform(0.5, 2.0)) for v in range(0, len(G), 3)}
    got = eggpu_amd.katz_centrality(G, alpha=0.01, beta=beta, normalized=False, tol=1e-9)
    ref, _ = O.katz_graph(G, alpha=0.01, beta=beta, normalized=False, tol=1e-9)
    rel = max(abs(got[v] - ref[v]) / abs(ref[v]) for v in ref)
    _say(f"beta dict, unnormalised: max relative {rel:.3e}")
    assert rel <= 1e-12


def test_katz_closed_forms(gpu_ready, gpu_env):
    # directed star 0 -> 1..5: tol 1.0 passes on the second update and keeps x1
    G = DiGraph()
    G.add_edges_from([(0, k) for k in range(1, 6)])
    got = eggpu_amd.katz_centrality(G, alpha=0.125, beta=2.0, tol=1.0, normalized=False)
    assert [got[v] for v in range(6)] == [2.625, 2.0, 2.0, 2.0, 2.0, 2.0]
    # a test that passes at once keeps x0 = 1
    got = eggpu_amd.katz_centrality(G, alpha=0.125, beta=2.0, tol=1e9)
    assert np.allclose([got[v] for v in range(6)], 1.0 / np.sqrt(6.0), rtol=0, atol=1e-15)
    # no arcs: x1 = beta (second test sees no change)
    H = Graph()
    H.add_nodes_from(range(4))
    got = eggpu_amd.katz_centrality(H, beta=3.0, normalized=False)
    assert [got[v] for v in range(4)] == [3.0] * 4


def test_katz_errors_and_empty(gpu_ready, gpu_env):
    C = DiGraph()
    C.add_edges_from([(0, 1), (1, 2), (2, 0)])
    with pytest.raises(RuntimeError, match="failed to converge in 40 iterations"):
        eggpu_amd.katz_centrality(C, alpha=1.0, max_iter=40)
    with pytest.raises(RuntimeError):
        eggpu_amd.katz_centrality(C, max_iter=0)
    assert eggpu_amd.katz_centrality(Graph()) == {}
    with pytest.raises(EasyGraphNotImplemented):
        eggpu_amd.katz_centrality(MultiGraph())


def test_katz_node_labels_follow_node_order(gpu_ready, gpu_env):
    G = Graph()
    G.add_edges_from([("a", "b"), ("b", (1, 2)), ((1, 2), 7), (7, "a"), ("b", 7)])
    got = eggpu_amd.katz_centrality(G, alpha=0.05)
    ref, _ = O.katz_graph(G, alpha=0.05)
    assert list(got) == list(ref) and max(abs(got[v] - ref[v]) for v in ref) <= 1e-12
