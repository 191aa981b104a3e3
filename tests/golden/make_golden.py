"""Writes the golden fixtures of the PageRank parity tests.

Inputs are the reference's own test graphs and known-answer vector; the
expected values come from networkx 3.4.2 (``networkx.pagerank``), an
independent implementation with the reference's stopping rule
(``err < N * tol``, uniform start, dangling mass spread uniformly,
row-normalised weights), and from this project's seeded R-MAT definition.

    python tests/golden/make_golden.py

Outputs (committed): golden_small.json, rmat14.npz, ppr_rmat10.npz.
The reference itself is never imported or run.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import oracle as O  # noqa: E402

# Easy-Graph/README.md:168-173 -- eg.pagerank of this undirected graph
README_EDGES = [(1, 2), (2, 3), (1, 3), (3, 4), (4, 5), (3, 5), (5, 6)]
README_KAT = {1: 0.14272233049003707, 2: 0.14272233049003694, 3: 0.2685427766200994,
              4: 0.14336430577918527, 5: 0.21634929087322705, 6: 0.0862989657474143}

CASES = [
    # name, directed, nodes (insertion order), edges (u, v[, w]), weight, alpha, tol, max_iter
    # Easy-Graph/README.md:168-173 at the default tol
    ("readme_undirected", False, None, README_EDGES, None, 0.85, 1e-6, 200),
    # scripts/correctness_smoke.py:38-51,137 (weighted, directed, alpha 0.75)
    ("smoke_weighted_directed", True, None,
     [(0, 1, 1.0), (1, 2, 2.0), (2, 0, 0.5), (2, 3, 1.5), (3, 4, 1.0), (4, 3, 0.75), (1, 4, 4.0)],
     "weight", 0.75, 1e-6, 200),
    # easygraph/tests/test_eggpu_bulk_csr.py:12,64-68,296-310 (undirected, alpha 0.75)
    ("bulk_tiny_undirected", False, list(range(6)),
     [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3)], None, 0.75, 1e-6, 200),
    # easygraph/tests/test_eggpu_bulk_csr.py:339-378 (alpha 0.90, tol 1e-7)
    ("alpha_change_directed", True, list(range(7)),
     [(0, 1), (0, 2), (1, 2), (2, 0), (2, 3), (3, 2), (3, 4), (4, 5), (5, 4), (5, 6)], None, 0.90, 1e-7, 200),
    ("alpha_change_directed_a075", True, list(range(7)),
     [(0, 1), (0, 2), (1, 2), (2, 0), (2, 3), (3, 2), (3, 4), (4, 5), (5, 4), (5, 6)], None, 0.75, 1e-7, 200),
    # easygraph/tests/run_eggpu_gpu_contract_regression.py:83-118 (before and after the weight mutation)
    ("regression_weighted_before", True, None, [(0, 1, 8.0), (0, 2, 1.0), (1, 2, 2.0), (2, 0, 1.0)],
     "weight", 0.85, 1e-6, 200),
    ("regression_weighted_after", True, None, [(0, 1, 0.25), (0, 2, 1.0), (1, 2, 2.0), (2, 0, 1.0)],
     "weight", 0.85, 1e-6, 200),
    # easygraph/functions/centrality/tests/test_pagerank.py:19-29
    ("directed_triangle", True, None, [(0, 1), (1, 2), (2, 0)], None, 0.85, 1e-6, 200),
    ("undirected_triangle", False, None, [(0, 1), (1, 2), (2, 0)], None, 0.85, 1e-6, 200),
    ("disconnected", True, None, [(0, 1), (2, 3)], None, 0.85, 1e-6, 200),
    ("self_loop", True, None, [(0, 0), (0, 1), (1, 2)], None, 0.85, 1e-6, 200),
]


def nx_graph(directed, nodes, edges, weight):
    G = nx.DiGraph() if directed else nx.Graph()
    if nodes is not None:
        G.add_nodes_from(nodes)
    for e in edges:
        if len(e) == 3:
            G.add_edge(e[0], e[1], **{weight or "weight": e[2]})
        else:
            G.add_edge(e[0], e[1])
    return G


def nx_pagerank_with_iterations(G, alpha, weight, tol, max_iter, personalization=None):
    """networkx result and the iteration count at which it stops (the
    smallest max_iter that does not raise PowerIterationFailedConvergence)."""
    for k in range(1, max_iter + 1):
        try:
            x = nx.pagerank(G, alpha=alpha, weight=weight, tol=tol, max_iter=k,
                            personalization=personalization)
            return x, k
        except nx.PowerIterationFailedConvergence:
            continue
    raise RuntimeError("no convergence")


def main():
    out = {"networkx": nx.__version__, "readme_kat": {str(k): v for k, v in README_KAT.items()},
           "readme_edges": README_EDGES, "cases": []}
    for name, directed, nodes, edges, weight, alpha, tol, max_iter in CASES:
        G = nx_graph(directed, nodes, edges, weight)
        x, k = nx_pagerank_with_iterations(G, alpha, weight, tol, max_iter)
        node_list = list(G.nodes)
        out["cases"].append({
            "name": name, "directed": directed, "nodes": node_list, "edges": [list(e) for e in edges],
            "weight": weight, "alpha": alpha, "tol": tol, "max_iter": max_iter,
            "nx_values": [x[v] for v in node_list], "nx_iterations": k,
        })
    # converged fixed point of the README graph (tol -> 0)
    G = nx_graph(False, None, README_EDGES, None)
    x = nx.pagerank(G, alpha=0.85, weight=None, tol=1e-15, max_iter=10000)
    out["readme_converged_nx"] = {str(v): x[v] for v in G.nodes}
    with open(os.path.join(HERE, "golden_small.json"), "w") as fh:
        json.dump(out, fh, indent=1)

    # C1: R-MAT scale 14, edge factor 16, seed 1, unweighted, alpha 0.85, tol 1e-6, max_iter 100
    V, E = O.rmat_csr(14, 16, seed=1)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    x, k = nx_pagerank_with_iterations(G, 0.85, None, 1e-6, 100)
    vals = np.array([x[i] for i in range(n)])
    np.savez_compressed(
        os.path.join(HERE, "rmat14.npz"), nx_values=vals, nx_iterations=k, m=len(E),
        V_sha256=hashlib.sha256(V.tobytes()).hexdigest(), E_sha256=hashlib.sha256(E.tobytes()).hexdigest(),
    )

    # personalised PageRank on R-MAT scale 10, weighted U(0.1,1), 4 seed sets
    V, E = O.rmat_csr(10, 16, seed=7)
    W = O.edge_weights(V, E, wseed=11)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    for s_, d_, w_ in zip(src.tolist(), dst.tolist(), W.tolist()):
        G.add_edge(s_, d_, weight=w_)
    seed_sets = [[0], [5, 17, 300], list(range(100, 116)), [1023, 512]]
    vals, iters = [], []
    for S in seed_sets:
        p = {v: (1.0 / len(S) if v in S else 0.0) for v in range(n)}
        x, k = nx_pagerank_with_iterations(G, 0.85, "weight", 1e-6, 200, personalization=p)
        vals.append([x[i] for i in range(n)])
        iters.append(k)
    np.savez_compressed(
        os.path.join(HERE, "ppr_rmat10.npz"), V=V, E=E, W=W,
        seed_sets=np.array([s + [-1] * (16 - len(s)) for s in seed_sets], dtype=np.int32),
        nx_values=np.array(vals), nx_iterations=np.array(iters),
    )
    print("golden fixtures written")


if __name__ == "__main__":
    main()
