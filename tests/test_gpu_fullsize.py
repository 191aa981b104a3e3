"""Full-size parity: the batched, multi-graph and served workloads at their
configured sizes against the FP64 oracle, with the error of each printed.

* batched personalised PageRank: directed R-MAT scale 22 with U(0.1, 1)
  FP32 weights, 64 personalisation vectors of 1-16 seeds in one batch; a
  sample of the columns against orc_ppr_push, at fixed K=20 and at tol 1e-6
  (where the reference's err < tol*n rule stops after one iteration);
* many graphs in shared launches: a sample of the real 512-graph
  heterogeneous set (its largest, middle and smallest members and others),
  solved together; members checked against orc_pagerank_push;
* serving: graphs of the real 32-graph pool (weighted and unweighted), a
  70/30 global/personalised request mix at alpha 0.85 and 0.9 through the
  server; every result against the oracle with the same iteration count.

Bars as everywhere: max-abs 1e-6 per node, L1 1e-5, identical iterations.
The oracle rebuilds every graph from its seeds (the generators are
bit-exact, tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

from eggpu_amd import native, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MAX_ABS, MAX_L1 = 1e-6, 1e-5


def _say(msg):
    print(f"[gpu-fullsize] {msg}", flush=True)


def _close(ours, ref, tag):
    e = np.abs(np.asarray(ours) - np.asarray(ref))
    _say(f"{tag}: max-abs {e.max():.3e} L1 {e.sum():.3e}")
    assert e.max() <= MAX_ABS and e.sum() <= MAX_L1, tag


def _p(n, seeds):
    p = np.zeros(n)
    p[list(seeds)] = 1.0 / len(seeds)
    return p


def _spec_csr(spec):
    if spec.kind == "er":
        V, E = O.er_csr(spec.n, spec.nedges, spec.seed)
    else:
        a, b, c = workloads.RMAT_ABC
        V, E = O.rmat_n_csr(spec.n, spec.nedges, spec.seed, a, b, c)
    W = O.edge_weights(V, E, spec.wseed) if spec.weighted else None
    return V, E, W


def test_batched_ppr_rmat22_64_columns(gpu_ready):
    g = native.DeviceGraph.rmat(22, 16, seed=1, weighted=True, wlo=0.1, whi=1.0, wseed=101)
    n = g.n
    V, E = O.rmat_csr(22, 16, seed=1)
    W = O.edge_weights(V, E, 101)
    assert n == len(V) - 1 and int(g.info()["m"]) == len(E)
    rng = np.random.default_rng(1)
    sets = [sorted(rng.choice(n, int(rng.integers(1, 17)), replace=False).tolist()) for _ in range(64)]
    _say(f"R-MAT 22 weighted: n={n} m={len(E)}, 64 columns")
    src, dst = O.csr_rows(V, E)
    pr, it, _ = g.ppr(sets, alpha=0.85, max_iter=20, tol=0.0)
    assert (it == 20).all()
    for j in (0, 31, 63):
        ref, it_ref, _ = O.ppr_rows(n, src, dst, W, _p(n, sets[j]), 0.85, 20, 0.0)
        assert it_ref == 20
        _close(pr[j], ref, f"C3 column {j} ({len(sets[j])} seeds) K=20")
    pr, it, _ = g.ppr(sets, alpha=0.85, max_iter=100, tol=1e-6)
    for j in (5, 40):
        ref, it_ref, _ = O.ppr_rows(n, src, dst, W, _p(n, sets[j]), 0.85, 100, 1e-6)
        assert it[j] == it_ref, (j, it[j], it_ref)
        _close(pr[j], ref, f"C3 column {j} tol 1e-6 ({it_ref} iterations)")
    g.close()


def test_hetero512_sample_vs_oracle(gpu_ready):
    specs = workloads.hetero_specs(512, seed=2024)
    by_size = sorted(range(len(specs)), key=lambda i: specs[i].nedges)
    rng = np.random.default_rng(3)
    pick = sorted(set(by_size[-3:] + by_size[250:253] + by_size[:3] +
                      rng.choice(len(specs), 23, replace=False).tolist()))
    graphs = [specs[i].build(native) for i in pick]
    arcs = sum(int(g.info()["m"]) for g in graphs)
    _say(f"{len(pick)} of the 512 graphs, {arcs} arcs, largest n={max(specs[i].n for i in pick)}")
    outs, iters, _ = native.multi_pagerank(graphs, max_iter=20, tol=0.0)
    outs_tol, iters_tol, _ = native.multi_pagerank(graphs, max_iter=100, tol=1e-6)
    assert (iters == 20).all()
    checked = [pick.index(i) for i in (by_size[-1], by_size[-2], by_size[251], by_size[0])]
    for k in checked:
        V, E, W = _spec_csr(specs[pick[k]])
        assert graphs[k].n == len(V) - 1 and int(graphs[k].info()["m"]) == len(E)
        ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 20, 0.0)
        _close(outs[k], ref, f"C4 graph {pick[k]} ({specs[pick[k]].kind}, n={len(V) - 1}, m={len(E)}) K=20")
        ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 100, 1e-6)
        assert iters_tol[k] == it_ref
        _close(outs_tol[k], ref, f"C4 graph {pick[k]} tol 1e-6 ({it_ref} iterations)")
    for g in graphs:
        g.close()


def test_serving_pool_sample_vs_oracle(gpu_ready):
    specs = workloads.serving_pool_specs(32, seed=77)
    # the pool's graphs up to 1.2e7 draws, weighted and unweighted
    sample = [i for i in range(len(specs)) if specs[i].nedges <= 12_000_000][:6]
    assert any(specs[i].weighted for i in sample) and not all(specs[i].weighted for i in sample)
    graphs = [specs[i].build(native) for i in sample]
    csr = [_spec_csr(specs[i]) for i in sample]
    _say("pool sample: " + ", ".join(f"{i}({'w' if specs[i].weighted else 'u'}, m={len(c[1])})"
                                     for i, c in zip(sample, csr)))
    rng = np.random.default_rng(9)
    reqs = []
    for k in range(20):
        gi = k % len(graphs)
        n = graphs[gi].n
        seeds = () if rng.random() < 0.7 else tuple(sorted(rng.choice(n, int(rng.integers(1, 17)),
                                                                      replace=False).tolist()))
        reqs.append((gi, seeds, 0.85 if rng.random() < 0.5 else 0.9))
    with native.Server(graphs) as srv:
        ids = [srv.submit(gi, s, alpha=a, max_iter=100, tol=1e-6) for gi, s, a in reqs]
        results = [srv.wait(rid) for rid in ids]
        st = srv.stats()
    assert st["failed"] == 0
    for (gi, seeds, a), (out, info) in zip(reqs, results):
        V, E, W = csr[gi]
        n = len(V) - 1
        src, dst = O.csr_rows(V, E)
        if seeds:
            ref, it, _ = O.ppr_rows(n, src, dst, W, _p(n, seeds), a, 100, 1e-6)
        else:
            ref, it, _ = O.pagerank_rows(n, src, dst, W, a, 100, 1e-6)
        assert info["iterations"] == it
        _close(out, ref, f"C5 pool graph {sample[gi]} seeds={len(seeds)} alpha={a} ({it} iterations)")
    for g in graphs:
        g.close()
