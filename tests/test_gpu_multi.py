"""GPU tests of many graphs solved in shared launches (egx_multi_pagerank)
and of hipGraph replay.

A graph's result must be the same alone and in the set: the multi-graph
kernels run the same block code over a table of per-graph arguments, so the
bar is bit-identity with the one-graph path (same iteration count too), and
the one-graph path is held to the oracle within 1e-6 max-abs elsewhere.
"""
import numpy as np
import pytest

from eggpu_amd import native, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-multi] {msg}", flush=True)


def _spec_csr(spec):
    """(V, E, W or None) of a graph spec, from the CPU oracle's generators."""
    if spec.kind == "er":
        V, E = O.er_csr(spec.n, spec.nedges, spec.seed)
    else:
        a, b, c = workloads.RMAT_ABC
        V, E = O.rmat_n_csr(spec.n, spec.nedges, spec.seed, a, b, c)
    W = O.edge_weights(V, E, spec.wseed) if spec.weighted else None
    return V, E, W


def _small_set(count, seed, weighted=False):
    specs = workloads.hetero_specs(count, seed=seed, n_lo=300, n_hi=60000)
    if weighted:
        specs = [workloads.GraphSpec(s.kind, s.n, s.nedges, s.seed, True, s.seed + 5) for s in specs]
    return specs


def test_rmat_n_generator_bit_exact(gpu_ready):
    for n, m, seed in ((1000, 9000, 3), (12345, 200000, 8), (4096, 65536, 1)):
        _say(f"rmat_n n={n} draws={m}")
        g = native.DeviceGraph.rmat_n(n, m, seed=seed, keep_csr=True)
        V, E, _ = g.download_csr()
        Vo, Eo = O.rmat_n_csr(n, m, seed)
        assert np.array_equal(V, Vo) and np.array_equal(E, Eo)
        g.close()
    # a power of two is the scale form
    g = native.DeviceGraph.rmat_n(1 << 10, 16 << 10, seed=5, keep_csr=True)
    V, E, _ = g.download_csr()
    Vs, Es = O.rmat_csr(10, 16, seed=5)
    assert np.array_equal(V, Vs) and np.array_equal(E, Es)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("precision", [native.PRECISION_FP32, native.PRECISION_FP64])
def test_multi_equals_single_bitwise(gpu_ready, weighted, precision):
    specs = _small_set(12, seed=11 + int(weighted), weighted=weighted)
    _say(f"multi vs single: {len(specs)} graphs weighted={weighted} precision={precision}")
    graphs = [s.build(native) for s in specs]
    for i, g in enumerate(graphs):  # mixed bin settings inside one launch
        g.set_bins(tile_sort=i % 2, tile_items=(2944, 5888, 4416)[i % 3])
    alone = [g.pagerank(0.85, 100, 1e-6, precision=precision) for g in graphs]
    for hip_graph, xcd in ((False, False), (True, False), (True, True)):
        outs, iters, st = native.multi_pagerank(graphs, 0.85, 100, 1e-6, precision=precision,
                                                hip_graph=hip_graph, xcd_placement=xcd)
        for i, g in enumerate(graphs):
            assert iters[i] == alone[i][1], (i, hip_graph)
            assert np.array_equal(outs[i], alone[i][0]), (i, hip_graph)
        _say(f"hip_graph={hip_graph}: iterations {iters.min()}..{iters.max()}, launches {st['spmv_launches']}")
    for g in graphs:
        g.close()


def test_multi_vs_oracle(gpu_ready):
    specs = _small_set(6, seed=5)
    graphs = [s.build(native) for s in specs]
    outs, iters, _ = native.multi_pagerank(graphs, 0.9, 100, 1e-6)
    for i, s in enumerate(specs):
        V, E, W = _spec_csr(s)
        ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.9, 100, 1e-6)
        err = np.abs(outs[i] - ref)
        _say(f"{s.kind} n={s.n} m={len(E)}: max-abs {err.max():.3e} iterations {iters[i]} / {it_ref}")
        assert err.max() <= 1e-6 and err.sum() <= 1e-5
        assert iters[i] == it_ref


def test_multi_fixed_iterations_and_empty_graph(gpu_ready):
    specs = _small_set(5, seed=9)
    graphs = [s.build(native) for s in specs]
    empty = native.DeviceGraph.from_csr(np.zeros(8, dtype=np.int32), np.zeros(0, dtype=np.int32))
    outs, iters, st = native.multi_pagerank(graphs + [empty], 0.85, 17, 0.0, check_every=4)
    assert list(iters[:5]) == [17] * 5 and iters[5] == 0
    assert np.allclose(outs[5], 1.0 / 7)
    for i, g in enumerate(graphs):
        ref, it, _ = g.pagerank(0.85, 17, 0.0)
        assert it == 17 and np.array_equal(outs[i], ref)


@pytest.mark.parametrize("shared", [0, 1])
@pytest.mark.parametrize("weighted", [False, True])
def test_multi_blocked_members_equal_single(gpu_ready, weighted, shared):
    """Members with the propagation-blocked layout run its scatter and sum
    inside the shared launches (table-driven), the others their pull work
    items: every member bit-identical to its solo solve, with the solo
    blocked solves against the oracle."""
    specs = _small_set(10, seed=21 + int(weighted), weighted=weighted)
    graphs = [s.build(native) for s in specs]
    blocked = [i for i in range(len(graphs)) if i % 3 != 1]
    for i in blocked:
        graphs[i].set_bins(blocked=1, blocked_shared=shared)
    assert all(bool(graphs[i].info()["blocked"]) == (i in blocked) for i in range(len(graphs)))
    _say(f"multi with blocked members {blocked} of {len(graphs)} weighted={weighted} shared={shared}")
    for K, tol in ((20, 0.0), (100, 1e-6)):
        alone = [g.pagerank(0.85, K, tol) for g in graphs]
        for hip_graph in (False, True):
            outs, iters, st = native.multi_pagerank(graphs, 0.85, K, tol, hip_graph=hip_graph)
            for i in range(len(graphs)):
                assert iters[i] == alone[i][1], (i, K, hip_graph)
                assert np.array_equal(outs[i], alone[i][0]), (i, K, hip_graph)
    for i in blocked[:3]:
        V, E, W = _spec_csr(specs[i])
        ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 100, 1e-6)
        pr, it, _ = graphs[i].pagerank(0.85, 100, 1e-6)
        assert it == it_ref
        d = np.abs(pr - ref)
        _say(f"blocked member {i} (n={len(V) - 1}, m={len(E)}) vs oracle: max-abs {d.max():.3e} L1 {d.sum():.3e}")
        assert d.max() <= 1e-6 and d.sum() <= 1e-5
    for g in graphs:
        g.close()


def test_multi_rejects_mixed_weighting(gpu_ready):
    a = native.DeviceGraph.er(1000, 8000, seed=1)
    b = native.DeviceGraph.er(1000, 8000, seed=2, weighted=True, wseed=3)
    with pytest.raises(native.EgxError):
        native.multi_pagerank([a, b])
    with pytest.raises(native.EgxError):
        native.multi_pagerank([a, a])


def test_single_graph_hip_graph_replay(gpu_ready):
    g = native.DeviceGraph.rmat(14, 16, seed=1)
    ref, it, _ = g.pagerank(0.85, 100, 1e-6)
    for ce in (1, 2, 3):
        out, st = g.pagerank_stats(0.85, 100, 1e-6, time_each_launch=False, download=True, hip_graph=True,
                                   check_every=ce)
        assert st["iterations"] == it == 5
        assert np.array_equal(out, ref), ce
    # the same graph object replays again after another solve
    out2, st2 = g.pagerank_stats(0.9, 100, 1e-7, time_each_launch=False, download=True, hip_graph=True,
                                 check_every=2)
    ref2, it2, _ = g.pagerank(0.9, 100, 1e-7)
    assert st2["iterations"] == it2 and np.array_equal(out2, ref2)
