"""GPU tests of the propagation-blocked iteration (egx_bin_opts.blocked):
source chunks staged in LDS scatter their terms into destination-binned
value slots, and every sub-bin sums its slots in LDS.

Parity bar as for the pull SpMV: FP32 iterates against the FP64 oracle at
max-abs 1e-6 and L1 1e-5 per vector with identical iteration counts.
Against the pull path on the same graph the results differ only by the
FP32 rounding of each stretch sum (the scatter stores one FP32 term per
stretch of equal rows in an 8-slot group; the pull kernel sums every row in
FP64): held to max-abs 1e-7.  Sub-bins too large for one wavefront are cut
into pieces summed by the reduce launch; a graph with a 100,000-arc row
exercises that path.  Two solves must agree bit for bit.
"""
import numpy as np
import pytest

from eggpu_amd import native
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_ABS = 1e-6
MAX_L1 = 1e-5


def _say(msg):
    print(f"[gpu-blocked] {msg}", flush=True)


def _assert_close(ours, ref, tag, tol=MAX_ABS, l1=MAX_L1):
    d = np.abs(np.asarray(ours) - np.asarray(ref))
    _say(f"{tag}: max-abs {d.max():.3e} L1 {d.sum():.3e}")
    assert d.max() <= tol and d.sum() <= l1, tag


def _in_star_graph(seed=5):
    # R-MAT scale 17 plus arcs u -> 7 from 100,000 sources: row 7 alone
    # holds more slots than one sum item takes
    V, E = O.rmat_csr(17, 16, seed=seed)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    extra = np.arange(1000, 101000, dtype=np.int64)
    return O.simple_csr_from_coo(n, np.concatenate([src, extra]), np.concatenate([dst, np.full(len(extra), 7)]))


CASES = {
    "rmat10": lambda: O.rmat_csr(10, 16, seed=10),
    "rmat14": lambda: O.rmat_csr(14, 16, seed=14),
    "rmat17": lambda: O.rmat_csr(17, 16, seed=17),
    "in_star": _in_star_graph,
    "er": lambda: O.er_csr(50000, 400000, seed=3),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tol,max_iter", [(1e-6, 100), (0.0, 30)])
def test_blocked_matches_oracle_and_pull(gpu_ready, case, tol, max_iter):
    V, E = CASES[case]()
    _say(f"{case}: n={len(V) - 1} m={len(E)} tol={tol} K={max_iter}")
    g = native.DeviceGraph.from_csr(V, E)
    pr0, it0, _ = g.pagerank(0.85, max_iter, tol)
    g.set_bins(blocked=1)
    assert g.info()["blocked"] == 1
    pr, it, _ = g.pagerank(0.85, max_iter, tol)
    ref, it_ref, _ = O.pagerank_csr(V, E, None, 0.85, max_iter, tol)
    _assert_close(pr, ref, f"{case} blocked vs oracle")
    assert it == it_ref
    # the pull path: same iterations, FP32 summation-order differences only
    assert it0 == it
    _assert_close(pr, pr0, f"{case} blocked vs pull", tol=1e-7, l1=1e-6)
    g.set_bins(blocked=0)
    assert g.info()["blocked"] == 0
    pr1, it1, _ = g.pagerank(0.85, max_iter, tol)
    assert it1 == it0 and np.array_equal(pr1, pr0)
    g.close()


def test_blocked_is_deterministic_and_alpha_independent_of_layout(gpu_ready):
    V, E = _in_star_graph(9)
    g = native.DeviceGraph.from_csr(V, E)
    g.set_bins(blocked=1)
    a, ia, _ = g.pagerank(0.85, 40, 0.0)
    b, ib, _ = g.pagerank(0.85, 40, 0.0)
    assert ia == ib and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    # another alpha on the same layout
    c, ic, _ = g.pagerank(0.9, 200, 1e-6)
    ref, it_ref, _ = O.pagerank_csr(V, E, None, 0.9, 200, 1e-6)
    _assert_close(c, ref, "in_star alpha 0.9")
    assert ic == it_ref
    g.close()


def test_blocked_single_vector_ppr_and_fp64(gpu_ready):
    V, E = O.rmat_csr(14, 16, seed=3)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    g = native.DeviceGraph.from_csr(V, E)
    g.set_bins(blocked=1)
    seeds = [5, 17, 123, 4000]
    p = np.zeros(n)
    p[seeds] = 1.0 / len(seeds)
    out, it, _ = g.ppr_single(seeds, alpha=0.85, max_iter=200, tol=1e-6)
    ref, it_ref = O.ppr_rows(n, src, dst, None, p, 0.85, 200, 1e-6)[:2]
    _assert_close(out, ref, "blocked ppr_single")
    assert it == it_ref
    # FP64 iterates run the pull kernels on the same graph
    pr64, it64, _ = g.pagerank(0.85, 20, 0.0, precision=native.PRECISION_FP64)
    ref64, _, _ = O.pagerank_csr(V, E, None, 0.85, 20, 0.0)
    _assert_close(pr64, ref64, "fp64 beside a blocked layout", tol=1e-12, l1=1e-10)
    g.close()


WEIGHTED = {
    "rmat14w": lambda: O.rmat_csr(14, 16, seed=41),
    "in_star_w": lambda: _in_star_graph(43),
    "er_w": lambda: O.er_csr(50000, 400000, seed=44),
}


@pytest.mark.parametrize("case", sorted(WEIGHTED))
def test_blocked_weighted_matches_oracle_and_pull(gpu_ready, case):
    # weighted graphs: each source slot carries its FP32 arc value
    # fl32(alpha * w / outW); terms are FP32 products as in the pull kernel
    V, E = WEIGHTED[case]()
    W = O.edge_weights(V, E, 7)
    _say(f"{case}: n={len(V) - 1} m={len(E)} weighted")
    g = native.DeviceGraph.from_csr(V, E, W)
    for tol, K in ((1e-6, 100), (0.0, 30)):
        g.set_bins(blocked=0)
        pr0, it0, _ = g.pagerank(0.85, K, tol)
        g.set_bins(blocked=1)
        assert g.info()["blocked"] == 1
        pr, it, _ = g.pagerank(0.85, K, tol)
        ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, K, tol)
        _assert_close(pr, ref, f"{case} blocked weighted vs oracle tol={tol}")
        assert it == it_ref == it0
        _assert_close(pr, pr0, f"{case} blocked weighted vs pull", tol=1e-7, l1=1e-6)
    # another alpha refreshes the per-slot values; then back again
    for alpha in (0.9, 0.85):
        pr, it, _ = g.pagerank(alpha, 200, 1e-6)
        ref, it_ref, _ = O.pagerank_csr(V, E, W, alpha, 200, 1e-6)
        _assert_close(pr, ref, f"{case} blocked weighted alpha={alpha}")
        assert it == it_ref
    g.close()


def test_blocked_weighted_ppr_and_determinism(gpu_ready):
    V, E = O.rmat_csr(15, 16, seed=45)
    W = O.edge_weights(V, E, 8)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    g = native.DeviceGraph.from_csr(V, E, W)
    g.set_bins(blocked=1)
    a, ia, _ = g.pagerank(0.85, 25, 0.0)
    b, ib, _ = g.pagerank(0.85, 25, 0.0)
    assert ia == ib and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    seeds = [1, 99, 2048]
    p = np.zeros(n)
    p[seeds] = 1.0 / len(seeds)
    out, it, _ = g.ppr_single(seeds, alpha=0.85, max_iter=200, tol=1e-6)
    ref, it_ref = O.ppr_rows(n, src, dst, W, p, 0.85, 200, 1e-6)[:2]
    _assert_close(out, ref, "blocked weighted ppr_single")
    assert it == it_ref
    g.close()


def test_blocked_matches_pull_at_scale_22(gpu_ready):
    # R-MAT 22 built on the device (too large for the oracle in seconds):
    # the blocked iteration must agree with the pull SpMV to FP32 stretch
    # rounding, at fixed K and at tol 1e-6 with the same iteration count,
    # repeat bit for bit, and sum to 1
    g = native.DeviceGraph.rmat(22, 16, seed=5)
    _say(f"rmat22: n={g.info()['n']} m={g.info()['m']}")
    # 65M arcs: blocked by default; the pull layout on request
    assert g.info()["blocked"] == 1
    g.set_bins(blocked=0)
    assert g.info()["blocked"] == 0
    a, ia, _ = g.pagerank(0.85, 10, 0.0)
    t, it_t, _ = g.pagerank(0.85, 100, 1e-6)
    g.set_bins(blocked=1)
    assert g.info()["blocked"] == 1
    b, ib, _ = g.pagerank(0.85, 10, 0.0)
    c, ic, _ = g.pagerank(0.85, 10, 0.0)
    u, iu, _ = g.pagerank(0.85, 100, 1e-6)
    assert ia == ib == ic and np.array_equal(b.view(np.uint64), c.view(np.uint64))
    _assert_close(b, a, "rmat22 blocked vs pull K=10", tol=1e-9, l1=1e-7)
    assert it_t == iu
    _assert_close(u, t, "rmat22 blocked vs pull tol 1e-6", tol=1e-9, l1=1e-7)
    assert abs(b.sum() - 1.0) < 1e-9
    g.close()


def test_blocked_default_policy(gpu_ready):
    # graphs of at least 2^23 arcs start blocked, weighted or not; smaller
    # ones start on the pull layout; set_bins overrides either way
    small = native.DeviceGraph.rmat(18, 16, seed=2)
    assert small.info()["m"] < 1 << 23 and small.info()["blocked"] == 0
    small.close()
    big = native.DeviceGraph.rmat(20, 16, seed=2)
    assert big.info()["m"] >= 1 << 23 and big.info()["blocked"] == 1
    a, ia, _ = big.pagerank(0.85, 100, 1e-6)
    big.set_bins(blocked=0)
    b, ib, _ = big.pagerank(0.85, 100, 1e-6)
    assert ia == ib
    _assert_close(a, b, "rmat20 default (blocked) vs pull", tol=1e-9, l1=1e-7)
    big.close()
    w = native.DeviceGraph.rmat(20, 16, seed=2, weighted=True, wseed=3)
    assert w.info()["blocked"] == 1
    a, ia, _ = w.pagerank(0.85, 100, 1e-6)
    w.set_bins(blocked=0)
    b, ib, _ = w.pagerank(0.85, 100, 1e-6)
    assert ia == ib
    _assert_close(a, b, "weighted rmat20 default (blocked) vs pull", tol=1e-9, l1=1e-7)
    w.close()


def test_blocked_layout_allocation_failure_falls_back(gpu_ready, monkeypatch):
    # a real device allocation failure inside the blocked-layout build (the
    # test hook asks hipMalloc for 4 PB): a graph that would start blocked
    # keeps the pull layout and solves correctly, with no stale HIP error
    # reaching its first launch; an explicit set_bins(blocked=1) reports the
    # failure and leaves the graph usable; without the hook it builds
    monkeypatch.setenv("EGGPU_AMD_TEST_FAIL_BLOCKED_ALLOC", "1")
    g = native.DeviceGraph.rmat(20, 16, seed=2)
    assert g.info()["m"] >= 1 << 23 and g.info()["blocked"] == 0
    a, ia, _ = g.pagerank(0.85, 100, 1e-6)
    with pytest.raises(native.EgxError) as ei:
        g.set_bins(blocked=1)
    assert ei.value.status == native.EGX_FAILED_TO_ALLOCATE_DEVICE_MEM
    assert g.info()["blocked"] == 0
    b, ib, _ = g.pagerank(0.85, 100, 1e-6)
    assert ia == ib and np.array_equal(a, b)
    monkeypatch.delenv("EGGPU_AMD_TEST_FAIL_BLOCKED_ALLOC")
    g.set_bins(blocked=1)
    assert g.info()["blocked"] == 1
    c, ic, _ = g.pagerank(0.85, 100, 1e-6)
    assert ic == ia
    _assert_close(c, a, "rmat20 blocked after the failed attempt vs pull", tol=1e-9, l1=1e-7)
    g.close()


def test_blocked_graph_in_multi_graph_and_batch_paths(gpu_ready):
    # the multi-graph launch and the SpMM batch sessions keep their own
    # kernels on a graph whose plain solves run blocked; all agree with the
    # oracle, and the plain solve still runs blocked afterwards
    V1, E1 = O.rmat_csr(12, 16, seed=31)
    V2, E2 = O.er_csr(3000, 30000, seed=32)
    g1 = native.DeviceGraph.from_csr(V1, E1)
    g2 = native.DeviceGraph.from_csr(V2, E2)
    g1.set_bins(blocked=1)
    outs, iters, _ = native.multi_pagerank([g1, g2], 0.85, 100, 1e-6)
    for (V, E), out, it in zip(((V1, E1), (V2, E2)), outs, iters):
        ref, it_ref, _ = O.pagerank_csr(V, E, None, 0.85, 100, 1e-6)
        _assert_close(out, ref, "multi-graph beside a blocked layout")
        assert it == it_ref
    n = len(V1) - 1
    src, dst = O.csr_rows(V1, E1)
    sets = [[], [3, 17, 900]]
    cols, its, _ = g1.ppr(sets, alpha=0.85, max_iter=100, tol=1e-6)
    for j, S in enumerate(sets):
        if S:
            p = np.zeros(n)
            p[S] = 1.0 / len(S)
            ref, it_ref = O.ppr_rows(n, src, dst, None, p, 0.85, 100, 1e-6)[:2]
        else:
            ref, it_ref, _ = O.pagerank_csr(V1, E1, None, 0.85, 100, 1e-6)
        _assert_close(cols[j], ref, f"batch column {j} beside a blocked layout")
        assert its[j] == it_ref
    assert g1.info()["blocked"] == 1
    pr, it, _ = g1.pagerank(0.85, 100, 1e-6)
    ref, it_ref, _ = O.pagerank_csr(V1, E1, None, 0.85, 100, 1e-6)
    _assert_close(pr, ref, "blocked after the batch")
    assert it == it_ref
    g1.close()
    g2.close()


def _edge_shape(name):
    rng = np.random.default_rng(17)
    if name == "chunk_plus_one":
        # 32,769 nodes: the last chunk holds one source
        n = 32769
        src = rng.integers(0, n, 200000)
        dst = rng.integers(0, n, 200000)
    elif name == "sparse_bins":
        # arcs only into every 5th sub-bin's rows and from a few chunks
        n = 70000
        src = rng.integers(0, 2000, 150000) * 31 % n
        dst = (rng.integers(0, n // 2048, 150000) // 5 * 5) * 2048 + rng.integers(0, 2048, 150000)
        dst = np.minimum(dst, n - 1)
    elif name == "one_row":
        # every arc into row 5 (one sub-bin split into many pieces)
        n = 300000
        src = np.arange(n)
        dst = np.full(n, 5)
    elif name == "huge_n_few_arcs":
        # 3M nodes and 500 arcs: almost every (chunk, sub-bin) run is empty
        n = 3000000
        src = rng.integers(0, n, 500)
        dst = rng.integers(0, n, 500)
    else:  # self_loops: only i -> i
        n = 5000
        src = np.arange(n)
        dst = np.arange(n)
    keep = src != dst if name != "self_loops" else np.ones(len(src), dtype=bool)
    if name == "self_loops":
        V = np.arange(n + 1, dtype=np.int32)
        return V, np.arange(n, dtype=np.int32)
    return O.simple_csr_from_coo(n, src[keep], dst[keep])


@pytest.mark.parametrize("name", ["chunk_plus_one", "sparse_bins", "one_row", "self_loops", "huge_n_few_arcs"])
def test_blocked_edge_shapes(gpu_ready, name):
    V, E = _edge_shape(name)
    _say(f"{name}: n={len(V) - 1} m={len(E)}")
    g = native.DeviceGraph.from_csr(V, E)
    g.set_bins(blocked=1)
    assert g.info()["blocked"] == 1
    for tol, K in ((1e-6, 100), (0.0, 15)):
        pr, it, _ = g.pagerank(0.85, K, tol)
        ref, it_ref, _ = O.pagerank_csr(V, E, None, 0.85, K, tol)
        _assert_close(pr, ref, f"{name} tol={tol}")
        assert it == it_ref
    g.close()
