"""GPU tests of the continuous-batching server (egx_server_*).

Each request's result must be the same alone and in the mix -- here
bit-identical, since a column's arithmetic does not depend on the other
columns of its session -- and agree with the oracle (global PageRank:
pagerank.py:75-139 restated; personalised: networkx 3.4.2 semantics) within
1e-6 max-abs per node with the same iteration count.
"""
import threading
import time

import numpy as np
import pytest

from eggpu_amd import native
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-serve] {msg}", flush=True)


@pytest.fixture(scope="module")
def pool(gpu_ready):
    specs = [(11, 3, False), (12, 5, True), (10, 8, False)]
    graphs, csr = [], []
    for scale, seed, weighted in specs:
        V, E = O.rmat_csr(scale, 16, seed=seed)
        W = O.edge_weights(V, E, seed + 1) if weighted else None
        graphs.append(native.DeviceGraph.from_csr(V, E, W))
        csr.append((V, E, W))
    yield graphs, csr
    for g in graphs:
        g.close()


def _requests(graphs, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        gi = int(rng.integers(0, len(graphs)))
        n = graphs[gi].n
        if rng.random() < 0.7:
            seeds = []
        else:
            seeds = sorted(rng.choice(n, int(rng.integers(1, 17)), replace=False).tolist())
        alpha = 0.85 if rng.random() < 0.5 else 0.9
        reqs.append((gi, tuple(seeds), alpha))
    return reqs


def test_results_alone_equal_results_in_the_mix(pool):
    graphs, csr = pool
    reqs = _requests(graphs, 90, seed=3)
    distinct = sorted(set(reqs))
    alone = {}
    with native.Server(graphs) as srv:
        for r in distinct:  # one at a time: nothing else on the device
            out, info = srv.wait(srv.submit(r[0], r[1], alpha=r[2], max_iter=100, tol=1e-6))
            alone[r] = (out, info["iterations"])
    _say(f"{len(distinct)} distinct requests solved alone")
    # the same requests as a Poisson stream (mean gap 0.5 ms) against a fresh server
    rng = np.random.default_rng(11)
    gaps = rng.exponential(5e-4, size=len(reqs))
    ids = [None] * len(reqs)
    with native.Server(graphs, step_iters=2) as srv:
        def feed():
            t = time.perf_counter()
            for k, r in enumerate(reqs):
                t += gaps[k]
                while time.perf_counter() < t:
                    pass
                ids[k] = srv.submit(r[0], r[1], alpha=r[2], max_iter=100, tol=1e-6)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        lat = []
        for k, r in enumerate(reqs):
            out, info = srv.wait(ids[k])
            ref, it = alone[r]
            assert info["iterations"] == it, k
            assert np.array_equal(out, ref), k
            lat.append(info["latency_seconds"])
        st = srv.stats()
    _say(f"mix: peak columns {st['peak_active_columns']}, sessions {st['sessions_opened']}, "
         f"p50 {np.percentile(lat, 50) * 1e3:.2f} ms p99 {np.percentile(lat, 99) * 1e3:.2f} ms")
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_active_columns"] > 1  # requests did share the device


def test_packed_delivery_equals_device_output(pool):
    """The server copies a finished column packed (its FP32 iterate and FP64
    sum) and finishes the FP64 values on the host; egx_batch_result finishes
    them on the device (k_spmm_output).  Same bits, weighted and unweighted,
    global and personalised."""
    graphs, _ = pool
    reqs = [(0, (), 0.85), (1, (), 0.9), (2, (3, 17, 40), 0.85), (1, (5,), 0.85), (0, tuple(range(0, 64, 4)), 0.9)]
    with native.Server(graphs) as srv:
        served = [srv.wait(srv.submit(gi, s, alpha=a, max_iter=100, tol=1e-6)) for gi, s, a in reqs]
    for (gi, s, a), (out, info) in zip(reqs, served):
        with graphs[gi].batch() as b:
            b.assign(0, s, alpha=a, max_iter=100, tol=1e-6)
            b.run(max_launches=100)
            ref = b.result(0)
            it = int(b.poll()["iterations"][0])
        assert info["iterations"] == it, (gi, s, a)
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (gi, s, a, np.abs(out - ref).max())
    _say(f"{len(reqs)} served results equal the device-finished ones bit for bit")


def test_served_results_match_oracle(pool):
    graphs, csr = pool
    with native.Server(graphs) as srv:
        cases = [(0, (), 0.85), (1, (), 0.9), (1, (3, 50, 700), 0.85), (2, (9,), 0.9)]
        ids = [srv.submit(g, s, alpha=a, max_iter=100, tol=1e-6) for g, s, a in cases]
        for (g, seeds, a), rid in zip(cases, ids):
            out, info = srv.wait(rid)
            V, E, W = csr[g]
            n = len(V) - 1
            src, dst = O.csr_rows(V, E)
            if seeds:
                p = np.zeros(n)
                p[list(seeds)] = 1.0 / len(seeds)
                ref, it, _ = O.ppr_rows(n, src, dst, W, p, a, 100, 1e-6)
            else:
                ref, it, _ = O.pagerank_rows(n, src, dst, W, a, 100, 1e-6)
            err = np.abs(out - ref)
            _say(f"graph {g} seeds={len(seeds)} alpha={a}: max-abs {err.max():.3e}, iterations {info['iterations']}")
            assert err.max() <= 1e-6 and err.sum() <= 1e-5
            assert info["iterations"] == it


def test_budget_admission(pool):
    graphs, _ = pool
    # too small for any session: the request fails cleanly
    with native.Server(graphs, hbm_budget_bytes=1 << 16) as srv:
        rid = srv.submit(0)
        with pytest.raises(native.EgxError):
            srv.wait(rid)
    # room for one session at a time: graphs take turns, idle sessions close
    with native.Server(graphs) as srv:
        one = int(srv.stats()["largest_session_bytes"] * 1.5)
    with native.Server(graphs, hbm_budget_bytes=one) as srv:
        ids = [srv.submit(k % 3, (), max_iter=30, tol=0.0) for k in range(9)]
        for rid in ids:
            _, info = srv.wait(rid)
            assert info["iterations"] == 30
        st = srv.stats()
    _say(f"one-session budget: opened {st['sessions_opened']}, closed {st['sessions_closed']}")
    assert st["sessions_closed"] >= 2 and st["peak_session_bytes"] <= one


def test_identical_requests_share_a_column(pool):
    graphs, _ = pool
    cases = [(1, (), 0.85)] * 16 + [(0, (5, 9, 40), 0.9)] * 8
    results = {}
    for coalesce in (True, False):
        with native.Server(graphs, coalesce=coalesce) as srv:
            ids = [srv.submit(g, s, alpha=a, max_iter=100, tol=1e-6) for g, s, a in cases]
            outs = [srv.wait(rid) for rid in ids]
            st = srv.stats()
        results[coalesce] = outs
        _say(f"coalesce={coalesce}: coalesced {st['coalesced']}, completed {st['completed']}")
        assert st["completed"] == len(cases) and st["failed"] == 0
        if coalesce:
            assert st["coalesced"] >= 1  # submitted back to back: later ones ride on the first column
        else:
            assert st["coalesced"] == 0
    for (a, ia), (b, ib) in zip(results[True], results[False]):
        assert np.array_equal(a, b) and ia["iterations"] == ib["iterations"]
