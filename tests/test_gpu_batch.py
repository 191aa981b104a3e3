"""GPU tests of batched / personalised PageRank (SpMM engine).

Semantics pinned by networkx 3.4.2 pagerank(personalization=...) fixtures
and the oracle's orc_ppr_push; a global column must equal plain PageRank.
A column's result must be the same alone and in a mixed batch (here:
bit-identical), also when other slots are re-assigned while it runs.
"""
import os

import numpy as np
import pytest

from eggpu_amd import native
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(f"[gpu-batch] {msg}", flush=True)


def _close(ours, ref, tag, tol=1e-6, l1=1e-5):
    e = np.abs(np.asarray(ours) - np.asarray(ref))
    _say(f"{tag}: max-abs {e.max():.3e} L1 {e.sum():.3e}")
    assert e.max() <= tol and e.sum() <= l1, tag


def _p(n, seeds):
    p = np.zeros(n)
    p[list(seeds)] = 1.0 / len(seeds)
    return p


def test_ppr_matches_networkx_fixture(gpu_ready):
    z = np.load(os.path.join(GOLDEN, "ppr_rmat10.npz"))
    V, E, W = z["V"], z["E"], z["W"]
    sets = [[int(s) for s in S if s >= 0] for S in z["seed_sets"]]
    g = native.DeviceGraph.from_csr(V, E, W)
    pr, it, _ = g.ppr(sets, alpha=0.85, max_iter=200, tol=1e-6)
    for j in range(len(sets)):
        ref = z["nx_values"][j] / z["nx_values"][j].sum()
        _close(pr[j], ref, f"nx ppr set {j}")
        assert it[j] == int(z["nx_iterations"][j])


@pytest.mark.parametrize("weighted", [False, True])
def test_mixed_columns_vs_oracle(gpu_ready, weighted):
    V, E = O.rmat_csr(12, 16, seed=31)
    W = O.edge_weights(V, E, 8) if weighted else None
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    rng = np.random.default_rng(5)
    sets, alphas = [], []
    for j in range(20):
        k = int(rng.integers(0, 17))  # 0 -> global column
        sets.append(sorted(rng.choice(n, k, replace=False).tolist()) if k else [])
        alphas.append(0.85 if j % 3 else 0.9)
    g = native.DeviceGraph.from_csr(V, E, W)
    pr, it, _ = g.ppr(sets, alphas=alphas, max_iter=200, tol=1e-6)
    for j, S in enumerate(sets):
        if S:
            ref, it_ref, _ = O.ppr_rows(n, src, dst, W, _p(n, S), alphas[j], 200, 1e-6)
        else:
            ref, it_ref, _ = O.pagerank_rows(n, src, dst, W, alphas[j], 200, 1e-6)
        _close(pr[j], ref, f"col {j} seeds={len(S)} alpha={alphas[j]}")
        assert it[j] == it_ref, f"col {j}"


def test_full_batch_of_64_fixed_iterations(gpu_ready):
    V, E = O.rmat_csr(13, 16, seed=2)
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    rng = np.random.default_rng(9)
    sets = [sorted(rng.choice(n, int(rng.integers(1, 17)), replace=False).tolist()) for _ in range(64)]
    g = native.DeviceGraph.from_csr(V, E)
    pr, it, _ = g.ppr(sets, max_iter=30, tol=0.0)
    assert (it == 30).all()
    for j in (0, 17, 40, 63):
        ref, _, _ = O.ppr_rows(n, src, dst, None, _p(n, sets[j]), 0.85, 30, 0.0)
        _close(pr[j], ref, f"64-batch col {j}")


def test_column_alone_equals_column_in_mix(gpu_ready):
    V, E = O.rmat_csr(12, 16, seed=77)
    W = O.edge_weights(V, E, 3)
    n = len(V) - 1
    g = native.DeviceGraph.from_csr(V, E, W)
    target = [5, 99, 1000]
    alone, it_alone, _ = g.ppr([target], alpha=0.9, max_iter=100, tol=1e-7)
    rng = np.random.default_rng(1)
    others = [sorted(rng.choice(n, 4, replace=False).tolist()) for _ in range(40)]
    mix_sets = others[:23] + [target] + others[23:]
    alphas = [0.85] * 23 + [0.9] + [0.85] * 17
    mixed, it_mix, _ = g.ppr(mix_sets, alphas=alphas, max_iter=100, tol=1e-7)
    assert it_mix[23] == it_alone[0]
    assert np.array_equal(mixed[23], alone[0])


def test_continuous_reassignment(gpu_ready):
    """Slots re-assigned while others run give the alone results."""
    V, E = O.rmat_csr(12, 16, seed=4)
    n = len(V) - 1
    g = native.DeviceGraph.from_csr(V, E)
    A, B, C = [3], [10, 20, 30], [2048, 7]
    ref = {}
    for name, S in (("A", A), ("B", B), ("C", C)):
        out, it, _ = g.ppr([S], max_iter=200, tol=1e-6)
        ref[name] = (out[0], it[0])
    with g.batch() as b:
        b.assign(0, A)
        b.assign(5, B)
        b.step(3)
        st = b.poll()
        assert st["iterations"][0] == 3 and st["iterations"][5] == 3
        b.assign(5, C)  # slot 5 restarts with a new request mid-flight
        b.run(max_launches=300)
        st = b.poll()
        assert st["done"][0] and st["done"][5]
        assert st["iterations"][0] == ref["A"][1] and st["iterations"][5] == ref["C"][1]
        assert np.array_equal(b.result(0), ref["A"][0])
        assert np.array_equal(b.result(5), ref["C"][0])
        b.release(5)
        b.assign(9, B)  # widens the X rows from 8 to 16 columns while slot 0 holds a result
        b.run(max_launches=300)
        assert np.array_equal(b.result(9), ref["B"][0])
        assert np.array_equal(b.result(0), ref["A"][0])
        b.release(9)
        b.assign(40, C)  # 64 columns
        b.run(max_launches=300)
        assert np.array_equal(b.result(40), ref["C"][0])
        assert np.array_equal(b.result(0), ref["A"][0])
        b.release(40)
        b.step(1)  # narrows back to 8 columns; slot 0 keeps its result
        assert np.array_equal(b.result(0), ref["A"][0])


@pytest.mark.parametrize("slot", [0, 12, 31, 63])
def test_column_bits_do_not_depend_on_width(gpu_ready, slot):
    """One request alone at slot `slot` (X rows 8, 16, 32 or 64 columns wide)
    gives the same bits as at slot 0."""
    V, E = O.rmat_csr(12, 16, seed=21)
    W = O.edge_weights(V, E, 2)
    g = native.DeviceGraph.from_csr(V, E, W)
    with g.batch() as b:
        b.assign(0, [7, 70, 700], alpha=0.9, max_iter=60, tol=1e-7)
        b.run(max_launches=100)
        ref = b.result(0)
        it_ref = b.poll()["iterations"][0]
    with g.batch() as b:
        b.assign(slot, [7, 70, 700], alpha=0.9, max_iter=60, tol=1e-7)
        b.run(max_launches=100)
        assert b.poll()["iterations"][slot] == it_ref
        assert np.array_equal(b.result(slot), ref)


def test_one_shot_ppr_reuses_its_session(gpu_ready):
    """egx_graph_ppr keeps its batch session on the handle: calls with 64,
    3 and 10 columns in a row (the later ones narrowing the X layout and
    leaving the high slots behind) give the bits of the same calls on
    fresh handles; a re-bin closes the kept session, and the next call
    opens a new one."""
    V, E = O.rmat_csr(12, 16, seed=23)
    W = O.edge_weights(V, E, 4)
    n = len(V) - 1
    rng = np.random.default_rng(5)
    sets = [sorted(rng.choice(n, int(rng.integers(1, 9)), replace=False).tolist()) for _ in range(64)]
    sets[5] = []
    g = native.DeviceGraph.from_csr(V, E, W)
    for cols, mi, tol in ((sets, 40, 0.0), (sets[:3], 60, 1e-7), (sets[10:20], 40, 1e-6)):
        out, it, _ = g.ppr(cols, alpha=0.85, max_iter=mi, tol=tol)
        fresh = native.DeviceGraph.from_csr(V, E, W)
        ref, it_ref, _ = fresh.ppr(cols, alpha=0.85, max_iter=mi, tol=tol)
        fresh.close()
        assert np.array_equal(it, it_ref) and np.array_equal(out, ref), len(cols)
    g.set_bins(tile_sort=1)
    out, it, _ = g.ppr(sets[:4], alpha=0.9, max_iter=50, tol=1e-6)
    for j in range(4):
        ref, it_ref, _ = O.ppr_rows(n, *O.csr_rows(V, E), W, _p(n, sets[j]), 0.9, 50, 1e-6)
        assert it[j] == it_ref
        _close(out[j], ref, f"one-shot ppr after a re-bin, column {j}")
    g.close()


def test_graph_replay_equals_eager_launches(gpu_ready, monkeypatch):
    """Steps replayed from a captured hipGraph give the bits of eager launches,
    replay after replay, with host work, re-assignment and a width change
    between the replays."""
    V, E = O.rmat_csr(12, 16, seed=17)
    W = O.edge_weights(V, E, 9)
    g = native.DeviceGraph.from_csr(V, E, W)
    runs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("EGGPU_AMD_BATCH_GRAPHS", mode)
        out = []
        with g.batch() as b:
            b.assign(0, [1, 2, 3], max_iter=40, tol=0.0)
            b.assign(3, [], alpha=0.9, max_iter=40, tol=0.0)
            for _ in range(3):
                b.step(4)
                b.poll()
            b.assign(12, [77], max_iter=40, tol=0.0)  # widens to 16 columns: a new capture
            b.run(max_launches=60, check_every=4)
            out = [b.result(0), b.result(3), b.result(12)]
            st = b.poll()
            assert st["iterations"][0] == st["iterations"][3] == st["iterations"][12] == 40
        runs[mode] = out
    for a, c in zip(runs["0"], runs["1"]):
        assert np.array_equal(a, c)


def test_global_column_equals_plain_pagerank(gpu_ready):
    V, E = O.rmat_csr(14, 16, seed=1)
    g = native.DeviceGraph.from_csr(V, E)
    pr, it, _ = g.ppr([[]], max_iter=100, tol=1e-6)
    plain, it_plain, _ = g.pagerank(0.85, 100, 1e-6)
    assert it[0] == it_plain == 5
    _close(pr[0], plain, "global column vs plain", tol=1e-9, l1=1e-7)


def test_scale20_ppr_properties(gpu_ready):
    _say("scale 20 weighted ppr, 64 columns")
    g = native.DeviceGraph.rmat(20, 16, seed=3, weighted=True, wseed=4)
    rng = np.random.default_rng(2)
    n = g.n
    sets = [sorted(rng.choice(n, int(rng.integers(1, 17)), replace=False).tolist()) for _ in range(64)]
    pr, it, _ = g.ppr(sets, max_iter=50, tol=0.0)
    assert (it == 50).all()
    assert np.abs(pr.sum(axis=1) - 1.0).max() < 1e-9
    assert (pr >= 0).all()


def test_single_vector_ppr_matches_networkx_fixture(gpu_ready):
    # the plain-SpMV form of one personalised vector (egx_graph_ppr_single)
    z = np.load(os.path.join(GOLDEN, "ppr_rmat10.npz"))
    V, E, W = z["V"], z["E"], z["W"]
    sets = [[int(s) for s in S if s >= 0] for S in z["seed_sets"]]
    g = native.DeviceGraph.from_csr(V, E, W)
    for j, S in enumerate(sets):
        out, it, _ = g.ppr_single(S, alpha=0.85, max_iter=200, tol=1e-6)
        ref = z["nx_values"][j] / z["nx_values"][j].sum()
        _close(out, ref, f"single-vector nx ppr set {j}")
        assert it == int(z["nx_iterations"][j])
    g.close()


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("precision", [native.PRECISION_FP32, native.PRECISION_FP64])
def test_single_vector_ppr_vs_oracle_and_batch(gpu_ready, weighted, precision):
    V, E = O.rmat_csr(12, 16, seed=33)
    W = O.edge_weights(V, E, 9) if weighted else None
    n = len(V) - 1
    src, dst = O.csr_rows(V, E)
    g = native.DeviceGraph.from_csr(V, E, W)
    rng = np.random.default_rng(6)
    for k in (1, 3, 16):
        S = sorted(rng.choice(n, k, replace=False).tolist())
        for alpha in (0.85, 0.9):
            out, it, _ = g.ppr_single(S + S[:1], alpha=alpha, tol=1e-6, precision=precision)  # a repeated seed counts once
            ref, it_ref, _ = O.ppr_rows(n, src, dst, W, _p(n, S), alpha, 200, 1e-6)
            _close(out, ref, f"single ppr k={k} alpha={alpha} weighted={weighted} fp64={precision}")
            assert it == it_ref
            col, it_col, _ = g.ppr([S], alpha=alpha, tol=1e-6)
            _close(out, col[0], f"single vs batch column k={k}")
            assert it == int(it_col[0])
    g.close()
