"""GPU tests of eggpu-csr-v1 artifacts streamed straight into HBM
(egx_graph_from_csr_files).  The device graph must be the one the host
route builds from the same arrays (pull CSC bit-exact against the oracle's
build_inbound, which restates pagerank.cu:286-347), and the validation must
reject what eggpu_bulk_graph.py / graph_convert.cpp:492-520 reject."""
import numpy as np
import pytest

from eggpu_amd import native
from eggpu_amd.bulk import read_eggpu_csr, write_eggpu_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _files(tmp_path, V, E, W=None, name="g"):
    mp = write_eggpu_csr(tmp_path / name, V, E, W)
    base = tmp_path / name
    w = str(base) + ".weights.f64" if W is not None else None
    return str(base) + ".offsets.i32", str(base) + ".indices.i32", w, mp


@pytest.mark.parametrize("weighted", [False, True])
def test_ingest_matches_oracle_build(gpu_ready, tmp_path, weighted):
    V, E = O.rmat_csr(13, 16, seed=12)
    W = O.edge_weights(V, E, 6) if weighted else None
    off, ind, wp, _ = _files(tmp_path, V, E, W)
    g, rep = native.DeviceGraph.from_csr_files(off, ind, wp, alpha=0.85, keep_csr=True)
    print(f"[gpu-ingest] {rep}", flush=True)
    assert rep["in_range"] == 1 and rep["finite"] == 1 and rep["simple_sorted"] == 1
    assert rep["bytes"] == 4 * (len(V) + len(E)) + (8 * len(E) if weighted else 0)
    Vd, Ed, Wd = g.download_csr()
    assert np.array_equal(Vd, V) and np.array_equal(Ed, E)
    if weighted:
        assert np.array_equal(Wd, W)
    in_row, in_col, in_val, dang = g.download_csc()
    r_row, r_col, r_val, r_dang, _ = O.build_inbound(V, E, W, 0.85)
    assert np.array_equal(in_row, r_row) and np.array_equal(in_col, r_col) and np.array_equal(dang, r_dang)
    assert np.array_equal(in_val.view(np.uint32), r_val.view(np.uint32))
    pr, it, _ = g.pagerank(0.85, 100, 1e-6)
    ref, it_ref, _ = O.pagerank_csr(V, E, W, 0.85, 100, 1e-6)
    assert np.abs(pr - ref).max() <= 1e-6 and it == it_ref


def test_ingest_validation(gpu_ready, tmp_path):
    V = np.array([0, 2, 3, 5], np.int32)
    E = np.array([1, 2, 0, 0, 1], np.int32)
    # offsets that decrease are always rejected
    off, ind, _, _ = _files(tmp_path, np.array([0, 3, 2, 5], np.int32), E, name="bad_off")
    with pytest.raises(native.EgxError, match="monotonic"):
        native.DeviceGraph.from_csr_files(off, ind, validate=False)
    # an endpoint outside [0, n): rejected with validation, skipped without
    Ebad = np.array([1, 7, 0, 0, 1], np.int32)
    off, ind, _, _ = _files(tmp_path, V, Ebad, name="bad_end")
    with pytest.raises(native.EgxError, match="outside"):
        native.DeviceGraph.from_csr_files(off, ind, validate=True)
    g, rep = native.DeviceGraph.from_csr_files(off, ind, validate=False)
    assert rep["in_range"] == 0
    in_row, in_col, _, _ = g.download_csc()
    r_row, r_col, _, _, _ = O.build_inbound(V, Ebad, None, 0.85)
    assert np.array_equal(in_row, r_row) and np.array_equal(in_col, r_col)
    # non-finite weights
    off, ind, wp, _ = _files(tmp_path, V, E, np.array([1.0, np.inf, 1.0, 1.0, 1.0]), name="bad_w")
    with pytest.raises(native.EgxError, match="finite"):
        native.DeviceGraph.from_csr_files(off, ind, wp, validate=True)
    # a self-loop is allowed but the artifact is not "simple sorted"
    off, ind, _, _ = _files(tmp_path, V, np.array([1, 2, 0, 0, 2], np.int32), name="loop")
    _, rep = native.DeviceGraph.from_csr_files(off, ind)
    assert rep["simple_sorted"] == 0 and rep["in_range"] == 1


def test_bulk_graph_uses_device_ingest(gpu_ready, tmp_path, monkeypatch):
    import eggpu_amd

    monkeypatch.setenv("EASYGRAPH_ENABLE_GPU", "TRUE")
    monkeypatch.setenv("EASYGRAPH_GPU_STRICT_ERRORS", "TRUE")
    V, E = O.rmat_csr(12, 16, seed=3)
    W = O.edge_weights(V, E, 1)
    _, _, _, mp = _files(tmp_path, V, E, W)
    G = read_eggpu_csr(mp)
    res = eggpu_amd.pagerank(G, alpha=0.85, weight="weight")
    assert G.ingest_report is not None and G.ingest_report["bytes"] > 0
    ref, _, _ = O.pagerank_csr(V, E, W, 0.85, 200, 1e-6)
    got = np.array([res[i] for i in range(len(V) - 1)])
    assert np.abs(got - ref).max() <= 1e-6
