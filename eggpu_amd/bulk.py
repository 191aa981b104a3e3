"""``eggpu-csr-v1`` bulk graphs: manifest + raw int32 CSR files.

Reader semantics follow ``easygraph/classes/eggpu_bulk_graph.py:170-334``
and ``cpp_easygraph/classes/graph_convert.cpp:447-526``:

* the JSON manifest must say ``format == "eggpu-csr-v1"``, int32 offsets and
  indices, ``node_labels == "zero_based_contiguous"``; file sizes must match
  ``num_nodes`` / ``num_entries`` (float64 weights when ``weights_path`` is
  given, all finite);
* offsets must start at 0 and end at ``num_entries``; with validation on
  (``EASYGRAPH_GPU_BULK_VALIDATE``, default true) they must be monotonic and
  every endpoint must lie in ``[0, num_nodes)``;
* the graph is immutable, its nodes are ``range(num_nodes)``, and the CSR is
  used as stored (undirected artifacts already hold both directions).

The files are memory-mapped, never turned into Python adjacency; the GPU
route streams them straight into HBM (``device_graph``) and builds there.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

_FORMAT = "eggpu-csr-v1"
_INT32_LIMIT = 2**31 - 1


def _resolved(base, value):
    p = Path(value).expanduser()
    return p if p.is_absolute() else (base.parent / p).resolve()


class EGGPUBulkGraph:
    """Read-only graph with node labels ``0..N-1`` backed by CSR files."""

    cflag = 0
    _eggpu_bulk_csr = True

    def __init__(self, metadata_path, validate=None):
        metadata_path = Path(metadata_path).expanduser().resolve()
        with metadata_path.open("r", encoding="utf-8") as fh:
            meta = json.load(fh)
        if meta.get("format") != _FORMAT:
            raise ValueError(f"unsupported EGGPU CSR format {meta.get('format')!r}; expected {_FORMAT!r}")
        if meta.get("offset_dtype") != "int32" or meta.get("index_dtype") != "int32":
            raise ValueError("EGGPU bulk CSR requires int32 offsets and int32 indices")
        if meta.get("node_labels") != "zero_based_contiguous":
            raise ValueError("EGGPU bulk CSR requires zero-based contiguous node labels")
        n = int(meta["num_nodes"])
        ne = int(meta["num_entries"])
        if n < 0 or n >= _INT32_LIMIT:
            raise ValueError("num_nodes exceeds the signed 32-bit EGGPU CSR limit")
        if ne < 0 or ne >= _INT32_LIMIT:
            raise ValueError("num_entries exceeds the signed 32-bit EGGPU CSR limit")
        offsets_path = _resolved(metadata_path, meta["offsets_path"])
        indices_path = _resolved(metadata_path, meta["indices_path"])
        wv = meta.get("weights_path")
        weights_path = _resolved(metadata_path, wv) if wv else None
        weight_key = str(meta.get("weight_key", "weight"))
        if offsets_path.stat().st_size != (n + 1) * 4:
            raise ValueError(f"offset file size mismatch: expected {(n + 1) * 4}, found {offsets_path.stat().st_size}")
        if indices_path.stat().st_size != ne * 4:
            raise ValueError(f"index file size mismatch: expected {ne * 4}, found {indices_path.stat().st_size}")
        if weights_path is not None:
            if meta.get("weight_dtype") != "float64":
                raise ValueError("EGGPU bulk CSR explicit weights require float64")
            if weights_path.stat().st_size != ne * 8:
                raise ValueError(f"weight file size mismatch: expected {ne * 8}, found {weights_path.stat().st_size}")
            if not weight_key:
                raise ValueError("weight_key must be non-empty")
        if validate is None:
            validate = os.environ.get("EASYGRAPH_GPU_BULK_VALIDATE", "TRUE").strip().upper() in {
                "1", "TRUE", "ON", "YES"}

        self._V = np.memmap(offsets_path, dtype=np.int32, mode="r", shape=(n + 1,))
        self._E = np.memmap(indices_path, dtype=np.int32, mode="r", shape=(ne,)) if ne else np.zeros(0, np.int32)
        self._W = None
        if weights_path is not None:
            self._W = np.memmap(weights_path, dtype=np.float64, mode="r", shape=(ne,)) if ne else np.zeros(0)
            if not np.isfinite(self._W).all():
                raise ValueError("bulk CSR weights must be finite")
        if self._V[0] != 0 or int(self._V[-1]) != ne:
            raise ValueError("invalid CSR offsets: first offset must be zero and final offset must equal num_entries")
        self.simple_sorted_topology = None
        if validate:
            d = np.diff(self._V.astype(np.int64))
            if (d < 0).any():
                raise ValueError("invalid CSR offsets: offsets must be monotonic and in range")
            if ne and ((self._E < 0).any() or (self._E >= n).any()):
                raise ValueError("invalid CSR index: endpoint is outside [0, num_nodes)")
            src = np.repeat(np.arange(n, dtype=np.int64), d)
            same_row = src[1:] == src[:-1] if ne > 1 else np.zeros(0, bool)
            increasing = np.asarray(self._E[1:], np.int64) > np.asarray(self._E[:-1], np.int64) if ne > 1 else same_row
            self.simple_sorted_topology = bool((src != self._E).all() and (increasing | ~same_row).all())

        self.metadata_path = metadata_path
        self.metadata = meta
        self.offsets_path = offsets_path
        self.indices_path = indices_path
        self.weights_path = weights_path
        self.weight_key = weight_key if weights_path is not None else None
        self.num_nodes = n
        self.num_entries = ne
        self.num_edges = int(meta.get("num_edges", ne))
        self._directed = bool(meta["directed"])
        self.nodes = range(n)
        self.cache = {}
        self.graph = {"name": meta.get("name", metadata_path.stem), "eggpu_bulk_csr": True}
        self.name = self.graph["name"]
        self._eggpu_bulk_signature = (str(metadata_path), int(meta.get("generation", 0)), n, ne, self._directed)
        self._validate = bool(validate)
        self._device = {}
        self.ingest_report = None

    def is_directed(self):
        return self._directed

    def is_multigraph(self):
        return False

    def __len__(self):
        return self.num_nodes

    def __iter__(self):
        return iter(self.nodes)

    def __contains__(self, node):
        return isinstance(node, (int, np.integer)) and 0 <= int(node) < self.num_nodes

    def csr_arrays(self, weight=None):
        """(V, E, W-or-None) as stored; W only for the manifest's weight key."""
        W = None
        if weight is not None:
            if self._W is None or str(weight) != self.weight_key:
                raise NotImplementedError(
                    f"weighted PageRank requires an EGGPU CSR artifact containing the {weight!r} edge attribute")
            W = self._W
        return self._V, self._E, W

    def device_graph(self, weight=None, alpha=0.85):
        """The graph resident in HBM, streamed from the artifact files on first
        use (egx_graph_from_csr_files: pinned double-buffered reads overlapped
        with the copies, validation on the device) and kept for later calls."""
        from . import native

        key = None if weight is None else str(weight)
        dg = self._device.get(key)
        if dg is None:
            wpath = None
            if weight is not None:
                if self._W is None or str(weight) != self.weight_key:
                    raise NotImplementedError(
                        f"weighted PageRank requires an EGGPU CSR artifact containing the {weight!r} edge attribute")
                wpath = self.weights_path
            dg, rep = native.DeviceGraph.from_csr_files(self.offsets_path, self.indices_path, wpath, self.num_nodes,
                                                        self.num_entries, alpha=alpha, validate=self._validate)
            self.ingest_report = rep
            if self.simple_sorted_topology is None:
                self.simple_sorted_topology = bool(rep["simple_sorted"])
            self._device[key] = dg
        return dg

    def __repr__(self):
        kind = "DiGraph" if self._directed else "Graph"
        return f"EGGPUBulk{kind}(name={self.name!r}, nodes={self.num_nodes}, csr_entries={self.num_entries})"


def read_eggpu_csr(metadata_path, validate=None):
    """Load an immutable CSR graph for the GPU route (``eggpu_bulk_graph.py:551-562``)."""
    return EGGPUBulkGraph(metadata_path, validate=validate)


def write_eggpu_csr(prefix, V, E, W=None, directed=True, name=None, weight_key="weight", generation=1):
    """Write ``prefix.offsets.i32``, ``prefix.indices.i32`` (and
    ``prefix.weights.f64``) plus the ``prefix.json`` manifest; returns the
    manifest path."""
    prefix = Path(prefix)
    V = np.ascontiguousarray(V, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.int32)
    off = prefix.with_name(prefix.name + ".offsets.i32")
    ind = prefix.with_name(prefix.name + ".indices.i32")
    V.tofile(off)
    E.tofile(ind)
    meta = {
        "format": _FORMAT,
        "generation": int(generation),
        "name": name or prefix.name,
        "directed": bool(directed),
        "num_nodes": int(len(V) - 1),
        "num_edges": int(len(E) if directed else len(E) // 2),
        "num_entries": int(len(E)),
        "node_labels": "zero_based_contiguous",
        "offset_dtype": "int32",
        "index_dtype": "int32",
        "offsets_path": off.name,
        "indices_path": ind.name,
    }
    if W is not None:
        wp = prefix.with_name(prefix.name + ".weights.f64")
        np.ascontiguousarray(W, dtype=np.float64).tofile(wp)
        meta.update({"weights_path": wp.name, "weight_dtype": "float64", "weight_key": weight_key})
    mp = prefix.with_name(prefix.name + ".json")
    mp.write_text(json.dumps(meta), encoding="utf-8")
    return mp


__all__ = ["EGGPUBulkGraph", "read_eggpu_csr", "write_eggpu_csr"]
