"""Eigenvector centrality on the device, after ``easygraph/functions/centrality/eigenvector.py:12-154``.

The reference computes eigenvector centrality on the CPU only.  Its Python
function builds ``A[v][u] = w(u, v)`` from ``G[u].items()`` in
``list(G.nodes)`` order (weight ``attrs.get(weight, 1.0)``, 1.0 without a
weight key), gives every node whose row and column sums are both 0 a 1e-4
self-loop, and then

* with ``nstart``: runs the power iteration ``v = A v / ||A v||`` from
  ``nstart`` (1.0 for nodes it omits) divided by its absolute sum, keeping
  ``v`` and stopping when ``||A v|| < 1e-10``, stopping when
  ``||v' - v||_2 < tol``, and raising ``EasyGraphError`` after ``max_iter``
  iterations without that (``eigenvector.py:110-131``);
* without: calls ARPACK, ``scipy.sparse.linalg.eigs(A, k=1, which='LR',
  maxiter=max_iter, tol=tol)`` (``eigenvector.py:134-138``);

and finally flips the sign if the sum is negative, divides by the L2 norm,
and sets the isolated nodes to 0 (renormalising when the sum stays positive).

The GPU route here runs the ``nstart`` iteration as the reference does, on
the SpMV engine (``egx_eigenvector_csr``, FP64 iterates and sums).  ARPACK's
iterates cannot be reproduced, so without ``nstart`` it runs the same
iteration on ``A + I`` from the all-ones vector.  ``A + I`` has the same
eigenvectors, and its principal eigenvalue is strictly dominant, so the
iteration also converges on bipartite graphs (trees, stars, grids), where the
plain iteration oscillates and ARPACK still returns a vector.  It stops when
``||v' - v||_2 < tol``.  It is given ``20 * max_iter`` iterations, ARPACK's
budget of ``max_iter`` restarts of a 20-vector Arnoldi basis.  Where the
principal eigenvector is unique, both routes return it to within their
tolerances.

Dispatch follows the PageRank route: ``EASYGRAPH_ENABLE_GPU`` selects the
device; otherwise (or on a non-strict failure) the call goes to the host
framework's CPU function, registered with :func:`set_cpu_eigenvector` or
EasyGraph's own when importable, and raises when there is none.
"""
from __future__ import annotations

import numpy as np

from . import gpu_backend as _backend
from .exception import EasyGraphError, EasyGraphNotImplemented, EasyGraphPointlessConcept
from .gpu_runtime import gpu_runtime_enabled, gpu_strict_errors
from .native import NativeLibraryMissing

__all__ = ["eigenvector_centrality", "set_cpu_eigenvector"]

_CPU_EIGENVECTOR = None
# Arnoldi basis size scipy's eigs uses for k = 1 (ncv = max(2k + 1, 20))
_ARNOLDI_VECTORS = 20


def set_cpu_eigenvector(fn):
    """Register the host framework's CPU eigenvector centrality
    ``fn(G, max_iter=..., tol=..., nstart=..., weight=...) -> dict``."""
    global _CPU_EIGENVECTOR
    _CPU_EIGENVECTOR = fn


def _cpu_route():
    if _CPU_EIGENVECTOR is not None:
        return _CPU_EIGENVECTOR
    try:  # the reference's own function, when EasyGraph is installed
        from easygraph.functions.centrality.eigenvector import eigenvector_centrality as fn  # type: ignore

        return fn
    except Exception:
        return None


def eigenvector_centrality(G, max_iter=100, tol=1.0e-6, nstart=None, weight=None):
    """Eigenvector centrality of every node of G (``eigenvector.py:12-154``).

    Returns a dict node -> value with unit L2 norm.  Raises
    ``EasyGraphPointlessConcept`` for graphs with fewer than two nodes and
    ``EasyGraphError`` when the iteration does not converge."""
    if G.is_multigraph():  # @not_implemented_for("multigraph"), eigenvector.py:12
        raise EasyGraphNotImplemented("not implemented for multigraph type")
    if len(G) == 0:
        raise EasyGraphPointlessConcept("cannot compute centrality for the null graph")
    if len(G) == 1:
        raise EasyGraphPointlessConcept("cannot compute eigenvector centrality for a single node graph")
    if gpu_runtime_enabled():
        try:
            return _eigenvector_gpu(G, max_iter, tol, nstart, weight)
        except (NativeLibraryMissing, EasyGraphError):
            raise
        except Exception:
            if gpu_strict_errors() or _cpu_route() is None:
                raise
    cpu = _cpu_route()
    if cpu is None:
        raise RuntimeError(
            "eggpu_amd provides the GPU route of eigenvector_centrality only: set EASYGRAPH_ENABLE_GPU=1, or "
            "register the host framework's CPU function with eggpu_amd.set_cpu_eigenvector()"
        )
    return cpu(G, max_iter=max_iter, tol=tol, nstart=nstart, weight=weight)


def with_isolated_loops(V, E, W):
    """(V, E, W, isolated): the CSR with a 1e-4 self-loop appended to the row
    of every node whose in- and out-weight sums are both 0
    (``eigenvector.py:94-109``); W is None when no weights were given and no
    node is isolated."""
    n = len(V) - 1
    deg = np.diff(V)
    w = np.ones(len(E)) if W is None else np.asarray(W, dtype=np.float64)
    rows = np.repeat(np.arange(n), deg)
    out_sum = np.bincount(rows, weights=w, minlength=n)
    in_sum = np.bincount(E, weights=w, minlength=n)
    isolated = np.flatnonzero((out_sum == 0) & (in_sum == 0))
    if len(isolated) == 0:
        return V, E, W, isolated
    extra = np.zeros(n, dtype=np.int64)
    extra[isolated] = 1
    V2 = np.zeros(n + 1, dtype=np.int64)
    V2[1:] = np.cumsum(deg + extra)
    pos = V2[rows] + (np.arange(len(E)) - V[rows])
    E2 = np.empty(V2[-1], dtype=np.int32)
    W2 = np.empty(V2[-1], dtype=np.float64)
    E2[pos] = E
    W2[pos] = w
    E2[V2[isolated + 1] - 1] = isolated
    W2[V2[isolated + 1] - 1] = 1.0e-4
    return V2.astype(np.int32), E2, W2, isolated


def _adjacency(G, weight):
    prepared = _backend._prepare_graph_cache(G)
    nodes = list(prepared["nodes"])
    V, E = _backend.adjacency_csr(G, prepared)
    W = _backend.adjacency_weights(G, prepared, weight) if weight else None
    return nodes, np.asarray(V, dtype=np.int64), np.asarray(E, dtype=np.int32), W


def _eigenvector_gpu(G, max_iter, tol, nstart, weight):
    from . import native

    nodes, V, E, W = _adjacency(G, weight)
    V, E, W, isolated = with_isolated_loops(V, E, W)
    if nstart is not None:
        x0 = np.array([nstart.get(v, 1.0) for v in nodes], dtype=np.float64)
        x, _, stop, _ = native.eigenvector_csr(V, E, W, x0, max_iter=int(max_iter), tol=float(tol), delta=2,
                                               norm_floor=1e-10, shift=0.0, start_norm=1)
        if stop == native.EIG_MAX_ITER:
            raise EasyGraphError(f"Eigenvector calculation did not converge in {max_iter} iterations")
    else:
        budget = max(int(max_iter), 1) * _ARNOLDI_VECTORS
        x, _, stop, _ = native.eigenvector_csr(V, E, W, None, max_iter=budget, tol=float(tol), delta=2,
                                               norm_floor=1e-10, shift=1.0, start_norm=1)
        if stop == native.EIG_MAX_ITER:
            raise EasyGraphError(f"Eigenvector calculation did not converge in {max_iter} iterations")
    return _finish(nodes, x, isolated)


def _finish(nodes, c, isolated):
    """Sign, L2 normalisation and isolated nodes (``eigenvector.py:140-154``)."""
    if c.sum() < 0:
        c = -c
    c = c / np.linalg.norm(c)
    if len(isolated):
        c[isolated] = 0.0
        if np.sum(c) > 0:
            c = c / np.linalg.norm(c)
    return {v: float(c[i]) for i, v in enumerate(nodes)}
