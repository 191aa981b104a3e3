"""Katz centrality on the device, after ``easygraph/functions/centrality/katz_centrality.py:7-110``.

The reference computes Katz centrality on the CPU only: its ``@hybrid``
decorator sends graphs without a C++ twin to this Python function, which
iterates ``x_new = alpha * A @ x + b`` from ``x = 1`` with ``A[u][v] = 1`` for
every adjacency key ``v`` of ``u``, stops when the L1 change is below ``tol``
(keeping the iterate before that update), raises ``RuntimeError`` after
``max_iter`` updates without convergence, and divides by the L2 norm when
``normalized``.  The GPU route here runs the same iteration on the SpMV
engine (``egx_katz_csr``): FP64 iterates and sums, alpha applied in FP64 to
each row sum.

Dispatch follows the PageRank route: ``EASYGRAPH_ENABLE_GPU`` selects the
device; otherwise (or on a non-strict failure) the call goes to the host
framework's CPU function, registered with :func:`set_cpu_katz` or
EasyGraph's own when importable, and raises when there is none.
"""
from __future__ import annotations

import numpy as np

from . import gpu_backend as _backend
from .exception import EasyGraphNotImplemented
from .gpu_runtime import gpu_runtime_enabled, gpu_strict_errors
from .native import NativeLibraryMissing

__all__ = ["katz_centrality", "set_cpu_katz"]

_CPU_KATZ = None


class _NotConverged(RuntimeError):
    """The reference's RuntimeError for an iteration that did not converge."""


def set_cpu_katz(fn):
    """Register the host framework's CPU Katz centrality
    ``fn(G, alpha=..., beta=..., max_iter=..., tol=..., normalized=...) -> dict``."""
    global _CPU_KATZ
    _CPU_KATZ = fn


def _cpu_route():
    if _CPU_KATZ is not None:
        return _CPU_KATZ
    try:  # the reference's own function, when EasyGraph is installed
        from easygraph.functions.centrality.katz_centrality import katz_centrality as fn  # type: ignore

        return fn
    except Exception:
        return None


def katz_centrality(G, alpha=0.1, beta=1.0, max_iter=1000, tol=1e-6, normalized=True):
    """Katz centrality of every node of G (``katz_centrality.py:7-110``).

    beta is a scalar or a dict node -> value (1.0 for nodes it omits).
    Returns a dict node -> score; raises ``RuntimeError`` when the iteration
    does not converge within ``max_iter`` updates."""
    if G.is_multigraph():  # @not_implemented_for("multigraph"), katz_centrality.py:6
        raise EasyGraphNotImplemented("not implemented for multigraph type")
    if gpu_runtime_enabled():
        try:
            return _katz_gpu(G, alpha, beta, max_iter, tol, normalized)
        except (NativeLibraryMissing, _NotConverged):
            raise
        except Exception:
            if gpu_strict_errors() or _cpu_route() is None:
                raise
    cpu = _cpu_route()
    if cpu is None:
        raise RuntimeError(
            "eggpu_amd provides the GPU route of katz_centrality only: set EASYGRAPH_ENABLE_GPU=1, or register "
            "the host framework's CPU function with eggpu_amd.set_cpu_katz()"
        )
    return cpu(G, alpha=alpha, beta=beta, max_iter=max_iter, tol=tol, normalized=normalized)


def _katz_gpu(G, alpha, beta, max_iter, tol, normalized):
    from . import native

    prepared = _backend._prepare_graph_cache(G)
    nodes = list(prepared["nodes"])
    V, E = _backend.adjacency_csr(G, prepared)
    b = np.array([beta.get(v, 1.0) for v in nodes], dtype=np.float64) if isinstance(beta, dict) else float(beta)
    x, _, converged, _ = native.katz_csr(V, E, alpha=float(alpha), beta=b, max_iter=int(max_iter),
                                         tol=float(tol), normalized=bool(normalized))
    if not converged:
        raise _NotConverged(f"Katz centrality failed to converge in {max_iter} iterations")
    return {v: float(x[i]) for i, v in enumerate(nodes)}
