"""``eg.pagerank`` drop-in (``easygraph/functions/centrality/pagerank.py:12-72,161-179``).

Same signature, same environment defaults and the same
``EASYGRAPH_ENABLE_GPU`` / ``EASYGRAPH_GPU_STRICT_ERRORS`` dispatch.  This
package supplies the GPU route only.  When the GPU route is off, or fails in
non-strict mode, the call goes to the CPU route of the host framework: the
function registered with :func:`set_cpu_pagerank`, else EasyGraph's own
``_pagerank_power_iteration`` when EasyGraph is importable.  With neither,
the call raises instead of computing anything on the CPU here.  A missing
native library always raises, in strict and non-strict mode alike.

C++ graphs (``G.cflag`` set: EasyGraph's ``GraphC`` / ``DiGraphC``) take the
reference's ``@hybrid("cpp_pagerank")`` route before any of this
(``pagerank.py:13``, ``decorators.py:140-163``): the call goes, with the
caller's own arguments, to ``cpp_easygraph.cpp_pagerank`` (or the function
registered with :func:`set_cpp_pagerank`), ``max_iter`` and ``tol`` renamed
``max_iterator`` and ``threshold``.  Without such a function the call stays
on the route above.
"""
from __future__ import annotations

import functools
import os

from . import gpu_backend as eggpu_backend
from .exception import EasyGraphNotImplemented
from .gpu_runtime import gpu_runtime_enabled, gpu_strict_errors
from .native import NativeLibraryMissing

__all__ = ["pagerank", "set_cpu_pagerank", "set_cpp_pagerank"]

_CPU_PAGERANK = None
_CPP_PAGERANK = None


def set_cpp_pagerank(fn):
    """Register the C++-graph route ``fn(G, *args, **kwargs)``, called as
    ``cpp_easygraph.cpp_pagerank`` would be (None: look that up again)."""
    global _CPP_PAGERANK
    _CPP_PAGERANK = fn


def _cpp_graph_route():
    if _CPP_PAGERANK is not None:
        return _CPP_PAGERANK
    try:
        import cpp_easygraph  # type: ignore

        return getattr(cpp_easygraph, "cpp_pagerank", None)
    except ImportError:
        return None


def _cpp_graphs_to_cpp(fn):
    """The @hybrid("cpp_pagerank") rule: a graph with cflag set is handed,
    with the caller's arguments as given, to the C++ route."""

    @functools.wraps(fn)
    def entry(*args, **kwargs):
        G = args[0] if args else kwargs.get("G")
        if G is not None and getattr(G, "cflag", 0):
            route = _cpp_graph_route()
            if route is not None:
                renamed = {("max_iterator" if k == "max_iter" and "max_iterator" not in kwargs else
                            "threshold" if k == "tol" and "threshold" not in kwargs else k): v
                           for k, v in kwargs.items()}
                return route(*args, **renamed)
        return fn(*args, **kwargs)

    return entry


def set_cpu_pagerank(fn):
    """Register the host framework's CPU PageRank
    ``fn(G, alpha=..., weight=..., max_iter=..., tol=...) -> Mapping``."""
    global _CPU_PAGERANK
    _CPU_PAGERANK = fn


def _cpu_route():
    if _CPU_PAGERANK is not None:
        return _CPU_PAGERANK
    try:  # the reference's own CPU path, when EasyGraph is installed
        from easygraph.functions.centrality.pagerank import _pagerank_power_iteration  # type: ignore

        return _pagerank_power_iteration
    except Exception:
        return None


def _not_for_multigraphs(fn):
    # @not_implemented_for("multigraph") (pagerank.py:12), outside the C++ route
    @functools.wraps(fn)
    def entry(*args, **kwargs):
        G = args[0] if args else kwargs.get("G")
        if G is not None and G.is_multigraph():
            raise EasyGraphNotImplemented("not implemented for multigraph type")
        return fn(*args, **kwargs)

    return entry


@_not_for_multigraphs
@_cpp_graphs_to_cpp
def pagerank(G, alpha=0.85, weight=None, max_iter=None, tol=None):
    """Returns the PageRank value of each node in G.

    Parameters follow ``pagerank.py:14-43``.  ``max_iter`` defaults to
    ``EASYGRAPH_CPU_PR_MAX_ITER`` (200) and ``tol`` to ``EASYGRAPH_CPU_PR_TOL``
    (1e-6); the iteration stops once the L1 change is below ``tol * n``.
    Returns a read-only ``Mapping`` node -> score."""
    if max_iter is None:
        max_iter_eff = max(1, int(os.environ.get("EASYGRAPH_CPU_PR_MAX_ITER", "200")))
    else:
        max_iter_eff = max(1, int(max_iter))
    if tol is None:
        tol_eff = max(0.0, float(os.environ.get("EASYGRAPH_CPU_PR_TOL", "1e-6")))
    else:
        tol_eff = max(0.0, float(tol))

    gpu_result = _pagerank_gpu_runtime_dispatch(G, alpha=alpha, weight=weight, max_iter=max_iter_eff,
                                                tol=tol_eff)
    if gpu_result is not None:
        return gpu_result
    if len(G) == 0:
        return {}
    cpu = _cpu_route()
    if cpu is None:
        raise RuntimeError(
            "eggpu_amd provides the GPU route of pagerank only: set EASYGRAPH_ENABLE_GPU=1, or register "
            "the host framework's CPU PageRank with eggpu_amd.set_cpu_pagerank()"
        )
    return cpu(G, alpha=alpha, weight=weight, max_iter=max_iter_eff, tol=tol_eff)


def _pagerank_gpu_runtime_dispatch(G, alpha, weight, max_iter, tol):
    """``pagerank.py:161-179``: GPU first; errors propagate in strict mode
    and for bulk graphs, otherwise the caller falls through to the CPU route."""
    if not gpu_runtime_enabled():
        return None
    try:
        if eggpu_backend.eggpu_backend_enabled():
            return eggpu_backend.pagerank(G, alpha=alpha, max_iter=int(max_iter), eps=float(tol), weight=weight)
    except NativeLibraryMissing:
        raise
    except Exception:
        if gpu_strict_errors() or getattr(G, "_eggpu_bulk_csr", False) or _cpu_route() is None:
            raise
    return None
