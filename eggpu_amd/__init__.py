"""eggpu_amd -- MI355X-native (gfx950) PageRank route for EasyGraph.

``pagerank(G, alpha=0.85, weight=None, max_iter=None, tol=None)`` is the
drop-in for ``eg.pagerank`` (GPU route); ``katz_centrality`` runs the
reference's CPU Katz iteration on the same SpMV engine; ``gpu_backend``
stands in for EasyGraph's GPU backend module; ``native`` is the ctypes
binding of ``libegx.so`` (C ABI in ``include/egx.h``).
"""
from . import gpu_backend, gpu_runtime, native
from .bulk import EGGPUBulkGraph, read_eggpu_csr, write_eggpu_csr
from .classes import DiGraph, Graph, MultiDiGraph, MultiGraph
from .exception import EasyGraphError, EasyGraphException, EasyGraphNotImplemented
from .native import DeviceGraph, NativeLibraryMissing
from .eigenvector import eigenvector_centrality, set_cpu_eigenvector
from .katz import katz_centrality, set_cpu_katz
from .pagerank import pagerank, set_cpp_pagerank, set_cpu_pagerank

__version__ = "0.1.0"

__all__ = [
    "pagerank",
    "set_cpu_pagerank",
    "set_cpp_pagerank",
    "eigenvector_centrality",
    "katz_centrality",
    "set_cpu_eigenvector",
    "set_cpu_katz",
    "DeviceGraph",
    "NativeLibraryMissing",
    "Graph",
    "DiGraph",
    "MultiGraph",
    "MultiDiGraph",
    "EGGPUBulkGraph",
    "read_eggpu_csr",
    "write_eggpu_csr",
    "EasyGraphException",
    "EasyGraphError",
    "EasyGraphNotImplemented",
    "gpu_backend",
    "gpu_runtime",
    "native",
]
