"""Opt-in switches and node indexing of the GPU route.

Same names and parsing as ``easygraph/utils/gpu_runtime.py:3-58``: a switch
is on when its value, stripped and upper-cased, is one of 1/TRUE/ON/YES.
"""
from __future__ import annotations

import os

_TRUE_VALUES = {"1", "TRUE", "ON", "YES"}


def _env_true(name):
    return os.environ.get(name, "").strip().upper() in _TRUE_VALUES


def gpu_runtime_enabled():
    """``EASYGRAPH_ENABLE_GPU`` (``gpu_runtime.py:10-11``)."""
    return _env_true("EASYGRAPH_ENABLE_GPU")


def gpu_strict_errors():
    """``EASYGRAPH_GPU_STRICT_ERRORS`` (``gpu_runtime.py:14-15``)."""
    return _env_true("EASYGRAPH_GPU_STRICT_ERRORS")


def graph_nodes_list(G):
    """Node order of the reference: ``list(G.nodes.keys())`` (``gpu_runtime.py:18-22``)."""
    nodes = G.nodes
    if hasattr(nodes, "keys"):
        return list(nodes.keys())
    return list(nodes)


def build_node_index(G):
    nodes = graph_nodes_list(G)
    return nodes, {node: idx for idx, node in enumerate(nodes)}
