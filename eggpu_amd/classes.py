"""Minimal graphs that speak EasyGraph's graph protocol.

The drop-in consumes any object with EasyGraph's protocol (``_node``,
``_adj``, ``nodes``, ``edges``, ``is_directed()``, ``is_multigraph()``,
``cache``, ``_mutation_generation``), including EasyGraph's own ``Graph`` /
``DiGraph`` and networkx graphs.  These classes give tests and benchmarks
such graphs without EasyGraph installed.  They keep the behaviours the GPU
route depends on:

* node order is insertion order (``easygraph/classes/graph.py``);
* ``DiGraph.edges`` lists ``(u, v, data)`` for u in ``_adj``, v in
  ``_adj[u]`` (``directed_graph.py:128-136``); ``Graph.edges`` lists each
  undirected pair once (``graph.py:261-282``);
* every mutation, including ``G[u][v][key] = value`` through an edge
  attribute dict, bumps ``_mutation_generation`` and drops ``cache``
  (``graph.py:17-78,1641-1644``).
"""
from __future__ import annotations

import copy
import weakref


def _mutator(name):
    """dict method `name`, followed by a notice to the owning graph."""
    base = getattr(dict, name)

    def method(self, *args, **kwargs):
        out = base(self, *args, **kwargs)
        owner = self._owner()
        if owner is not None:
            owner._clear_cache()
        return out

    method.__name__ = name
    method.__qualname__ = f"_EdgeData.{name}"
    return method


class _EdgeData(dict):
    """The attribute dict of one edge.  A plain dict for every read; every
    dict method that can write is wrapped so that the owning graph drops its
    derived state (cache and generation) afterwards.  A write that turns out
    to change nothing still notifies, which only costs a rebuild.  Pickles and
    deep-copies as the plain dict it holds, without the owner."""

    __slots__ = ("_owner",)

    def __init__(self, owner, *args, **kwargs):
        dict.__init__(self, *args, **kwargs)
        self._owner = weakref.ref(owner)

    for _name in ("__setitem__", "__delitem__", "__ior__", "update", "setdefault", "pop", "popitem", "clear"):
        locals()[_name] = _mutator(_name)
    del _name

    def __reduce__(self):
        return (dict, (dict(self),))

    def __deepcopy__(self, memo):
        return copy.deepcopy(dict(self), memo)


class Graph:
    """Undirected simple graph."""

    cflag = 0

    def __init__(self, incoming_graph_data=None, **graph_attr):
        self.graph = dict(graph_attr)
        self._node = {}
        self._adj = {}
        self.cache = {}
        self._mutation_generation = 0
        if incoming_graph_data is not None:
            self.add_edges_from(incoming_graph_data)

    # -- protocol -------------------------------------------------------------
    def _clear_cache(self):
        self._mutation_generation = getattr(self, "_mutation_generation", 0) + 1
        self.cache = {}

    def is_directed(self):
        return False

    def is_multigraph(self):
        return False

    @property
    def nodes(self):
        return self._node

    @property
    def adj(self):
        return self._adj

    @property
    def edges(self):
        out = []
        seen = set()
        for u, nbrs in self._adj.items():
            for v, data in nbrs.items():
                if (u, v) in seen:
                    continue
                seen.add((u, v))
                seen.add((v, u))
                out.append((u, v, data))
        return out

    def __len__(self):
        return len(self._node)

    def __iter__(self):
        return iter(self._node)

    def __contains__(self, node):
        try:
            return node in self._node
        except TypeError:
            return False

    def __getitem__(self, node):
        return self._adj[node]

    def number_of_nodes(self):
        return len(self._node)

    def number_of_edges(self):
        return len(self.edges)

    # -- mutation -------------------------------------------------------------
    def _ensure_node(self, n, attr=None):
        if n not in self._node:
            self._node[n] = dict(attr or {})
            self._adj[n] = {}
            self._on_new_node(n)
        elif attr:
            self._node[n].update(attr)

    def _on_new_node(self, n):
        pass

    def add_node(self, node_for_adding, **node_attr):
        self._ensure_node(node_for_adding, node_attr)
        self._clear_cache()

    def add_nodes(self, nodes_for_adding, nodes_attr=None):
        nodes_attr = nodes_attr or [None] * len(list(nodes_for_adding))
        for n, a in zip(list(nodes_for_adding), nodes_attr):
            self._ensure_node(n, a)
        self._clear_cache()

    def add_nodes_from(self, nodes_for_adding, **attr):
        for n in nodes_for_adding:
            self._ensure_node(n, attr)
        self._clear_cache()

    def _link(self, u, v, attrs):
        data = self._adj[u].get(v)
        if data is None:
            data = _EdgeData(self)
            self._adj[u][v] = data
            self._adj[v][u] = data
        dict.update(data, attrs)

    def add_edge(self, u_of_edge, v_of_edge, **edge_attr):
        self._ensure_node(u_of_edge)
        self._ensure_node(v_of_edge)
        self._link(u_of_edge, v_of_edge, edge_attr)
        self._clear_cache()

    def add_edges(self, edges_for_adding, edges_attr=None):
        edges_for_adding = list(edges_for_adding)
        edges_attr = edges_attr or [{}] * len(edges_for_adding)
        for (u, v), a in zip(edges_for_adding, edges_attr):
            self._ensure_node(u)
            self._ensure_node(v)
            self._link(u, v, a or {})
        self._clear_cache()

    def add_edges_from(self, ebunch_to_add, **attr):
        for e in ebunch_to_add:
            if len(e) == 3:
                u, v, dd = e
                a = dict(attr)
                a.update(dd)
            else:
                u, v = e
                a = dict(attr)
            self._ensure_node(u)
            self._ensure_node(v)
            self._link(u, v, a)
        self._clear_cache()

    def add_weighted_edges_from(self, ebunch_to_add, weight="weight", **attr):
        self.add_edges_from(((u, v, {weight: d}) for u, v, d in ebunch_to_add), **attr)

    def remove_edge(self, u, v):
        del self._adj[u][v]
        if u != v:
            del self._adj[v][u]
        self._clear_cache()

    def remove_node(self, n):
        for nbr in list(self._adj[n]):
            if nbr != n:
                del self._adj[nbr][n]
        del self._adj[n]
        del self._node[n]
        self._clear_cache()

    def remove_nodes_from(self, nodes):
        for n in list(nodes):
            if n in self._node:
                self.remove_node(n)
        self._clear_cache()


class DiGraph(Graph):
    """Directed simple graph (self-loops allowed)."""

    def __init__(self, incoming_graph_data=None, **graph_attr):
        self._pred = {}
        super().__init__(incoming_graph_data, **graph_attr)

    def is_directed(self):
        return True

    def _on_new_node(self, n):
        self._pred[n] = {}

    @property
    def edges(self):
        return [(u, v, data) for u, nbrs in self._adj.items() for v, data in nbrs.items()]

    def _link(self, u, v, attrs):
        data = self._adj[u].get(v)
        if data is None:
            data = _EdgeData(self)
            self._adj[u][v] = data
            self._pred[v][u] = data
        dict.update(data, attrs)

    def remove_edge(self, u, v):
        del self._adj[u][v]
        del self._pred[v][u]
        self._clear_cache()

    def remove_node(self, n):
        for v in list(self._adj[n]):
            del self._pred[v][n]
        for u in list(self._pred[n]):
            del self._adj[u][n]
        del self._adj[n]
        del self._pred[n]
        del self._node[n]
        self._clear_cache()


class MultiGraph(Graph):
    """Marker class: PageRank rejects multigraphs (``pagerank.py:12``)."""

    def is_multigraph(self):
        return True


class MultiDiGraph(DiGraph):
    def is_multigraph(self):
        return True


__all__ = ["Graph", "DiGraph", "MultiGraph", "MultiDiGraph"]
