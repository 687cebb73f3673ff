"""Python mirror of ``Zog.Centrality`` (lib/zog/centrality.ex), the
copy-in/copy-out entry points: every call builds a graph from the builder's
arrays, runs one algorithm and drops the graph (centrality.ex:62-129 build an
ArrayGraph per call; here a device resource per call, on the same kernels as
``Zog.ResourceGraph``).  Scores come back as a map from label to score
(map_scores, centrality.ex:319-324); the two betweenness functions halve the
scores of an undirected builder first (maybe_scale_undirected, :326-330).

  * ``betweenness_unweighted(builder)``        centrality.ex:191-197
  * ``betweenness_f64(builder)``               centrality.ex:203-209
  * ``pagerank(builder, damping=0.85, max_iterations=100, tolerance=0.0001)``
                                               centrality.ex:237-247
  * ``eigenvector(builder, max_iterations=100, tolerance=0.0001)``
                                               centrality.ex:253-274 (the
                                               zero-vector fallback included)
  * ``katz(builder, alpha=0.1, beta=1.0, max_iterations=100, tolerance=0.0001)``
                                               centrality.ex:281-291
  * ``alpha_centrality(builder, alpha=0.5, initial=1.0, max_iterations=100,
    tolerance=0.0001)``                        centrality.ex:300-312
"""
from __future__ import annotations

import math
from contextlib import contextmanager

from . import resource_graph as rg
from .soa import SoA


@contextmanager
def _resource(builder: SoA):
    g = rg.new(builder)
    try:
        yield g
    finally:
        rg.destroy(g)


def _labelled(builder: SoA, scores):
    return dict(zip(builder.all_labels(), scores))


def betweenness_unweighted(builder: SoA):
    with _resource(builder) as g:
        raw = rg.betweenness_unweighted_array(g)
    return _labelled(builder, rg._maybe_scale_undirected(builder, raw).tolist())


def betweenness_f64(builder: SoA):
    with _resource(builder) as g:
        raw = rg.betweenness_f64_array(g)
    return _labelled(builder, rg._maybe_scale_undirected(builder, raw).tolist())


def pagerank(builder: SoA, damping: float = 0.85, max_iterations: int = 100, tolerance: float = 0.0001, *,
             precision: int = 32):
    with _resource(builder) as g:
        scores = rg.pagerank_array(g, damping, max_iterations, tolerance, precision=precision)
    return _labelled(builder, scores.tolist())


def eigenvector(builder: SoA, max_iterations: int = 100, tolerance: float = 0.0001):
    with _resource(builder) as g:
        mapped = rg.eigenvector(g, max_iterations, tolerance)
    n = builder.node_count()
    if n > 1 and all(v == 0.0 for v in mapped.values()):
        # centrality.ex:263-272: a zero result falls back to uniform
        u = 1.0 / math.sqrt(n)
        return {k: u for k in mapped}
    return mapped


def katz(builder: SoA, alpha: float = 0.1, beta: float = 1.0, max_iterations: int = 100,
         tolerance: float = 0.0001):
    with _resource(builder) as g:
        return rg.katz(g, alpha, beta, max_iterations, tolerance)


def alpha_centrality(builder: SoA, alpha: float = 0.5, initial: float = 1.0, max_iterations: int = 100,
                     tolerance: float = 0.0001):
    with _resource(builder) as g:
        return rg.alpha_centrality(g, alpha, initial, max_iterations, tolerance)
