"""Python mirror of ``Zog.ResourceGraph`` for the ``:soa`` backend's two
accelerated paths (lib/zog/resource_graph.ex), calling the gfx950 library
through its C ABI (include/zog_hip.h).

Reference surface kept:
  * ``new(builder, backend="soa")``            resource_graph.ex:1393-1404
  * ``read_edgelist(path, directed=True, backend="soa", integer_labels=False)``
                                               resource_graph.ex:1474-1486; parse
                                               :990-1084 and undirected dedupe
                                               :953-988 in the library's native
                                               reader (zg_edgelist_read)
  * ``destroy(graph)``                         resource_graph.ex:1577-1580
  * ``betweenness_unweighted(graph, raw=False)``  resource_graph.ex:1590-1601
    (x0.5 for undirected builders before the raw branch, :2633-2637)
  * ``betweenness_f64(graph, raw=False)``       resource_graph.ex:1610-1621
    (arc weights from the builder, as buildGraph keeps them, :120-129)
  * ``pagerank(graph, damping=0.85, max_iterations=100, tolerance=0.0001,
    raw=False)``                               resource_graph.ex:1671-1685
  * ``eigenvector(graph, max_iterations=100, tolerance=0.0001, raw=False)``
                                               resource_graph.ex:1697-1709
  * ``katz(graph, alpha=0.1, beta=1.0, max_iterations=100, tolerance=0.0001,
    raw=False)``                               resource_graph.ex:1723-1737
  * ``alpha_centrality(graph, alpha=0.5, initial=1.0, max_iterations=100,
    tolerance=0.0001, raw=False)``             resource_graph.ex:1751-1765
  * ``node_degrees(graph)``                    resource_graph.ex:2130-2136
  * ``core_numbers(graph, raw=False)``         resource_graph.ex:2171-2192
  * ``raw=True`` returns the dense list indexed by u32 id; otherwise scores
    are mapped to labels (map_scores, :2609-2619): integer keys for an
    integer-labelled builder, else labels in insertion order.

Extensions (keyword-only, absent in the reference): ``sources=`` for sampled
Brandes, ``precision=`` for PageRank, ``info=`` dicts with per-run metrics,
and ``betweenness_centrality`` as an alias of ``betweenness_unweighted``.
Errors: a failing status raises ``ZogError`` (``NodeNotFoundError`` for an
out-of-range id), as Zigler raises a returned Zig error in Elixir.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import numpy as np

from . import _lib
from .soa import SoA

BACKENDS = ("soa",)


class Resource:
    """Owner of one ``zg_graph`` handle (the GraphRes resource)."""

    def __init__(self, n_cap: int, src: np.ndarray, dst: np.ndarray, live_mask=None, weights=None):
        L = _lib.lib()
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        if len(src) != len(dst):
            raise ValueError("from/to arrays differ in length")
        live = None if live_mask is None else np.ascontiguousarray(live_mask, dtype=np.uint8)
        h = ctypes.c_void_p()
        if weights is None:
            rc = L.zg_graph_create(int(n_cap), _ptr(live), _ptr(src), _ptr(dst), len(src), ctypes.byref(h))
            _lib.check(rc, "zg_graph_create")
        else:
            w = np.ascontiguousarray(weights, dtype=np.float64)
            if len(w) != len(src):
                raise ValueError("weight array differs in length from the arcs")
            rc = L.zg_graph_create_weighted(int(n_cap), _ptr(live), _ptr(src), _ptr(dst), _ptr(w), len(src),
                                            ctypes.byref(h))
            _lib.check(rc, "zg_graph_create_weighted")
        self.weighted = weights is not None
        self._h = h
        self._lock = threading.Lock()
        self.n_cap = int(n_cap)
        self._live = np.ones(self.n_cap, bool) if live is None else live[: self.n_cap].astype(bool)

    def live_mask(self) -> np.ndarray:
        """bool[n_cap]: the live ids (nodeIds, array_graph.zig:268-284)."""
        return self._live

    @property
    def handle(self):
        if not self._h:
            raise _lib.ZogError(_lib.ZG_ERR_INVALID_HANDLE, "released resource")
        return self._h

    def release(self):
        with self._lock:
            if self._h:
                _lib.lib().zg_graph_destroy(self._h)
                self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def info(self):
        V = ctypes.c_uint32()
        n = ctypes.c_uint32()
        m = ctypes.c_uint64()
        _lib.check(_lib.lib().zg_graph_info(self.handle, ctypes.byref(V), ctypes.byref(n), ctypes.byref(m)))
        return int(V.value), int(n.value), int(m.value)


def _ptr(a) -> Optional[ctypes.c_void_p]:
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class BatchResource(Resource):
    """One handle over several independent graphs (zg_graph_create_batch)."""

    def __init__(self, graphs):
        L = _lib.lib()
        ns = np.ascontiguousarray([int(g[0]) for g in graphs], dtype=np.uint32)
        ms = np.ascontiguousarray([len(g[1]) for g in graphs], dtype=np.uint64)
        src = np.ascontiguousarray(np.concatenate([np.asarray(g[1], np.uint32) for g in graphs]) if graphs
                                   else np.zeros(0, np.uint32), dtype=np.uint32)
        dst = np.ascontiguousarray(np.concatenate([np.asarray(g[2], np.uint32) for g in graphs]) if graphs
                                   else np.zeros(0, np.uint32), dtype=np.uint32)
        h = ctypes.c_void_p()
        rc = L.zg_graph_create_batch(len(graphs), _ptr(ns), _ptr(ms), _ptr(src), _ptr(dst), ctypes.byref(h))
        _lib.check(rc, "zg_graph_create_batch")
        self._h = h
        self._lock = threading.Lock()
        self.weighted = False
        self.n_cap = int(ns.sum())
        self._live = np.ones(self.n_cap, bool)
        self.offsets = np.concatenate([[0], np.cumsum(ns.astype(np.int64))])


def betweenness_unweighted_batch(graphs, info: Optional[dict] = None, groups: Optional[int] = None):
    """Exact unweighted BC of every graph in `graphs` ((n, src, dst) triples
    of local ids) in one batched sweep; returns one f64 array per graph
    (directed sums, as betweenness_unweighted_array)."""
    res = BatchResource(graphs)
    try:
        if groups:
            rc = _lib.lib().zg_set_batch_groups(res.handle, int(groups))
            if rc < 0:
                _lib.check(rc, "zg_set_batch_groups")
        out = np.zeros(res.n_cap, dtype=np.float64)
        inf = _lib.ZgBcInfo()
        rc = _lib.lib().zg_bc_unweighted(res.handle, None, 0, _ptr(out), ctypes.byref(inf))
        _lib.check(rc, "zg_bc_unweighted")
        if info is not None:
            info.update({k: getattr(inf, k) for k, _ in _lib.ZgBcInfo._fields_})
        return [out[res.offsets[i]: res.offsets[i + 1]] for i in range(len(graphs))]
    finally:
        res.release()


class ResourceGraph(dict):
    """``%{resource: ref, builder: SoA}`` of the reference."""

    @property
    def resource(self) -> Resource:
        return self["resource"]

    @property
    def builder(self) -> SoA:
        return self["builder"]


def new(builder: SoA, backend: str = "soa", live_mask=None) -> ResourceGraph:
    """new/5 with the builder's arcs and weights (resource_graph.ex:1393-1404,
    buildGraph :120-129)."""
    if backend not in BACKENDS:
        raise ValueError(f"backend {backend!r} is not accelerated here (only :soa)")
    src, dst, w = builder.to_edge_arrays()
    res = Resource(builder.node_count(), src, dst, live_mask, weights=w)
    return ResourceGraph(resource=res, builder=builder)


def destroy(graph: ResourceGraph) -> str:
    graph.resource.release()
    return "ok"


# ---------------------------------------------------------------------------
# edge-list ingest (nif_read_edgelist, resource_graph.ex:990-1084)

def _read_edgelist_native(path: str, directed: bool, integer_labels: bool):
    """Parse and expand an edge list with the library's reader
    (zg_edgelist_read): (src u32[], dst u32[], weight f64[], node count,
    labels in id order or None for integer labels)."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.zg_edgelist_read(os.fsencode(path), int(bool(directed)), int(bool(integer_labels)), ctypes.byref(h))
    if rc == _lib.ZG_ERR_IO:
        raise FileNotFoundError(path)
    _lib.check(rc, "zg_edgelist_read")
    try:
        pf, pt, pw = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        m = ctypes.c_uint64()
        nn = ctypes.c_uint32()
        _lib.check(L.zg_edgelist_arcs(h, ctypes.byref(pf), ctypes.byref(pt), ctypes.byref(pw), ctypes.byref(m),
                                      ctypes.byref(nn)), "zg_edgelist_arcs")
        M = int(m.value)

        def arr(p, ct, dt):
            if M == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(M,)).copy()

        src = arr(pf, ctypes.c_uint32, np.uint32)
        dst = arr(pt, ctypes.c_uint32, np.uint32)
        w = arr(pw, ctypes.c_double, np.float64)
        labels = None
        if not integer_labels:
            pb, po = ctypes.c_void_p(), ctypes.c_void_p()
            nl = ctypes.c_uint32()
            _lib.check(L.zg_edgelist_labels(h, ctypes.byref(pb), ctypes.byref(po), ctypes.byref(nl)),
                       "zg_edgelist_labels")
            n_l = int(nl.value)
            offs = np.ctypeslib.as_array(ctypes.cast(po, ctypes.POINTER(ctypes.c_uint64)), shape=(n_l + 1,))
            blob = ctypes.string_at(pb, int(offs[-1])) if n_l else b""
            labels = [blob[int(offs[i]):int(offs[i + 1])].decode() for i in range(n_l)]
        return src, dst, w, int(nn.value), labels
    finally:
        L.zg_edgelist_free(h)


def read_edgelist(path: str, directed: bool = True, backend: str = "soa",
                  integer_labels: bool = False) -> ResourceGraph:
    if backend not in BACKENDS:
        raise ValueError(f"backend {backend!r} is not accelerated here (only :soa)")
    src, dst, wts, next_id, labels = _read_edgelist_native(os.path.expanduser(path), directed, integer_labels)
    kind = "directed" if directed else "undirected"
    # build_from_labels (resource_graph.ex:1531-1575): the builder carries only
    # labels; its edge list stays empty
    b = SoA(kind)
    if integer_labels:
        b.integer_labels = True
        b.next_id = next_id
    else:
        for lab in labels:
            b.ensure_node(lab)
    res = Resource(next_id, src, dst, weights=wts)
    return ResourceGraph(resource=res, builder=b)


# ---------------------------------------------------------------------------
# scores

def _map_scores(builder: SoA, scores):
    """map_scores (resource_graph.ex:2609-2619)."""
    if builder.integer_labels:
        return {i: s for i, s in enumerate(scores)}
    return dict(zip(builder.all_labels(), scores))


def _maybe_scale_undirected(builder: SoA, scores: np.ndarray) -> np.ndarray:
    """maybe_scale_undirected (resource_graph.ex:2633-2637)."""
    return scores * 0.5 if builder.kind == "undirected" else scores


def betweenness_unweighted_array(graph: ResourceGraph, sources=None, info: Optional[dict] = None) -> np.ndarray:
    """Unscaled directed-sum BC as f64[n_cap] (the NIF's return value)."""
    res = graph.resource
    out = np.zeros(res.n_cap, dtype=np.float64)
    inf = _lib.ZgBcInfo()
    if sources is None:
        rc = _lib.lib().zg_bc_unweighted(res.handle, None, 0, _ptr(out), ctypes.byref(inf))
    else:
        s = np.ascontiguousarray(sources, dtype=np.uint32)
        rc = _lib.lib().zg_bc_unweighted(res.handle, _ptr(s), len(s), _ptr(out), ctypes.byref(inf))
    _lib.check(rc, "zg_bc_unweighted")
    if info is not None:
        info.update({k: getattr(inf, k) for k, _ in _lib.ZgBcInfo._fields_})
    return out


def betweenness_unweighted(graph: ResourceGraph, raw: bool = False, *, sources=None,
                           info: Optional[dict] = None):
    raw_scores = betweenness_unweighted_array(graph, sources=sources, info=info)
    scores = _maybe_scale_undirected(graph.builder, raw_scores)
    if raw:
        return scores.tolist()
    return _map_scores(graph.builder, scores.tolist())


def betweenness_f64_array(graph: ResourceGraph, sources=None, info: Optional[dict] = None) -> np.ndarray:
    """Unscaled directed-sum weighted BC as f64[n_cap] (nif_betweenness_f64,
    resource_graph.ex:217-226)."""
    res = graph.resource
    out = np.zeros(res.n_cap, dtype=np.float64)
    inf = _lib.ZgBcInfo()
    if sources is None:
        rc = _lib.lib().zg_bc_f64(res.handle, None, 0, _ptr(out), ctypes.byref(inf))
    else:
        s = np.ascontiguousarray(sources, dtype=np.uint32)
        rc = _lib.lib().zg_bc_f64(res.handle, _ptr(s), len(s), _ptr(out), ctypes.byref(inf))
    _lib.check(rc, "zg_bc_f64")
    if info is not None:
        info.update({k: getattr(inf, k) for k, _ in _lib.ZgBcInfo._fields_})
    return out


def betweenness_f64(graph: ResourceGraph, raw: bool = False, *, sources=None, info: Optional[dict] = None):
    """Weighted betweenness (resource_graph.ex:1610-1621): x0.5 for an
    undirected builder, then the raw list or label-mapped scores."""
    raw_scores = betweenness_f64_array(graph, sources=sources, info=info)
    scores = _maybe_scale_undirected(graph.builder, raw_scores)
    if raw:
        return scores.tolist()
    return _map_scores(graph.builder, scores.tolist())


def betweenness_centrality(graph: ResourceGraph, raw: bool = False, **kw):
    """Alias of ``betweenness_unweighted`` (same semantics)."""
    return betweenness_unweighted(graph, raw=raw, **kw)


def _power(graph: ResourceGraph, fn: str, args, raw: bool, info: Optional[dict]):
    res = graph.resource
    out = np.zeros(res.n_cap, dtype=np.float64)
    inf = _lib.ZgPrInfo()
    L = _lib.lib()
    _lib.check(getattr(L, fn)(res.handle, *args, _ptr(out), ctypes.byref(inf)), fn)
    if info is not None:
        info.update({k: getattr(inf, k) for k, _ in _lib.ZgPrInfo._fields_})
    if raw:
        return out.tolist()
    return _map_scores(graph.builder, out.tolist())


def eigenvector(graph: ResourceGraph, max_iterations: int = 100, tolerance: float = 0.0001, raw: bool = False,
                *, info: Optional[dict] = None):
    """Eigenvector centrality (resource_graph.ex:1697-1709; NIF :266-275)."""
    if max_iterations < 0:
        raise ValueError("max_iterations must be >= 0")
    return _power(graph, "zg_eigenvector", (int(max_iterations), float(tolerance)), raw, info)


def katz(graph: ResourceGraph, alpha: float = 0.1, beta: float = 1.0, max_iterations: int = 100,
         tolerance: float = 0.0001, raw: bool = False, *, info: Optional[dict] = None):
    """Katz centrality (resource_graph.ex:1723-1737; NIF :277-286)."""
    if max_iterations < 0:
        raise ValueError("max_iterations must be >= 0")
    return _power(graph, "zg_katz", (float(alpha), float(beta), int(max_iterations), float(tolerance)), raw, info)


def alpha_centrality(graph: ResourceGraph, alpha: float = 0.5, initial: float = 1.0, max_iterations: int = 100,
                     tolerance: float = 0.0001, raw: bool = False, *, info: Optional[dict] = None):
    """Alpha centrality (resource_graph.ex:1751-1765; NIF :288-297)."""
    if max_iterations < 0:
        raise ValueError("max_iterations must be >= 0")
    return _power(graph, "zg_alpha_centrality",
                  (float(alpha), float(initial), int(max_iterations), float(tolerance)), raw, info)


def node_degrees(graph: ResourceGraph):
    """Live out-degrees by u32 id (resource_graph.ex:2130-2136; NIF :562-599)."""
    res = graph.resource
    out = np.zeros(res.n_cap, dtype=np.uint32)
    _lib.check(_lib.lib().zg_node_degrees(res.handle, _ptr(out)), "zg_node_degrees")
    return out.tolist()


def core_numbers(graph: ResourceGraph, raw: bool = False):
    """Core numbers (resource_graph.ex:2171-2192; NIF :622-628): the raw
    list by u32 id, or labels in insertion order mapped to their core number."""
    res = graph.resource
    out = np.zeros(res.n_cap, dtype=np.uint32)
    _lib.check(_lib.lib().zg_core_numbers(res.handle, _ptr(out)), "zg_core_numbers")
    cores = out.tolist()
    if raw:
        return cores
    labels = graph.builder.all_labels()
    return {labels[i]: c for i, c in enumerate(cores)}


def pagerank_array(graph: ResourceGraph, damping: float = 0.85, max_iterations: int = 100,
                   tolerance: float = 0.0001, *, precision: int = 32,
                   info: Optional[dict] = None) -> np.ndarray:
    res = graph.resource
    if max_iterations < 0:
        raise ValueError("max_iterations must be non-negative")
    out = np.zeros(res.n_cap, dtype=np.float64)
    inf = _lib.ZgPrInfo()
    rc = _lib.lib().zg_pagerank(res.handle, float(damping), int(max_iterations), float(tolerance),
                                int(precision), _ptr(out), ctypes.byref(inf))
    _lib.check(rc, "zg_pagerank")
    if info is not None:
        info.update({k: getattr(inf, k) for k, _ in _lib.ZgPrInfo._fields_})
    return out


def pagerank(graph: ResourceGraph, damping: float = 0.85, max_iterations: int = 100,
             tolerance: float = 0.0001, raw: bool = False, *, precision: int = 32,
             info: Optional[dict] = None):
    scores = pagerank_array(graph, damping, max_iterations, tolerance, precision=precision, info=info)
    if raw:
        return scores.tolist()
    return _map_scores(graph.builder, scores.tolist())


def bc_paths(graph: ResourceGraph, sources, max_levels: int = 0):
    """Per-source BFS depth (u32, 0xFFFFFFFF unreached), stored path counts
    (f64) and per-level shifts; for parity checks."""
    res = graph.resource
    s = np.ascontiguousarray(sources, dtype=np.uint32)
    V = res.n_cap
    depth = np.empty((len(s), V), dtype=np.uint32)
    sigma = np.empty((len(s), V), dtype=np.float64)
    levels = np.empty(len(s), dtype=np.uint32)
    ml = int(max_levels) if max_levels else V + 1
    shift = np.zeros((len(s), ml), dtype=np.int32)
    rc = _lib.lib().zg_bc_paths(res.handle, _ptr(s), len(s), _ptr(depth), _ptr(sigma), _ptr(shift), ml,
                                _ptr(levels))
    _lib.check(rc, "zg_bc_paths")
    return depth, sigma, shift, levels


def set_batch_groups(graph: ResourceGraph, groups: int) -> int:
    rc = _lib.lib().zg_set_batch_groups(graph.resource.handle, int(groups))
    if rc < 0:
        _lib.check(rc, "zg_set_batch_groups")
    return rc


def set_push_alpha(graph: ResourceGraph, alpha: float) -> None:
    """Forward levels run push-style when frontier arcs * alpha < arcs
    (0 = always pull)."""
    _lib.check(_lib.lib().zg_set_tuning(graph.resource.handle, _lib.ZG_TUNE_PUSH_ALPHA, float(alpha)),
               "zg_set_tuning")


def set_pull_grid(graph: ResourceGraph, blocks: int) -> None:
    """Workgroups per 64-source group on pull levels (0 = one wavefront per
    predecessor chunk)."""
    _lib.check(_lib.lib().zg_set_tuning(graph.resource.handle, _lib.ZG_TUNE_PULL_GRID, float(blocks)),
               "zg_set_tuning")


def set_engine(graph: ResourceGraph, engine: str) -> None:
    """'auto' (default), 'lanes' (64 sources per wavefront) or 'queue'
    (per-source queues, for high-diameter graphs)."""
    code = {"auto": _lib.ZG_ENGINE_AUTO, "lanes": _lib.ZG_ENGINE_LANES, "queue": _lib.ZG_ENGINE_QUEUE}[engine]
    _lib.check(_lib.lib().zg_set_tuning(graph.resource.handle, _lib.ZG_TUNE_ENGINE, float(code)), "zg_set_tuning")


def set_lane_order(graph: ResourceGraph, on: bool = True) -> None:
    """True (default): the 64-lane engine runs on a copy of the resource whose
    vertices are renumbered by decreasing degree within each graph (built at
    its first use); False: on the resource's own ids.  Scores come back in the
    resource's ids either way."""
    _lib.check(_lib.lib().zg_set_tuning(graph.resource.handle, _lib.ZG_TUNE_LANE_ORDER, 1.0 if on else 0.0),
               "zg_set_tuning")


def set_source_order(graph: ResourceGraph, mode: int) -> int:
    """0 = sweep sources in the given order, 1 = group them by distance to the
    hub core (default), 2 = by discovery rank in a search from the hub core.
    Returns the previous mode."""
    rc = _lib.lib().zg_set_source_order(graph.resource.handle, int(mode))
    if rc < 0:
        _lib.check(rc, "zg_set_source_order")
    return rc


def set_profiling(graph: ResourceGraph, on: bool = True) -> None:
    _lib.check(_lib.lib().zg_set_profiling(graph.resource.handle, 1 if on else 0), "zg_set_profiling")


def get_profile(graph: ResourceGraph) -> dict:
    """{kernel: (total device ms, launches)} since the last read."""
    p = _lib.ZgProfile()
    _lib.check(_lib.lib().zg_get_profile(graph.resource.handle, ctypes.byref(p)), "zg_get_profile")
    return {name: (float(p.ms[i]), int(p.launches[i])) for i, name in enumerate(_lib.PROF_SLOTS)}
