"""The work tables the load derives on the device (chunk lists, hub slots,
PageRank row bins and binned columns) against their definitions, restated
here in numpy from the exported CSR, entry by entry.  The load is buildGraph
(lib/zog/resource_graph.ex:120-129): one load, then many runs.

Cases: tombstones, self-loops and parallel arcs; an R-MAT graph with hubs of
more than 256 successors (several successor chunks) and rows of more than
2048 predecessors (long PageRank rows cut into 4096-arc segments); a grid,
whose PageRank ids follow the vertex ids (bins in vertex order); a resource
of several graphs; a graph with no arcs."""
import ctypes
import sys

import numpy as np
import pytest

from tests import graphs
from zog_amd import _lib
from zog_amd import generators as gen
from zog_amd import resource_graph as rg

pytestmark = pytest.mark.gpu

CHUNK, SMALL, M4, M16, MED, SEG = 256, 4, 32, 256, 2048, 4096
NONE = 0xFFFFFFFF


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def table(res, which):
    L = _lib.lib()
    n = ctypes.c_uint64()
    _lib.check(L.zg_graph_table(res.handle, which, None, 0, ctypes.byref(n)), "zg_graph_table")
    out = np.zeros(max(n.value, 1), np.uint32)
    _lib.check(L.zg_graph_table(res.handle, which, ctypes.c_void_p(out.ctypes.data), out.size, ctypes.byref(n)),
               "zg_graph_table")
    return out[:n.value]


def csr(res):
    V, _, E = res.info()
    a = [np.zeros(V + 1, np.uint32), np.zeros(max(E, 1), np.uint32), np.zeros(V + 1, np.uint32),
         np.zeros(max(E, 1), np.uint32), np.zeros(max(V, 1), np.uint32)]
    rc = _lib.lib().zg_graph_csr(res.handle, *[ctypes.c_void_p(x.ctypes.data) for x in a])
    assert rc in (0, 1)
    return a[0].astype(np.int64), a[2].astype(np.int64), a[3][:E], a[4][:V].astype(np.int64)


def chunks(ptr, counts, width):
    """(vertex, first arc, end) of every chunk and the first chunk of each
    vertex: counts[v] chunks of `width` arcs from ptr[v]."""
    first = np.concatenate([[0], np.cumsum(counts)])
    v = np.repeat(np.arange(len(counts)), counts)
    k = np.arange(first[-1]) - first[:-1][v]
    e0 = ptr[v] + k * width
    e1 = np.minimum(e0 + width, ptr[v + 1])
    return v, e0, e1, first


def expected(V, live, out_ptr, in_ptr, in_col, pr_id):
    ideg, odeg = np.diff(in_ptr), np.diff(out_ptr)
    ex = {}
    # predecessor chunks: live vertices with predecessors
    ic = np.where(live & (ideg > 0), (ideg + CHUNK - 1) // CHUNK, 0)
    v, e0, e1, first = chunks(in_ptr, ic, CHUNK)
    ex[_lib.ZG_TABLE_ICH] = np.concatenate([v, e0, e1])
    ex[_lib.ZG_TABLE_ICH_FIRST] = first
    # successor chunks: every vertex, at least one
    oc = np.maximum(1, (odeg + CHUNK - 1) // CHUNK)
    v, e0, e1, first = chunks(out_ptr, oc, CHUNK)
    ex[_lib.ZG_TABLE_OCH] = np.concatenate([v, e0, e1])
    ex[_lib.ZG_TABLE_OCH_FIRST] = first
    slots = np.where(oc > 1, oc, 0)
    ex[_lib.ZG_TABLE_OHUB] = np.where(oc > 1, np.concatenate([[0], np.cumsum(slots)[:-1]]), NONE)
    # PageRank row bins: rows in PageRank-id order unless the ids follow the
    # vertex ids for at least half of the vertices
    order = np.argsort(pr_id, kind="stable")
    follow = int(np.sum(pr_id[1:] == pr_id[:-1] + 1)) if V > 1 else 0
    seq = order if follow * 2 < V else np.arange(V)
    d = ideg[seq]
    b = np.select([d <= SMALL, d <= M4, d <= M16, d <= MED], [0, 1, 2, 3], 4)
    b = np.where(live[seq], b, 5)
    rows = np.concatenate([seq[b == k] for k in range(5)])
    counts = [int(np.sum(b == k)) for k in range(5)]
    ex[_lib.ZG_TABLE_PR_BINS] = np.array(counts)
    nsm = sum(counts[:4])
    ln = ideg[rows]
    pos = np.concatenate([[0], np.cumsum(ln)])
    ex[_lib.ZG_TABLE_PR_ROWS] = rows[:nsm]
    ex[_lib.ZG_TABLE_PR_DESC] = np.stack([pr_id[rows[:nsm]], pos[:nsm], ln[:nsm], odeg[rows[:nsm]]], 1).reshape(-1)
    lr = rows[nsm:]
    nseg = (ideg[lr] + SEG - 1) // SEG
    seg0 = np.concatenate([[0], np.cumsum(nseg)])
    ex[_lib.ZG_TABLE_PR_LARGE] = np.concatenate([pr_id[lr], odeg[lr], seg0])
    j = np.repeat(np.arange(len(lr)), nseg)
    c = np.arange(seg0[-1]) - seg0[:-1][j]
    s0 = pos[nsm:][j] + c * SEG
    s1 = np.minimum(s0 + SEG, pos[nsm:][j] + ideg[lr][j])
    ex[_lib.ZG_TABLE_PR_SEG] = np.concatenate([j, s0, s1])
    ex[_lib.ZG_TABLE_PR_COL] = np.concatenate([pr_id[in_col[in_ptr[r]:in_ptr[r + 1]]] for r in rows]) \
        if len(in_col) else np.zeros(0)
    assert pos[-1] == len(in_col)
    return ex, counts


def check(res, live, name):
    V, _, E = res.info()
    out_ptr, in_ptr, in_col, pr_id = csr(res)
    live = np.ones(V, bool) if live is None else live.astype(bool)
    ex, counts = expected(V, live, out_ptr, in_ptr, in_col, pr_id)
    for which, want in ex.items():
        got = table(res, which)
        assert got.shape == want.shape, (name, which, got.shape, want.shape)
        assert np.array_equal(got, want.astype(np.uint32)), (name, which)
    say(name, "V", V, "arcs", E, "bins", counts, "successor chunks", int(len(ex[_lib.ZG_TABLE_OCH]) // 3))
    return counts


@pytest.mark.parametrize("gpu_build", [True, False])
def test_tables_multigraph_with_tombstones(gpu_build):
    rng = np.random.default_rng(3)
    n, s, d = graphs.random_multigraph(rng, 4000, 50000)  # self-loops, parallel arcs
    # a hub with many successors and one with many predecessors
    s = np.concatenate([s, np.full(3000, 7, np.uint32), rng.integers(0, n, 5000).astype(np.uint32)])
    d = np.concatenate([d, rng.integers(0, n, 3000).astype(np.uint32), np.full(5000, 9, np.uint32)])
    live = (rng.random(n) > 0.05).astype(np.uint8)
    live[[7, 9]] = 1
    L = _lib.lib()
    L.zg_set_build_threshold(0 if gpu_build else (1 << 62))
    try:
        res = rg.Resource(n, s, d, live)
    finally:
        L.zg_set_build_threshold(1 << 18)
    counts = check(res, live, "multigraph")
    assert counts[4] >= 1  # the 5000-predecessor row is long
    res.release()


def test_tables_rmat_and_grid():
    n, s, d = gen.rmat(16, 1 << 20, 12)
    res = rg.Resource(n, s, d)
    counts = check(res, None, "rmat16")
    assert counts[4] > 0 and counts[0] > 0
    res.release()
    n, s, d = gen.grid_2d(300, 300, 4, delete=2000)
    res = rg.Resource(n, s, d)
    check(res, None, "grid300")
    res.release()


def test_tables_batch_and_empty():
    res = rg.BatchResource(gen.ba_batch(12, 3, 50, 3000, 40))
    check(res, None, "batch of 12")
    res.release()
    res = rg.Resource(10, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    check(res, None, "no arcs")
    res.release()


def test_queue_order_keeps_results():
    # the queue engine on the breadth-first renumbering (default) and on the
    # vertex ids: every row keeps its arc order, so the same bits; and against
    # the oracle
    from oracle import oracle as O
    from zog_amd.soa import SoA
    n, s, d = gen.grid_2d(120, 120, 6, delete=900)
    # a second component and isolated vertices, numbered after the search
    s = np.concatenate([s, np.array([n, n + 1, n + 2], np.uint32)])
    d = np.concatenate([d, np.array([n + 1, n + 2, n], np.uint32)])
    n += 5
    sources = np.concatenate([gen.seeded_sources(n - 5, s[:-3], d[:-3], 60, 9), [n - 5, n - 4]]).astype(np.uint32)
    g = rg.new(SoA.from_arrays("directed", n, s, d))
    rg.set_engine(g, "queue")
    got = rg.betweenness_unweighted_array(g, sources=sources)
    L = _lib.lib()
    _lib.check(L.zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_QUEUE_ORDER, 0.0), "zg_set_tuning")
    plain = rg.betweenness_unweighted_array(g, sources=sources)
    _lib.check(L.zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_QUEUE_ORDER, 1.0), "zg_set_tuning")
    again = rg.betweenness_unweighted_array(g, sources=sources)
    assert np.array_equal(got, plain) and np.array_equal(got, again)
    ref, _ = O.Graph(n, s, d).betweenness(sources=sources, mode=O.SIGMA_EXP)  # counts pass 2^64
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.all(np.abs(got - ref) <= 1e-9 * np.maximum(np.abs(ref), 1e-9 * scale))
    say("queue order: 62 sources, max", float(got.max()))
    rg.destroy(g)
