"""Resource build on the GPU (buildGraph, lib/zog/resource_graph.ex:120-129)
against the host build and the oracle: the same successor rows (newest arc
first, array_graph.zig:88-108, :166-180), predecessor rows (ascending
source, utils.zig:52-73) and PageRank ids, bit for bit, with tombstones,
parallel arcs, self-loops, weights and bad ids."""
import ctypes
import sys

import numpy as np
import pytest

from tests import graphs
from zog_amd import _lib
from zog_amd import generators as gen
from zog_amd import resource_graph as rg

pytestmark = pytest.mark.gpu

DEFAULT_THRESHOLD = 1 << 18


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def export(res):
    V, _, E = res.info()
    a = [np.zeros(V + 1, np.uint32), np.zeros(max(E, 1), np.uint32), np.zeros(V + 1, np.uint32),
         np.zeros(max(E, 1), np.uint32), np.zeros(max(V, 1), np.uint32)]
    rc = _lib.lib().zg_graph_csr(res.handle, *[ctypes.c_void_p(x.ctypes.data) for x in a])
    assert rc in (0, 1)
    return rc, a[0], a[1][:E], a[2], a[3][:E], a[4][:V]


def build(n, s, d, live=None, w=None, gpu=True):
    L = _lib.lib()
    L.zg_set_build_threshold(0 if gpu else (1 << 62))
    try:
        return rg.Resource(n, s, d, live, weights=w)
    finally:
        L.zg_set_build_threshold(DEFAULT_THRESHOLD)


def reference_rows(n, s, d, live):
    """Successor rows newest arc first and predecessor rows by ascending
    source, live arcs only."""
    ok = np.ones(len(s), bool) if live is None else (live[s] & live[d]).astype(bool)
    idx = np.nonzero(ok)[0]
    order = idx[np.lexsort((-idx, s[idx]))]  # by source, newest first
    out_ptr = np.concatenate([[0], np.cumsum(np.bincount(s[order], minlength=n))]).astype(np.uint32)
    out_col = d[order]
    src_of = s[order]
    o2 = np.argsort(out_col, kind="stable")
    in_ptr = np.concatenate([[0], np.cumsum(np.bincount(out_col, minlength=n))]).astype(np.uint32)
    return out_ptr, out_col, in_ptr, src_of[o2]


@pytest.mark.parametrize("case", ["multigraph", "rmat", "sparse_ids", "no_arcs"])
def test_gpu_build_matches_host_and_definition(case):
    say("build", case)
    rng = np.random.default_rng(7)
    live = None
    if case == "multigraph":
        n, s, d = graphs.random_multigraph(rng, 5000, 60000)  # self-loops, parallel arcs
        live = (rng.random(n) > 0.05).astype(np.uint8)
    elif case == "rmat":
        n, s, d = gen.rmat(14, 1 << 18, 5)
    elif case == "sparse_ids":
        n = 3_000_000
        s = rng.integers(0, n, 4000).astype(np.uint32)
        d = rng.integers(0, n, 4000).astype(np.uint32)
    else:
        n, s, d = 10, np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    rg_gpu = build(n, s, d, live, gpu=True)
    rg_host = build(n, s, d, live, gpu=False)
    eg = export(rg_gpu)
    eh = export(rg_host)
    assert eg[0] == 1 and eh[0] == 0
    for x, y in zip(eg[1:], eh[1:]):
        assert np.array_equal(x, y)
    ref = reference_rows(n, s, d, None if live is None else live.astype(bool))
    for x, y in zip(eg[1:5], ref):
        assert np.array_equal(x, y)
    # PageRank ids: decreasing out-degree, ties by id
    deg = np.diff(eg[1].astype(np.int64))
    order = np.lexsort((np.arange(n), -deg))
    pr_id = np.empty(n, np.uint32)
    pr_id[order] = np.arange(n, dtype=np.uint32)
    assert np.array_equal(eg[5], pr_id)
    if case == "multigraph":
        # and against the oracle's linked-list successor walk
        from oracle import oracle as O
        og = O.Graph(n, s, d, live)
        for v in rng.choice(n, 300, replace=False).tolist():
            assert np.array_equal(og.successors(v), eg[2][eg[1][v]:eg[1][v + 1]])
    rg_gpu.release()
    rg_host.release()


def test_gpu_build_weights_and_errors():
    say("build weights")
    rng = np.random.default_rng(9)
    n, s, d = graphs.random_multigraph(rng, 3000, 30000)
    live = (rng.random(n) > 0.05).astype(np.uint8)
    w = rng.uniform(0.1, 2.0, len(s))
    from zog_amd.soa import SoA
    outs = []
    for gpu in (True, False):
        res = build(n, s, d, live, w, gpu=gpu)
        g = rg.ResourceGraph(resource=res, builder=SoA.from_arrays("directed", n, s, d, w))
        outs.append(rg.betweenness_f64_array(g, sources=np.nonzero(live)[0][::7].astype(np.uint32)))
        res.release()
    assert np.allclose(outs[0], outs[1], rtol=1e-12, atol=1e-12)
    # a bad live weight: only the weighted entry point refuses it
    w2 = w.copy()
    w2[np.nonzero(live[s] & live[d])[0][0]] = -1.0
    res = build(n, s, d, live, w2, gpu=True)
    g = rg.ResourceGraph(resource=res, builder=SoA.from_arrays("directed", n, s, d, w2))
    with pytest.raises(_lib.ZogError):
        rg.betweenness_f64_array(g)
    res.release()
    # an id >= n_cap: NodeNotFound from both builds
    s2 = s.copy()
    s2[-1] = n
    for gpu in (True, False):
        with pytest.raises(_lib.NodeNotFoundError):
            build(n, s2, d, gpu=gpu)
