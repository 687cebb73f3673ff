"""Weighted Brandes (betweennessF64, priv/zog/src/centrality.zig:529-667,
:1355-1357) on the GPU against the oracle: the reference's own cases, random
multigraphs with ties, self-loops and tombstones, high-diameter grids,
several source batches, unit weights against the unweighted engine, and the
entry point's error behaviour."""
import sys

import numpy as np
import pytest

from tests import graphs
from tests.test_oracle import brute_bc_weighted
from zog_amd import _lib
from zog_amd import generators as gen
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu

# FP64 sums in another order than the reference's: relative tolerance
RTOL = 1e-9


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make(n, src, dst, w, live=None, kind="directed"):
    return rg.new(SoA.from_arrays(kind, n, src, dst, w), live_mask=live)


def oracle_bc(n, src, dst, w, live=None, sources=None):
    from oracle import oracle as O
    og = O.Graph(n, src, dst, live)
    og.set_weights(w)
    return og.betweenness_f64(sources=sources)


def close(got, ref):
    scale = max(1.0, float(np.abs(ref).max()))
    err = np.abs(got - ref)
    ok = np.all(err <= RTOL * np.maximum(np.abs(ref), 1e-3 * scale))
    return ok, float(err.max())


def test_reference_cases():
    say("weighted reference cases")
    # centrality_test.exs:55-65 (centrality.zig:1436-1454): a chain's middle
    b = SoA.from_list("directed", [("A", "B", 1.0), ("B", "C", 1.0)])
    g = rg.new(b)
    s = rg.betweenness_f64(g)
    assert s["B"] > s["A"] and s["B"] > s["C"]
    rg.destroy(g)
    # centrality_test.exs:67-79: the 10.0 shortcut loses to A->B->C
    b = SoA.from_list("directed", [("A", "B", 1.0), ("B", "C", 1.0), ("A", "C", 10.0)])
    g = rg.new(b)
    assert rg.betweenness_f64(g) == {"A": 0.0, "B": 1.0, "C": 0.0}
    assert rg.betweenness_f64(g, raw=True) == [0.0, 1.0, 0.0]
    rg.destroy(g)
    b = SoA.from_list("directed", [("A", "B", 1.0), ("B", "C", 1.0), ("A", "C", 1.5)])
    g = rg.new(b)
    assert rg.betweenness_f64(g, raw=True) == [0.0, 0.0, 0.0]
    rg.destroy(g)
    # undirected builder: both arcs, scores x0.5 (resource_graph.ex:2633-2637)
    b = SoA.from_list("undirected", [(0, 1, 2.0), (1, 2, 2.0), (0, 2, 5.0)])
    g = rg.new(b)
    assert rg.betweenness_f64(g, raw=True) == [0.0, 1.0, 0.0]
    rg.destroy(g)


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_small(seed):
    # integer weights so that ties are exact; parallel arcs count separately
    rng = np.random.default_rng(900 + seed)
    n, m = 8, 20
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.integers(1, 4, len(src)).astype(np.float64)
    g = make(n, src, dst, w)
    got = rg.betweenness_f64_array(g)
    rg.destroy(g)
    ref = brute_bc_weighted(n, src, dst, w)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (got, ref)


@pytest.mark.parametrize("n,m,wmax", [(300, 3000, 3), (2000, 24000, 5), (1500, 6000, 0)])
def test_random_multigraphs(n, m, wmax):
    # self-loops, parallel arcs, tombstones; wmax > 0: integer weights
    # 1..wmax (many ties), 0: uniform real weights
    say("weighted multigraph", n, m, wmax)
    rng = np.random.default_rng(n + m)
    n, src, dst = graphs.random_multigraph(rng, n, m)
    w = (rng.integers(1, wmax + 1, len(src)).astype(np.float64) if wmax
         else rng.uniform(0.05, 4.0, len(src)))
    live = (rng.random(n) > 0.04).astype(np.uint8)
    g = make(n, src, dst, w, live)
    info = {}
    got = rg.betweenness_f64_array(g, info=info)
    rg.destroy(g)
    ref = oracle_bc(n, src, dst, w, live)
    ok, err = close(got, ref)
    assert ok, err
    assert np.all(got[live == 0] == 0.0)
    assert info["n_sources"] == int(live.sum())


def test_several_source_batches_and_sampled_sources():
    # more sources than one batch holds, repeated and listed sources
    say("weighted batches")
    rng = np.random.default_rng(77)
    n, src, dst = graphs.random_multigraph(rng, 2600, 13000)
    w = rng.uniform(0.5, 2.0, len(src))
    g = make(n, src, dst, w)
    got = rg.betweenness_f64_array(g)
    ok, err = close(got, oracle_bc(n, src, dst, w))
    assert ok, err
    sources = rng.integers(0, n, 300).astype(np.uint32)
    sources[:4] = sources[4]
    got = rg.betweenness_f64_array(g, sources=sources)
    ref = oracle_bc(n, src, dst, w, sources=sources)
    ok, err = close(got, ref)
    assert ok, err
    # a narrower batch width: five batches, the last one partly empty
    L = _lib.lib()
    _lib.check(L.zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_WEIGHTED_BATCH, 64.0))
    info = {}
    got = rg.betweenness_f64_array(g, sources=sources, info=info)
    assert info["batches"] == 5
    ok, err = close(got, ref)
    assert ok, err
    assert L.zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_WEIGHTED_BATCH, 100.0) == _lib.ZG_ERR_INVALID_ARG
    rg.destroy(g)


def test_grid_high_diameter():
    # a 48x48 grid with random weights: hundreds of relax rounds
    say("weighted grid")
    rng = np.random.default_rng(5)
    n, src, dst = gen.grid_2d(48, 48)
    w = rng.uniform(1.0, 2.0, len(src))
    g = make(n, src, dst, w)
    info = {}
    got = rg.betweenness_f64_array(g, info=info)
    rg.destroy(g)
    ok, err = close(got, oracle_bc(n, src, dst, w))
    assert ok, err
    assert info["max_levels"] >= 94


def test_rmat_uniform_weights():
    say("weighted rmat")
    rng = np.random.default_rng(12)
    n, src, dst = gen.rmat(12, 1 << 16, 31)
    w = rng.uniform(0.1, 1.0, len(src))
    sources = rng.choice(n, 400, replace=False).astype(np.uint32)
    g = make(n, src, dst, w)
    got = rg.betweenness_f64_array(g, sources=sources)
    rg.destroy(g)
    ok, err = close(got, oracle_bc(n, src, dst, w, sources=sources))
    assert ok, err


@pytest.mark.parametrize("delta", [0.01, 0.3, float("inf")])
def test_near_far_bound_steps(delta):
    # the bound step changes only the amount of work, never the result
    say("weighted delta", delta)
    rng = np.random.default_rng(21)
    n, src, dst = gen.rmat(11, 1 << 15, 17)
    w = rng.uniform(0.01, 1.0, len(src))
    sources = rng.choice(n, 200, replace=False).astype(np.uint32)
    g = make(n, src, dst, w)
    _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_WEIGHTED_DELTA, delta))
    info = {}
    got = rg.betweenness_f64_array(g, sources=sources, info=info)
    rg.destroy(g)
    ok, err = close(got, oracle_bc(n, src, dst, w, sources=sources))
    assert ok, err
    say("  rounds", info["push_levels"], "arcs relaxed / arcs", info["arcs_relaxed"] / info["arcs_traversed"])


def test_hub_pieces():
    # two hubs of 9,000 and 20,000 arcs (cut into pieces for many
    # wavefronts) on a random graph, with small bound steps so that the hubs
    # are relaxed in rounds of their own
    say("weighted hubs")
    rng = np.random.default_rng(8)
    n = 30000
    hs = np.concatenate([np.zeros(9000, np.uint32), np.full(20000, 1, np.uint32)])
    hd = np.concatenate([rng.choice(n, 9000, replace=False), rng.choice(n, 20000, replace=False)]).astype(np.uint32)
    rs = rng.integers(0, n, 60000).astype(np.uint32)
    rd = rng.integers(0, n, 60000).astype(np.uint32)
    src = np.concatenate([rs[:30000], hs, rs[30000:]])
    dst = np.concatenate([rd[:30000], hd, rd[30000:]])
    w = rng.uniform(0.05, 1.0, len(src))
    sources = np.concatenate([[0, 1], rng.choice(n, 150, replace=False)]).astype(np.uint32)
    g = make(n, src, dst, w)
    ref = oracle_bc(n, src, dst, w, sources=sources)
    for delta in (0.02, float("inf")):
        _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_WEIGHTED_DELTA, delta))
        got = rg.betweenness_f64_array(g, sources=sources)
        ok, err = close(got, ref)
        assert ok, (delta, err)
    rg.destroy(g)


def test_unit_weights_equal_unweighted_engine():
    # all weights 1.0: the same shortest paths and path counts as BFS
    # (the unweighted engine is pinned against the oracle separately)
    say("weighted unit = unweighted")
    n, src, dst = gen.rmat(15, 1 << 19, 3)
    rng = np.random.default_rng(2)
    sources = rng.choice(n, 2048, replace=False).astype(np.uint32)
    g = make(n, src, dst, np.ones(len(src)))
    a = rg.betweenness_f64_array(g, sources=sources)
    b = rg.betweenness_unweighted_array(g, sources=sources)
    rg.destroy(g)
    ok, err = close(a, b)
    assert ok, err


def test_zero_weights_on_a_dag_match_brute_force():
    # weight-0 arcs in an acyclic graph: every tight arc counts (the
    # reference's result here depends on its heap order, so the check is the
    # brute-force definition)
    rng = np.random.default_rng(41)
    n, m = 9, 24
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    src = np.minimum(a, b)[keep].astype(np.uint32)
    dst = np.maximum(a, b)[keep].astype(np.uint32)
    w = rng.integers(0, 3, len(src)).astype(np.float64)
    g = make(n, src, dst, w)
    got = rg.betweenness_f64_array(g)
    rg.destroy(g)
    ref = brute_bc_weighted(n, src, dst, w)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (got, ref)


def test_errors_and_edges():
    say("weighted errors")
    src = np.array([0, 1, 2], np.uint32)
    dst = np.array([1, 2, 0], np.uint32)
    # a negative or NaN weight on a live arc
    for bad in (-1.0, float("nan"), float("inf")):
        g = make(3, src, dst, np.array([1.0, bad, 1.0]))
        with pytest.raises(_lib.ZogError) as e:
            rg.betweenness_f64_array(g)
        assert e.value.code == _lib.ZG_ERR_INVALID_ARG
        assert np.allclose(rg.betweenness_unweighted_array(g), [1.0, 1.0, 1.0])  # unweighted still works
        rg.destroy(g)
    # ... but not when the arc touches a tombstone
    g = make(3, src, dst, np.array([1.0, -1.0, 1.0]), live=np.array([1, 1, 0], np.uint8))
    assert rg.betweenness_f64_array(g).tolist() == [0.0, 0.0, 0.0]
    rg.destroy(g)
    # a resource created without weights
    res = rg.Resource(3, src, dst)
    g = rg.ResourceGraph(resource=res, builder=SoA.from_arrays("directed", 3, src, dst))
    with pytest.raises(_lib.ZogError):
        rg.betweenness_f64_array(g)
    rg.destroy(g)
    # a non-live source, an empty graph, a graph without arcs
    g = make(3, src, dst, np.ones(3), live=np.array([1, 0, 1], np.uint8))
    with pytest.raises(_lib.NodeNotFoundError):
        rg.betweenness_f64_array(g, sources=np.array([1], np.uint32))
    rg.destroy(g)
    g = make(0, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    assert rg.betweenness_f64_array(g).tolist() == []
    rg.destroy(g)
    g = make(4, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    assert rg.betweenness_f64_array(g).tolist() == [0.0] * 4
    rg.destroy(g)
