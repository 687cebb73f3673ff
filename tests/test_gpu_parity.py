"""Parity of the gfx950 path (through the C ABI) with the CPU oracle.

Bars: BFS depths and path counts bit-exact; BC within 1e-9 relative (per
vertex, with an absolute floor of 1e-9 times the largest score); PageRank
L1 <= 1e-6 against the FP64 oracle at the same iteration count.
"""
import sys

import numpy as np
import pytest

from tests import graphs
from zog_amd import generators as gen
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9
PR_L1 = 1e-6


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make(n, src, dst, live=None, groups=None):
    g = rg.new(SoA.from_arrays("directed", n, src, dst), live_mask=live)
    if groups:
        rg.set_batch_groups(g, groups)
    return g


def assert_bc_close(got, ref):
    scale = max(1.0, float(np.abs(ref).max()) if len(ref) else 1.0)
    err = np.abs(got - ref)
    ok = err <= BC_RTOL * np.maximum(np.abs(ref), 1e-9 * scale)
    assert ok.all(), f"max err {err.max()} at {int(np.argmax(err))}"


def check_bc(n, src, dst, sources=None, live=None, groups=None, mode=0):
    from oracle import oracle as O
    say("bc case", n, len(src), None if sources is None else len(sources), groups)
    og = O.Graph(n, src, dst, live)
    ref, ovf = og.betweenness(sources=sources, mode=mode)
    g = make(n, src, dst, live, groups)
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert_bc_close(got, ref)
    rg.destroy(g)
    return got, ref, info


def check_pr(n, src, dst, live=None, iters=50, tol=0.0, damping=0.85, precision=32):
    from oracle import oracle as O
    say("pr case", n, len(src), iters, tol, precision)
    og = O.Graph(n, src, dst, live)
    ref, it_ref, _ = og.pagerank(damping=damping, max_iterations=iters, tolerance=tol)
    g = make(n, src, dst, live)
    info = {}
    got = rg.pagerank_array(g, damping, iters, tol, precision=precision, info=info)
    rg.destroy(g)
    return got, ref, it_ref, info


# --- reference cases through the Zog.ResourceGraph mirror ------------------

def test_reference_cases_resource_api():
    # resource_graph_test.exs:41-54 path graph (undirected)
    g = rg.new(SoA.undirected().add_edge("A", "B", 1.0).add_edge("B", "C", 1.0))
    s = rg.betweenness_unweighted(g)
    assert s["B"] > s["A"] and s["B"] > s["C"] and s["B"] == 1.0
    rg.destroy(g)
    # centrality_test.exs:26-39 star: centre 3.0, leaves exactly 0.0
    g = rg.new(SoA.undirected().add_edge("center", "a").add_edge("center", "b").add_edge("center", "c"))
    s = rg.betweenness_unweighted(g)
    assert s == {"center": 3.0, "a": 0.0, "b": 0.0, "c": 0.0}
    p = rg.pagerank(g)
    assert all(p["center"] > p[k] for k in "abc")
    rg.destroy(g)
    # resource_graph_test.exs:459-541: several algorithms on one resource, raw lists
    b = SoA.undirected().add_edge("A", "B").add_edge("B", "C").add_edge("C", "A").add_edge("B", "D")
    g = rg.new(b)
    assert len(rg.betweenness_unweighted(g)) == 4 and len(rg.pagerank(g)) == 4
    raw = rg.pagerank(g, raw=True)
    assert isinstance(raw, list) and len(raw) == 4 and all(isinstance(x, float) for x in raw)
    assert isinstance(rg.betweenness_unweighted(g, raw=True), list)
    assert rg.betweenness_centrality(g) == rg.betweenness_unweighted(g)
    rg.destroy(g)
    with pytest.raises(rg._lib.ZogError):
        rg.pagerank(g)
    # centrality_test.exs:12-24, 138-162 (directed)
    g = rg.new(SoA.directed().add_edge("A", "B").add_edge("B", "C"))
    s = rg.betweenness_unweighted(g)
    assert s == {"A": 0.0, "B": 1.0, "C": 0.0}
    assert rg.pagerank(g, damping=0.99)["A"] > 0.0
    rg.destroy(g)
    g = rg.new(SoA.directed().add_edge("A", "B").add_edge("C", "B"))
    p = rg.pagerank(g)
    assert p["B"] > p["A"] and p["B"] > p["C"]
    rg.destroy(g)


def test_empty_and_single_node():
    g = rg.new(SoA.directed())
    assert rg.betweenness_unweighted(g) == {} and rg.pagerank(g) == {}
    rg.destroy(g)
    g = rg.new(SoA.directed().add_node("A"))
    assert rg.betweenness_unweighted(g) == {"A": 0.0}
    assert rg.pagerank(g) == {"A": 1.0}
    rg.destroy(g)


def test_integer_labels_edgelist(tmp_path):
    # resource_graph_test.exs:545-571
    p = tmp_path / "e.txt"
    p.write_text("0 1 1.5\n1 2 2.0\n2 0 1.0\n")
    g = rg.read_edgelist(str(p), integer_labels=True)
    assert g.builder.integer_labels and g.builder.next_id == 3
    pr = rg.pagerank(g)
    assert set(pr) == {0, 1, 2}
    assert len(rg.pagerank(g, raw=True)) == 3
    rg.destroy(g)
    # sparse ids: placeholders are live, isolated and dangling
    p.write_text("0 5\n5 2\n")
    g = rg.read_edgelist(str(p), integer_labels=True, directed=False)
    from oracle import oracle as O
    og = O.Graph(6, *graphs.arcs([(0, 5), (5, 0), (2, 5), (5, 2)]))
    assert np.allclose(rg.betweenness_unweighted(g, raw=True), 0.5 * og.betweenness()[0], rtol=1e-12)
    ref, _, _ = og.pagerank()
    assert np.abs(np.array(rg.pagerank(g, raw=True)) - ref).sum() < PR_L1
    rg.destroy(g)


def test_node_not_found():
    with pytest.raises(rg._lib.NodeNotFoundError):
        make(3, np.array([0, 5], np.uint32), np.array([1, 1], np.uint32))
    g = make(3, *graphs.arcs([(0, 1)]))
    with pytest.raises(rg._lib.NodeNotFoundError):
        rg.betweenness_unweighted_array(g, sources=[7])
    rg.destroy(g)


# --- Brandes parity -----------------------------------------------------------

def test_wiki_vote_fixture():
    n, src, dst = graphs.wiki_vote()
    check_bc(n, src, dst)


@pytest.mark.parametrize("seed", range(6))
def test_random_multigraphs(seed):
    rng = np.random.default_rng(500 + seed)
    n = int(rng.integers(2, 300))
    m = int(rng.integers(0, 6 * n))
    n, src, dst = graphs.random_multigraph(rng, n, m)
    check_bc(n, src, dst, groups=int(rng.choice([1, 2, 4])))


def test_tombstoned_nodes():
    rng = np.random.default_rng(9)
    n, src, dst = graphs.random_multigraph(rng, 200, 900)
    live = (rng.random(n) > 0.2).astype(np.uint8)
    check_bc(n, src, dst, live=live)
    got, ref, _, _ = check_pr(n, src, dst, live=live)
    assert np.abs(got - ref).sum() <= PR_L1
    assert np.all(got[live == 0] == 0.0)


def test_er_parity_config():
    n, src, dst = gen.erdos_renyi_directed(10_000, 8.0 / 9_999, gen.SEED_BASE + 1)
    got, ref, info = check_bc(n, src, dst)
    assert info["n_sources"] == n


def test_hubs_split_chunks():
    # in- and out-hubs far above the 256-arc chunk, parallel arcs on the hubs
    rng = np.random.default_rng(21)
    n = 3000
    hub_in = np.full(2000, 7, np.uint32)
    hub_in_src = rng.integers(0, n, 2000).astype(np.uint32)
    hub_out_dst = rng.integers(0, n, 1500).astype(np.uint32)
    hub_out = np.full(1500, 11, np.uint32)
    _, s2, d2 = graphs.random_multigraph(rng, n, 9000)
    src = np.concatenate([hub_in_src, hub_out, s2, hub_out[:300]])
    dst = np.concatenate([hub_in, hub_out_dst, d2, hub_out_dst[:300]])
    check_bc(n, src, dst, sources=np.arange(0, n, 3, dtype=np.uint32), groups=4)


def test_rmat_small_sources_subset():
    n, src, dst = gen.rmat(14, 1 << 18, gen.SEED_BASE + 77)
    sources = gen.seeded_sources(n, src, dst, 300, 5)
    check_bc(n, src, dst, sources=sources, groups=4)


def test_batch_widths_and_source_order_agree():
    n, src, dst = gen.rmat(11, 1 << 14, 3)
    sources = np.arange(n, dtype=np.uint32)
    outs = []
    for gr, order in ((1, 1), (3, 0), (16, 1), (4, 0), (5, 2), (64, 2)):
        g = make(n, src, dst, groups=gr)
        rg.set_source_order(g, order)
        outs.append(rg.betweenness_unweighted_array(g, sources=sources))
        rg.destroy(g)
    for o in outs[1:]:
        assert_bc_close(o, outs[0])


@pytest.mark.parametrize("alpha", [0.0, 1e-9, 8.0])
@pytest.mark.parametrize("graph", [1, 0])
def test_push_pull_directions(alpha, graph):
    # all-pull, all-push and the default switch give the oracle's BC, depths
    # and path counts (hubs and parallel arcs included)
    from oracle import oracle as O
    n, src, dst = gen.rmat(12, 1 << 15, 1234)
    src = np.concatenate([src, src[:500]])
    dst = np.concatenate([dst, dst[:500]])
    sources = gen.seeded_sources(n, src, dst, 200, 8)
    og = O.Graph(n, src, dst)
    ref, _ = og.betweenness(sources=sources)
    _, _, arcs = og.betweenness(sources=sources, return_arcs=True)
    g = make(n, src, dst)
    rg.set_push_alpha(g, alpha)
    from zog_amd import _lib
    _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_GRAPH, float(graph)), "zg_set_tuning")
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert_bc_close(got, ref)
    assert info["arcs_traversed"] == int(arcs.sum())
    if alpha == 0.0:
        assert info["push_levels"] == 0
    if alpha == 1e-9:
        assert info["pull_levels"] == 0
    depth, sigma, _, _ = rg.bc_paths(g, sources[:64], max_levels=64)
    for i, s in enumerate(sources[:64].tolist()):
        d_ref, su, *_ = og.bfs_paths(s)
        assert np.array_equal(depth[i], d_ref)
        r = d_ref != O.NONE_U32
        assert np.array_equal(sigma[i][r], su[r].astype(np.float64))
    rg.destroy(g)


@pytest.mark.parametrize("groups", [1, 4, 16, 64])
def test_batch_of_uneven_graphs(groups):
    # one resource holding graphs of uneven size: every graph's BC equals the
    # oracle on that graph alone
    from oracle import oracle as O
    rng = np.random.default_rng(77)
    graphs_ = [gen.barabasi_albert(int(n), 4, 100 + i) for i, n in enumerate([5, 70, 900, 130, 2500, 40])]
    graphs_.append(graphs.random_multigraph(rng, 300, 1200))
    graphs_.append((1, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    info = {}
    outs = rg.betweenness_unweighted_batch(graphs_, info=info, groups=groups)
    assert info["n_sources"] == sum(g[0] for g in graphs_)
    for (n, src, dst), got in zip(graphs_, outs):
        ref, _ = O.Graph(n, src, dst).betweenness()
        assert_bc_close(got, ref)


def test_batch_groups_spanning_several_graphs():
    # one handle over several graphs; sources of one graph, then sources
    # scattered over all graphs (a group spans several graphs and the
    # per-group state grows), then depths and path counts of scattered groups
    from oracle import oracle as O
    graphs_ = [gen.barabasi_albert(n, 4, 300 + i) for i, n in enumerate([200, 1500, 60, 800, 5])]
    res = rg.BatchResource(graphs_)
    off = res.offsets
    n = int(off[-1])
    src = np.concatenate([g[1] + np.uint32(o) for g, o in zip(graphs_, off[:-1])])
    dst = np.concatenate([g[2] + np.uint32(o) for g, o in zip(graphs_, off[:-1])])
    og = O.Graph(n, src, dst)
    g = rg.ResourceGraph(resource=res, builder=SoA.from_arrays("directed", n, src, dst))
    rng = np.random.default_rng(5)
    one = np.arange(off[1], off[2], dtype=np.uint32)
    scattered = rng.permutation(n)[:700].astype(np.uint32)
    for order in (1, 0, 2):
        rg.set_source_order(g, order)
        for sources in (one, scattered, one):
            ref, _, arcs = og.betweenness(sources=sources, return_arcs=True)
            info = {}
            got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
            assert_bc_close(got, ref)
            assert info["arcs_traversed"] == int(arcs.sum())
    pick = scattered[:70]
    depth, sigma, _, _ = rg.bc_paths(g, pick, max_levels=64)
    for i, s in enumerate(pick.tolist()):
        d_ref, su, *_ = og.bfs_paths(s)
        assert np.array_equal(depth[i], d_ref)
        r = d_ref != O.NONE_U32
        assert np.array_equal(sigma[i][r], su[r].astype(np.float64))
    res.release()


@pytest.mark.parametrize("case", ["rmat", "multigraph", "hubs", "grid"])
def test_queue_engine(case):
    # the per-source queue engine against the oracle on the same inputs as the
    # 64-lane engine: parallel arcs, self-loops, hubs, counts beyond 2^64
    from oracle import oracle as O
    rng = np.random.default_rng(31)
    mode = O.SIGMA_U64
    if case == "rmat":
        n, src, dst = gen.rmat(12, 1 << 15, 4321)
        sources = gen.seeded_sources(n, src, dst, 300, 3)
    elif case == "multigraph":
        n, src, dst = graphs.random_multigraph(rng, 400, 2500)
        sources = None
    elif case == "hubs":
        n = 2000
        src = np.concatenate([rng.integers(0, n, 3000), np.full(1200, 9), rng.integers(0, n, 5000)]).astype(np.uint32)
        dst = np.concatenate([np.full(3000, 4), rng.integers(0, n, 1200), rng.integers(0, n, 5000)]).astype(np.uint32)
        sources = np.arange(0, n, 7, dtype=np.uint32)
    else:
        n, src, dst = gen.grid_2d(200, 200, seed=5, delete=2000)
        sources = gen.seeded_sources(n, src, dst, 70, 9)
        mode = O.SIGMA_SCALED
    og = O.Graph(n, src, dst)
    ref, _ = og.betweenness(sources=sources, mode=mode)
    g = make(n, src, dst)
    rg.set_engine(g, "queue")
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert_bc_close(got, ref)
    _, _, arcs = og.betweenness(sources=sources, mode=mode, return_arcs=True)
    assert info["arcs_traversed"] == int(arcs.sum())
    rg.destroy(g)


def test_depth_sigma_bit_exact():
    from oracle import oracle as O
    n, src, dst = gen.rmat(12, 1 << 15, 99)
    sources = gen.seeded_sources(n, src, dst, 70, 1)
    og = O.Graph(n, src, dst)
    g = make(n, src, dst)
    depth, sigma, shift, levels = rg.bc_paths(g, sources, max_levels=64)
    for i, s in enumerate(sources.tolist()):
        d_ref, su, sf, sh, arcs, ovf = og.bfs_paths(s)
        assert not ovf
        assert np.array_equal(depth[i], d_ref)
        reached = d_ref != O.NONE_U32
        assert np.array_equal(sigma[i][reached], su[reached].astype(np.float64))
        assert np.all(sigma[i][reached] == su[reached])
        assert levels[i] == len(sh)
    rg.destroy(g)


def test_grid_overflowing_path_counts():
    # sigma exceeds u64 (and the FP64 range at the far corner): compare with
    # the oracle's scaled-sigma mode and check the shifted counts
    from oracle import oracle as O
    n, src, dst = gen.grid_2d(300, 300)
    sources = np.array([0, 299, 150 * 300 + 150, 89999], np.uint32)
    check_bc(n, src, dst, sources=sources, mode=O.SIGMA_SCALED)
    og = O.Graph(n, src, dst)
    g = make(n, src, dst)
    depth, sigma, shift, levels = rg.bc_paths(g, sources[:1], max_levels=700)
    d_ref, _, sf_ref, sh_ref, _, _ = og.bfs_paths(0)
    assert np.array_equal(depth[0], d_ref)
    assert np.array_equal(shift[0][: len(sh_ref)], sh_ref)
    assert np.allclose(sigma[0], sf_ref, rtol=1e-11)
    rg.destroy(g)


def test_lane_engine_deeper_than_level_table():
    # the 64-lane engine on a graph with ~240 BFS levels: the backward pass
    # keeps per-level lane masks for the first 32 levels and falls back to
    # the deepest-level filter past them; path counts pass 2^64 (scaled mode)
    from oracle import oracle as O
    n, src, dst = gen.grid_2d(120, 120, seed=7, delete=800)
    sources = np.array([0, 119, 60 * 120 + 60, 120 * 120 - 1, 5, 777, 9000, 14000], np.uint32)
    say("lane engine deep grid", n, len(src), len(sources))
    ref, _ = O.Graph(n, src, dst).betweenness(sources=sources, mode=O.SIGMA_SCALED)
    g = make(n, src, dst)
    rg.set_engine(g, "lanes")
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert info["max_levels"] > 64
    assert_bc_close(got, ref)
    rg.destroy(g)


def test_queue_engine_counts_beyond_one_level_fp64_range():
    # 1600x1600 grid: a level's path counts span more than the FP64 range
    # (1 on the edges, ~2^1590 mid-grid from a corner); the queue engine's
    # per-vertex exponents against the oracle's, BC finite, depths equal
    from oracle import oracle as O
    R = 1600
    n, src, dst = gen.grid_2d(R, R)
    sources = np.array([0, R - 1, (R // 2) * R + R // 2, R * R - 1], np.uint32)
    og = O.Graph(n, src, dst)
    ref, _ = og.betweenness(sources=sources, mode=O.SIGMA_EXP)
    assert np.all(np.isfinite(ref))
    g = make(n, src, dst)
    rg.set_engine(g, "queue")
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    rg.destroy(g)
    assert np.all(np.isfinite(got))
    assert_bc_close(got, ref)
    depth = og.bfs_paths(0)[0]
    assert info["max_levels"] == int(depth.max()) + 1


def test_dependency_checksum_rmat():
    # sum_v BC(v) over a source set = sum over sources of sum_t (dist(s,t) - 1)
    n, src, dst = gen.rmat(13, 1 << 17, 42)
    sources = gen.seeded_sources(n, src, dst, 200, 2)
    g = make(n, src, dst)
    bc = rg.betweenness_unweighted_array(g, sources=sources)
    depth, _, _, _ = rg.bc_paths(g, sources, max_levels=64)
    reached = depth != 0xFFFFFFFF
    checksum = float(np.where(reached & (depth > 0), depth.astype(np.float64) - 1.0, 0.0).sum())
    assert abs(bc.sum() - checksum) <= 1e-9 * checksum
    rg.destroy(g)


# --- PageRank parity ------------------------------------------------------------

@pytest.mark.parametrize("precision", [32, 64])
def test_pagerank_fixed_iterations(precision):
    n, src, dst = gen.erdos_renyi_directed(10_000, 8.0 / 9_999, gen.SEED_BASE + 1)
    got, ref, it_ref, info = check_pr(n, src, dst, iters=60, tol=0.0, precision=precision)
    assert info["iterations"] == it_ref == 60
    l1 = np.abs(got - ref).sum()
    assert l1 <= (PR_L1 if precision == 32 else 1e-12), l1


def test_pagerank_tolerance_stop_and_residual():
    n, src, dst = gen.rmat(14, 1 << 18, gen.SEED_BASE + 7)
    got, ref, it_ref, info = check_pr(n, src, dst, iters=100, tol=1e-6)
    assert info["iterations"] == it_ref
    assert np.abs(got - ref).sum() <= PR_L1
    assert info["last_l1"] < 1e-6
    # a-posteriori bound from the FP64 residual of the returned vector
    assert info["error_bound_l1"] < 1e-4
    assert abs(got.sum() - 1.0) < 1e-5


def test_pagerank_hub_rows_and_wiki_vote():
    n, src, dst = graphs.wiki_vote()
    got, ref, it_ref, info = check_pr(n, src, dst, iters=100, tol=1e-4)
    assert info["iterations"] == it_ref and np.abs(got - ref).sum() <= PR_L1
    # rows far beyond one segment (4096 predecessors)
    rng = np.random.default_rng(4)
    src = np.concatenate([rng.integers(0, 20000, 30000), rng.integers(0, 20000, 40000)]).astype(np.uint32)
    dst = np.concatenate([np.full(30000, 5), rng.integers(0, 20000, 40000)]).astype(np.uint32)
    got, ref, it_ref, info = check_pr(20000, src, dst, iters=40, tol=0.0)
    assert np.abs(got - ref).sum() <= PR_L1


def test_pagerank_every_row_bin():
    # predecessor counts in every bin and at its bounds: 0, <=4, <=32, <=256,
    # <=2048, longer rows; parallel arcs and self-loops included
    rng = np.random.default_rng(12)
    n = 6000
    degs = [0, 3, 4, 5, 16, 17, 32, 33, 64, 65, 200, 256, 257, 1500, 2048, 2049, 4096, 4097, 9000]
    src, dst = [], []
    for i, dgr in enumerate(degs):
        src.append(rng.integers(0, n, dgr))
        dst.append(np.full(dgr, 100 + i))
    s2 = rng.integers(0, n, 30000)
    d2 = rng.integers(0, n, 30000)
    src = np.concatenate(src + [s2, np.array([7, 7, 7])]).astype(np.uint32)
    dst = np.concatenate(dst + [d2, np.array([7, 7, 8])]).astype(np.uint32)
    for prec, bar in ((32, PR_L1), (64, 1e-12)):
        got, ref, it_ref, info = check_pr(n, src, dst, iters=30, tol=0.0, precision=prec)
        assert np.abs(got - ref).sum() <= bar


def test_pagerank_mixed_precision_tight_tolerance():
    # a tolerance below the FP32 rounding floor: FP32 Kahan down to 1e-6, then
    # FP64 from that iterate; the result meets the FP64 oracle's fixed point
    from oracle import oracle as O
    n, src, dst = gen.grid_2d(60, 60, seed=3, delete=300)
    og = O.Graph(n, src, dst)
    ref, it_ref, l1_ref = og.pagerank(0.85, 1000, 1e-10)
    g = make(n, src, dst)
    info = {}
    got = rg.pagerank_array(g, 0.85, 1000, 1e-10, precision=96, info=info)
    assert info["last_l1"] < 1e-10 and info["fp64_iterations"] > 0
    assert abs(info["iterations"] - it_ref) <= 2
    assert np.abs(got - ref).sum() <= 1e-8
    assert info["error_bound_l1"] <= 1e-8
    rg.destroy(g)


def test_pagerank_max_iterations_zero_and_damping():
    n, src, dst = gen.erdos_renyi_directed(500, 0.01, 3)
    got, ref, it_ref, info = check_pr(n, src, dst, iters=0)
    assert info["iterations"] == 0 and np.allclose(got, 1.0 / n, rtol=1e-7)
    got, ref, it_ref, info = check_pr(n, src, dst, iters=100, tol=0.0, damping=0.99, precision=64)
    assert np.abs(got - ref).sum() <= 1e-12


def test_autotuner_keeps_results():
    # every candidate launch shape gives the oracle's BC; the tuner leaves the
    # fastest one set on the handle
    from oracle import oracle as O
    from zog_amd import autotune
    n, src, dst = gen.rmat(12, 1 << 15, 4321)
    sources = gen.seeded_sources(n, src, dst, 300, 2)
    ref, _ = O.Graph(n, src, dst).betweenness(sources=sources)
    g = make(n, src, dst)
    out = autotune.tune(g.resource, sources[:128], groups=(1, 2), alphas=(0.0, 8.0), pull_grids=(64, 0),
                        orders=(1, 2), repeats=1)
    assert len(out["trials"]) == 16 and out["best"] in out["trials"]
    got = rg.betweenness_unweighted_array(g, sources=sources)
    assert_bc_close(got, ref)
    rg.destroy(g)


def test_batched_sweep_work_counters():
    # row visits, row arcs and per-lane tree arcs of the 64-lane sweep (the
    # inputs of its byte model) against the same counts taken from the
    # oracle's BFS depths, sources swept in the given order, 64 per group
    from oracle import oracle as O
    from zog_amd import _lib
    n, src, dst = gen.rmat(11, 1 << 14, 71)
    src = np.concatenate([src, src[:200], np.array([5, 5], np.uint32)])  # parallel arcs, a self-loop
    dst = np.concatenate([dst, dst[:200], np.array([5, 9], np.uint32)])
    sources = gen.seeded_sources(n, src, dst, 150, 3)
    og = O.Graph(n, src, dst)
    outdeg = np.bincount(src, minlength=n).astype(np.int64)
    visits = row_arcs = tree = 0
    for g0 in range(0, len(sources), 64):
        depths = np.stack([og.bfs_paths(int(s))[0] for s in sources[g0:g0 + 64]])
        reached = depths != O.NONE_U32
        for L in range(1, int(depths[reached].max()) + 1):
            on = np.any(depths == L, axis=0)
            visits += int(on.sum())
            row_arcs += int(outdeg[on].sum())
        for d in depths:
            r = (d[src] != O.NONE_U32) & (d[src] >= 1)  # the backward pass starts at level 1
            tree += int(np.count_nonzero(r & (d[dst] == d[src] + 1)))
    g = make(n, src, dst, groups=2)
    rg.set_source_order(g, 0)
    _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_COUNT_TREE_ARCS, 1.0), "zg_set_tuning")
    info = {}
    rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert (info["row_visits"], info["row_arcs"], info["tree_arcs"]) == (visits, row_arcs, tree)
    _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_COUNT_TREE_ARCS, 0.0), "zg_set_tuning")
    info = {}
    rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert info["tree_arcs"] == 0 and info["row_visits"] == visits
    rg.destroy(g)


def _lane_order(g, on):
    rg.set_lane_order(g, on)


def test_lane_view_matches_resource_ids():
    # the 64-lane engine on the degree-ordered view and on the resource's own
    # ids: BC against the oracle (tombstones, parallel arcs, a self-loop),
    # per-source dependency totals and source weights, work counters, and
    # the kernel profile read through the parent handle
    from oracle import oracle as O
    from zog_amd import pivots
    rng = np.random.default_rng(404)
    n, src, dst = gen.rmat(12, 1 << 15, 404)
    src = np.concatenate([src, src[:300], np.array([3, 3], np.uint32)])
    dst = np.concatenate([dst, dst[:300], np.array([3, 8], np.uint32)])
    live = (rng.random(n) > 0.1).astype(np.uint8)
    og = O.Graph(n, src, dst, live)
    sources = np.flatnonzero(live)[::5].astype(np.uint32)
    ref, _, arcs = og.betweenness(sources=sources, return_arcs=True)
    w = rng.random(len(sources)) + 0.5
    out = {}
    for on in (True, False):
        g = make(n, src, dst, live, groups=4)
        _lane_order(g, on)
        rg.set_profiling(g, True)
        info = {}
        got = rg.betweenness_unweighted_array(g, sources=sources, info=info)
        prof = rg.get_profile(g)
        assert_bc_close(got, ref)
        assert info["arcs_traversed"] == int(arcs.sum())
        assert sum(c for _, c in prof.values()) > 0
        bc_w, tot = pivots.bc_weighted(g, sources, w, want_bc=True, want_totals=True)
        out[on] = (got, bc_w, tot)
        assert np.all(got[live == 0] == 0.0)
        rg.destroy(g)
    assert_bc_close(out[True][1], out[False][1])
    np.testing.assert_allclose(out[True][2], out[False][2], rtol=1e-12)


def test_lane_view_on_a_batch_of_graphs():
    # component ranges survive the renumbering: every graph of a batched
    # resource keeps the oracle's BC, with groups spanning several graphs
    from oracle import oracle as O
    graphs_ = [gen.barabasi_albert(n, 3, 900 + i) for i, n in enumerate([300, 40, 1200, 7, 650])]
    res = rg.BatchResource(graphs_)
    off = res.offsets
    n = int(off[-1])
    src = np.concatenate([g[1] + np.uint32(o) for g, o in zip(graphs_, off[:-1])])
    dst = np.concatenate([g[2] + np.uint32(o) for g, o in zip(graphs_, off[:-1])])
    og = O.Graph(n, src, dst)
    g = rg.ResourceGraph(resource=res, builder=SoA.from_arrays("directed", n, src, dst))
    scattered = np.random.default_rng(6).permutation(n)[:500].astype(np.uint32)
    for sources in (None, scattered):
        ref, _ = og.betweenness(sources=sources)
        for on in (True, False):
            _lane_order(g, on)
            assert_bc_close(rg.betweenness_unweighted_array(g, sources=sources), ref)
    res.release()


def test_spans_end_at_the_last_vertex_with_an_arc():
    # vertices without arcs after (and among) the others, some of them sources
    # (the group's span then reaches past the last vertex with an arc), in
    # one graph and in a batch of graphs, on the view and on the resource's
    # own ids
    from oracle import oracle as O
    rng = np.random.default_rng(77)
    n = 3000
    _, src, dst = graphs.random_multigraph(rng, 1800, 9000)
    src = (src * 3 // 2).astype(np.uint32)  # ids up to ~2700 with gaps; 2701.. have no arcs
    dst = (dst * 3 // 2).astype(np.uint32)
    og = O.Graph(n, src, dst)
    sources = np.concatenate([np.arange(0, n, 7), [n - 1, n - 2, 1]]).astype(np.uint32)
    ref, _ = og.betweenness(sources=sources)
    for on in (True, False):
        g = make(n, src, dst, groups=2)
        _lane_order(g, on)
        assert_bc_close(rg.betweenness_unweighted_array(g, sources=sources), ref)
        assert_bc_close(rg.betweenness_unweighted_array(g), og.betweenness()[0])
        rg.destroy(g)
    graphs_ = []
    for i, m in enumerate([40, 700, 3, 250]):
        k, s2, d2 = gen.barabasi_albert(m + 4, 2, 40 + i)
        graphs_.append((k + 5 + i, s2, d2))  # 5 + i vertices without arcs at each graph's end
    res = rg.BatchResource(graphs_)
    off = res.offsets
    nb = int(off[-1])
    bs = np.concatenate([gg[1] + np.uint32(o) for gg, o in zip(graphs_, off[:-1])])
    bd = np.concatenate([gg[2] + np.uint32(o) for gg, o in zip(graphs_, off[:-1])])
    ob = O.Graph(nb, bs, bd)
    gb = rg.ResourceGraph(resource=res, builder=SoA.from_arrays("directed", nb, bs, bd))
    tails = np.array([o - 1 for o in off[1:]], np.uint32)  # every graph's last vertex: no arcs
    picks = np.concatenate([tails, np.random.default_rng(3).permutation(nb)[:150].astype(np.uint32)])
    for sources in (None, picks):
        refb, _ = ob.betweenness(sources=sources)
        for on in (True, False):
            _lane_order(gb, on)
            assert_bc_close(rg.betweenness_unweighted_array(gb, sources=sources), refb)
    depth, sigma, _, _ = rg.bc_paths(gb, picks[:8], max_levels=64)
    for i, s0 in enumerate(picks[:8].tolist()):
        d_ref, su, *_ = ob.bfs_paths(s0)
        assert np.array_equal(depth[i], d_ref)
    res.release()


@pytest.mark.parametrize("precision", [32, 64])
def test_pagerank_block_shapes_and_lds_copy(precision):
    # every iteration-block shape, with and without the LDS copy of the
    # most-gathered values, meets the oracle, repeats bit for bit, and
    # agrees with the others (only the per-block FP64 order of the dangling
    # mass differs); rows longer than one segment are included
    from oracle import oracle as O
    from zog_amd import _lib
    n, src, dst = gen.rmat(15, 1 << 19, gen.SEED_BASE + 21)
    rng = np.random.default_rng(6)
    src = np.concatenate([src, rng.integers(0, n, 9000)]).astype(np.uint32)
    dst = np.concatenate([dst, np.full(9000, 11)]).astype(np.uint32)
    ref, _, _ = O.Graph(n, src, dst).pagerank(0.85, 40, 0.0)
    g = make(n, src, dst)
    h = g.resource.handle
    outs = []
    for block, cache in ((256, None), (512, None), (1024, None), (1024, 0), (1024, 1000), (256, 0)):
        say("pagerank shape", block, cache, precision)
        _lib.check(_lib.lib().zg_set_tuning(h, _lib.ZG_TUNE_PR_BLOCK, float(block)), "zg_set_tuning")
        _lib.check(_lib.lib().zg_set_tuning(h, _lib.ZG_TUNE_PR_CACHE, float(1 << 30 if cache is None else cache)),
                   "zg_set_tuning")
        a = rg.pagerank_array(g, 0.85, 40, 0.0, precision=precision)
        b = rg.pagerank_array(g, 0.85, 40, 0.0, precision=precision)
        assert np.array_equal(a, b)
        assert np.abs(a - ref).sum() <= (PR_L1 if precision == 32 else 1e-12)
        outs.append(a)
    for a in outs[1:]:
        assert np.abs(a - outs[0]).sum() <= (1e-6 if precision == 32 else 1e-13)
    rg.destroy(g)


def test_pagerank_long_rows_with_a_grid_smaller_than_the_plan():
    # many rows past 2,048 predecessors, dealt to block slots at load; a
    # tuned grid of 3 blocks (fewer than the slots) makes every block take
    # slots b, b + 3, ... and then the row bins: still the oracle's ranks
    from oracle import oracle as O
    from zog_amd import _lib
    rng = np.random.default_rng(9)
    n = 40_000
    heads = rng.choice(n, 40, replace=False)
    src = [rng.integers(0, n, 120_000)]
    dst = [rng.integers(0, n, 120_000)]
    for i, h in enumerate(heads):
        k = int(2_100 + 700 * i)
        src.append(rng.integers(0, n, k))
        dst.append(np.full(k, h))
    src = np.concatenate(src).astype(np.uint32)
    dst = np.concatenate(dst).astype(np.uint32)
    ref, _, _ = O.Graph(n, src, dst).pagerank(0.85, 30, 0.0)
    g = make(n, src, dst)
    h = g.resource.handle
    for grid, block in ((0, 0), (3, 1024), (3, 256), (1, 512)):
        say("pagerank long rows", grid, block)
        _lib.check(_lib.lib().zg_set_tuning(h, _lib.ZG_TUNE_PR_GRID, float(grid)), "zg_set_tuning")
        _lib.check(_lib.lib().zg_set_tuning(h, _lib.ZG_TUNE_PR_BLOCK, float(block)), "zg_set_tuning")
        for prec, bar in ((32, PR_L1), (64, 1e-12)):
            got = rg.pagerank_array(g, 0.85, 30, 0.0, precision=prec)
            assert np.abs(got - ref).sum() <= bar, (grid, block, prec)
    rg.destroy(g)
