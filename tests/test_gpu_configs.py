"""Parity of the gfx950 path on the benchmark workloads at their real sizes.

Each workload is built with its own seed and launch shape, exactly as the
benchmark builds it, and the HIP path (through the C ABI) is checked against
the CPU oracle on the part of the workload the oracle finishes in seconds,
plus size-independent properties over the whole workload:

* R-MAT scale 20 (2^24 distinct arcs), 4,096 seeded sources swept 256 per
  launch: 64 of the sources against the oracle (BC within 1e-9, arcs scanned
  equal), depths and path counts bit-exact for 4 sources, the full 4,096-source
  sweep equal to the sum of its 256-source slices; PageRank FP32 at
  tolerance 1e-6 against the FP64 oracle (L1 <= 1e-6, equal iteration counts).
* 512 Barabasi-Albert graphs (m=4, 1,000..100,000 nodes) in one batched
  resource: every source of the smallest graphs and 64 sources of the largest,
  each graph's BC against the oracle on that graph alone; the other 500-odd
  graphs' scores stay exactly 0.
* 2048x2048 grid with 5% of edges deleted, 1,024 seeded sources: 2 sources
  against the oracle's per-vertex-exponent counts (the counts pass 2^4000),
  the 1,024-source sweep equal to the sum of its slices; mixed-precision
  PageRank to 1e-8 against the FP64 oracle.
* R-MAT scale 22 (2^26 distinct arcs), the learned-pivot workload: core
  numbers and node features against the oracle, the network's forward pass on
  a sample of rows, and the weighted (Horvitz-Thompson) sweep over 4 learned
  pivots against the oracle.

What is checked follows priv/zog/src/centrality.zig:694-798 (Brandes) and
:871-965 (PageRank).
"""
import os
import sys
import time

import numpy as np
import pytest

from zog_amd import generators as gen
from zog_amd import resource_graph as rg
from zog_amd import workloads as W
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9
THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def assert_bc_close(got, ref, rtol=BC_RTOL):
    scale = max(1.0, float(np.abs(ref).max()) if len(ref) else 1.0)
    err = np.abs(got - ref)
    ok = err <= rtol * np.maximum(np.abs(ref), 1e-9 * scale)
    assert ok.all(), f"max err {err.max()} at {int(np.argmax(err))} (ref {ref[int(np.argmax(err))]})"


def set_shape(g, groups, alpha, pull_grid, order):
    """The launch shape the benchmark's autotuner picks for the workload."""
    rg.set_batch_groups(g, groups)
    rg.set_push_alpha(g, alpha)
    rg.set_pull_grid(g, pull_grid)
    rg.set_source_order(g, order)


# --- R-MAT scale 20, 4,096 sources, PageRank tol 1e-6 -------------------------

@pytest.fixture(scope="module")
def rmat20():
    from oracle import oracle as O
    t0 = time.time()
    n, src, dst = W.rmat20()
    sources = W.rmat20_sources(n, src, dst, 4096)
    g = rg.new(SoA.from_arrays("directed", n, src, dst))
    set_shape(g, 4, 16.0, 16384, 2)
    og = O.Graph(n, src, dst)
    say(f"rmat20 built in {time.time() - t0:.1f}s: n={n} arcs={len(src)} sources={len(sources)}")
    yield n, src, dst, sources, g, og
    rg.destroy(g)
    og.close()


def test_rmat20_bc_sources_against_oracle(rmat20):
    n, src, dst, sources, g, og = rmat20
    sample = np.ascontiguousarray(sources[::64])  # 64 of the 4,096, spread over the seeded order
    say("rmat20 bc", len(sample), "sources, oracle on", THREADS, "threads")
    ref, ovf, arcs = og.betweenness(sources=sample, threads=THREADS, return_arcs=True)
    assert not ovf
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=sample, info=info)
    assert_bc_close(got, ref)
    assert info["arcs_traversed"] == int(arcs.sum())
    assert info["n_sources"] == len(sample)


def test_rmat20_depth_sigma_bit_exact(rmat20):
    from oracle import oracle as O
    n, src, dst, sources, g, og = rmat20
    pick = np.ascontiguousarray(sources[[0, 1000, 2000, 4095]])
    say("rmat20 depth/sigma", pick.tolist())
    depth, sigma, _, levels = rg.bc_paths(g, pick, max_levels=64)
    for i, s in enumerate(pick.tolist()):
        d_ref, su, _, sh, _, ovf = og.bfs_paths(s)
        assert not ovf
        assert np.array_equal(depth[i], d_ref)
        r = d_ref != O.NONE_U32
        assert np.array_equal(sigma[i][r], su[r].astype(np.float64))
        assert np.all(sigma[i][r] == su[r])  # counts stay below 2^53 here
        assert levels[i] == len(sh)


def test_rmat20_full_sweep_is_sum_of_launch_slices(rmat20):
    # the benchmark's step (4,096 sources, 256 per launch) against the sum of
    # its 256-source slices swept one at a time
    n, src, dst, sources, g, og = rmat20
    say("rmat20 full 4096-source sweep")
    info = {}
    full = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert info["n_sources"] == 4096 and info["batches"] == 16
    acc = np.zeros(n)
    arcs = 0
    for a in range(0, len(sources), 256):
        inf = {}
        acc += rg.betweenness_unweighted_array(g, sources=sources[a:a + 256], info=inf)
        arcs += inf["arcs_traversed"]
    assert arcs == info["arcs_traversed"]
    assert_bc_close(full, acc)
    assert np.all(full >= 0.0)


def test_rmat20_pagerank_fp32_tol_1e6(rmat20):
    n, src, dst, sources, g, og = rmat20
    say("rmat20 pagerank")
    ref, it_ref, l1_ref = og.pagerank(0.85, 100, 1e-6, threads=THREADS)
    info = {}
    got = rg.pagerank_array(g, 0.85, 100, 1e-6, precision=32, info=info)
    l1 = float(np.abs(got - ref).sum())
    say(f"  iterations {info['iterations']} (oracle {it_ref}), L1 vs FP64 {l1:.3e}, "
        f"error bound {info['error_bound_l1']:.3e}")
    assert info["iterations"] == it_ref
    assert l1 <= 1e-6
    assert info["last_l1"] < 1e-6 and l1_ref < 1e-6


# --- 512 Barabasi-Albert graphs in one batched resource ------------------------

@pytest.fixture(scope="module")
def ba512():
    t0 = time.time()
    graphs = gen.ba_batch(512, 4, 1000, 100000, W.SEEDS["ba512"])
    res = rg.BatchResource(graphs)
    g = rg.ResourceGraph(resource=res, builder=None)
    set_shape(g, 64, 16.0, 4096, 2)
    say(f"ba512 built in {time.time() - t0:.1f}s: {len(graphs)} graphs, {int(res.offsets[-1])} nodes")
    yield graphs, res, g
    res.release()


def test_ba512_batch_against_each_graph_alone(ba512):
    from oracle import oracle as O
    graphs, res, g = ba512
    off = res.offsets
    sizes = np.array([gr[0] for gr in graphs])
    by_size = np.argsort(sizes, kind="stable")
    every = [int(i) for i in by_size[:3]]          # all sources of the 3 smallest graphs
    largest = int(by_size[-1])                      # 64 sources of the largest
    big_local = np.linspace(0, sizes[largest] - 1, 64).astype(np.uint32)
    parts = [np.arange(off[i], off[i + 1], dtype=np.uint32) for i in every]
    parts.append(big_local + np.uint32(off[largest]))
    sources = np.concatenate(parts)
    say("ba512 batch", len(sources), "sources in graphs", every + [largest], "sizes", sizes[every + [largest]].tolist())
    info = {}
    out = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    arcs_ref = 0
    for i, local in zip(every + [largest], [None, None, None, big_local]):
        n, s, d = graphs[i]
        ref, _, arcs = O.Graph(n, s, d).betweenness(sources=local, threads=THREADS, return_arcs=True)
        arcs_ref += int(arcs.sum())
        assert_bc_close(out[off[i]:off[i + 1]], ref)
    assert info["arcs_traversed"] == arcs_ref
    touched = np.zeros(len(out), bool)
    for i in every + [largest]:
        touched[off[i]:off[i + 1]] = True
    assert np.all(out[~touched] == 0.0)


def test_ba512_all_sources_of_mid_graphs(ba512):
    # every source of a few mid-size graphs of the batch (a group spans one
    # graph's whole span), each graph against the oracle alone
    from oracle import oracle as O
    graphs, res, g = ba512
    off = res.offsets
    sizes = np.array([gr[0] for gr in graphs])
    mids = [int(i) for i in np.nonzero((sizes > 4000) & (sizes < 9000))[0][:4]]
    assert mids
    sources = np.concatenate([np.arange(off[i], off[i + 1], dtype=np.uint32) for i in mids])
    say("ba512 mid graphs", mids, sizes[mids].tolist())
    out = rg.betweenness_unweighted_array(g, sources=sources)
    for i in mids:
        n, s, d = graphs[i]
        ref, _ = O.Graph(n, s, d).betweenness(threads=THREADS)
        assert_bc_close(out[off[i]:off[i + 1]], ref)


# --- 2048x2048 grid, 5% of edges deleted ---------------------------------------

@pytest.fixture(scope="module")
def grid2048():
    from oracle import oracle as O
    t0 = time.time()
    n, src, dst = W.grid2048()
    sources = W.grid2048_sources(n, src, dst, 1024)
    g = rg.new(SoA.from_arrays("directed", n, src, dst))
    og = O.Graph(n, src, dst)
    say(f"grid2048 built in {time.time() - t0:.1f}s: n={n} arcs={len(src)}")
    yield n, src, dst, sources, g, og
    rg.destroy(g)
    og.close()


def test_grid2048_bc_against_exponent_oracle(grid2048):
    from oracle import oracle as O
    n, src, dst, sources, g, og = grid2048
    pick = np.ascontiguousarray(sources[:2])
    say("grid2048 bc", pick.tolist())
    ref, _, arcs = og.betweenness(sources=pick, mode=O.SIGMA_EXP, threads=2, return_arcs=True)
    assert np.all(np.isfinite(ref))
    info = {}
    got = rg.betweenness_unweighted_array(g, sources=pick, info=info)
    assert np.all(np.isfinite(got))
    assert_bc_close(got, ref)
    assert info["arcs_traversed"] == int(arcs.sum())
    deepest = 0
    for s in pick.tolist():
        depth = og.bfs_paths_exp(s)[0]
        deepest = max(deepest, int(depth[depth != O.NONE_U32].max()))
    assert info["max_levels"] == deepest + 1


def test_grid2048_full_sweep_is_sum_of_slices(grid2048):
    n, src, dst, sources, g, og = grid2048
    say("grid2048 full 1024-source sweep")
    info = {}
    full = rg.betweenness_unweighted_array(g, sources=sources, info=info)
    assert info["n_sources"] == 1024 and np.all(np.isfinite(full))
    acc = np.zeros(n)
    arcs = 0
    for a in range(0, len(sources), 256):
        inf = {}
        acc += rg.betweenness_unweighted_array(g, sources=sources[a:a + 256], info=inf)
        arcs += inf["arcs_traversed"]
    assert arcs == info["arcs_traversed"]
    assert_bc_close(full, acc)


def test_grid2048_pagerank_mixed_to_1e8(grid2048):
    n, src, dst, sources, g, og = grid2048
    say("grid2048 pagerank")
    ref, it_ref, l1_ref = og.pagerank(0.85, 1000, 1e-8, threads=THREADS)
    info = {}
    got = rg.pagerank_array(g, 0.85, 1000, 1e-8, precision=96, info=info)
    l1 = float(np.abs(got - ref).sum())
    say(f"  iterations {info['iterations']} ({info['fp64_iterations']} FP64; oracle {it_ref}), "
        f"L1 vs FP64 {l1:.3e}, error bound {info['error_bound_l1']:.3e}")
    assert info["last_l1"] < 1e-8 and info["fp64_iterations"] > 0
    assert abs(info["iterations"] - it_ref) <= 1
    assert l1 <= 1e-8


# --- R-MAT scale 22: the learned-pivot workload --------------------------------

@pytest.fixture(scope="module")
def rmat22():
    from oracle import oracle as O
    t0 = time.time()
    n, src, dst = gen.rmat(22, 16 << 22, W.SEEDS["rmat22"])
    g = rg.new(SoA.from_arrays("directed", n, src, dst))
    og = O.Graph(n, src, dst)
    say(f"rmat22 built in {time.time() - t0:.1f}s: n={n} arcs={len(src)}")
    yield {"n": n, "src": src, "dst": dst, "g": g, "og": og}
    rg.destroy(g)
    og.close()


@pytest.fixture(scope="module")
def rmat22_core(rmat22):
    t0 = time.time()
    core = rmat22["og"].core_numbers()
    say(f"rmat22 oracle core numbers in {time.time() - t0:.1f}s")
    return core


def test_rmat22_core_numbers(rmat22, rmat22_core):
    from zog_amd import pivots as PV
    got = PV.core_numbers(rmat22["g"])
    assert np.array_equal(got, rmat22_core)


def test_rmat22_features_and_forward(rmat22, rmat22_core):
    import torch
    from oracle import mlp as M
    from zog_amd import pivots as PV
    n, src, dst, og = rmat22["n"], rmat22["src"], rmat22["dst"], rmat22["og"]
    t0 = time.time()
    pr, _, _ = og.pagerank(0.85, 100, 1e-6, threads=THREADS)
    ref = M.features(n, src, dst, None, pr, rmat22_core)
    say(f"rmat22 oracle features in {time.time() - t0:.1f}s")
    feats = PV.node_features(rmat22["g"])
    got = feats.cpu().numpy()
    err = np.abs(got - ref).max(0)
    say("  feature column errors", np.round(err, 6).tolist())
    assert err.max() < 2e-3, err
    # the network's forward pass on a sample of the graph's rows
    rng = np.random.default_rng(22)
    rows = np.sort(rng.choice(n, 100_000, replace=False))
    sc = PV.Scorer(seed=9)
    p = rng.normal(0, 0.2, M.P).astype(np.float32)
    sc.params = p
    x = feats[torch.from_numpy(rows).cuda()].contiguous()
    y = sc.forward(x).cpu().numpy()
    yref, _ = M.forward(p, x.cpu().numpy())
    sc.close()
    assert np.abs(y - yref).max() <= 1e-3 * max(1.0, float(np.abs(yref).max()))


def test_rmat22_weighted_sweep_over_learned_pivots(rmat22):
    # a scorer trained briefly on small R-MAT graphs proposes 4 pivots with
    # Horvitz-Thompson weights 1/pi; the weighted sweep against the oracle
    from oracle import oracle as O
    from zog_amd import pivots as PV
    n, og, g = rmat22["n"], rmat22["og"], rmat22["g"]
    X, T = PV.rmat_training_set(scales=(10, 11), per_scale=3, seed=91, workers=2)
    sc = PV.Scorer(seed=5)
    p = sc.params
    p[-1] = float(T.mean())
    sc.params = p
    sc.fit(X, T, epochs=10, batch=8192, lr=2e-3)
    pred, live_ids = PV.predict_totals(g, sc)
    sc.close()
    rng = np.random.default_rng(4)
    pick, w = PV.learned_pivots(pred, live_ids, 4, rng)
    assert 3 <= len(pick) <= 5 and np.all(np.isfinite(w)) and np.all(w >= 1.0)
    say("rmat22 pivots", pick.tolist(), "weights", np.round(w, 3).tolist())
    info = {}
    bc, tot = PV.bc_weighted(g, pick, w, want_bc=True, want_totals=True, info=info)
    t0 = time.time()

    def one(s):
        # the oracle releases the GIL inside its C calls: one source per thread
        b, _ = og.betweenness(sources=np.array([s], np.uint32))
        depth, *_ = og.bfs_paths(s)
        r = depth != O.NONE_U32
        return b, float(np.sum(depth[r].astype(np.int64) - 1) + 1)  # the source itself has depth 0

    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(max_workers=len(pick)) as ex:
        per = list(ex.map(one, pick.tolist()))
    say(f"  oracle in {time.time() - t0:.1f}s")
    ref = np.zeros(n)
    for i, (b, d_ref) in enumerate(per):
        ref += w[i] * b
        assert tot[i] == d_ref
    assert_bc_close(bc, ref)
