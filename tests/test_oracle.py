"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

Each check is independent of the oracle's own code path:
  * the reference's own test cases (test/centrality_test.exs:12-51,138-162,
    test/resource_graph_test.exs:41-127, priv/zog/src/models/array_graph.zig:416-445),
  * closed forms (directed path, star, cycle, complete graph, grid path counts),
  * brute-force enumeration of shortest paths on tiny multigraphs,
  * networkx 3.4.2 on simple digraphs,
  * the PageRank fixed point solved as a dense linear system,
  * big-integer path counts for graphs whose counts overflow u64,
  * the identity sum_v delta_s(v) = sum_t (dist(s, t) - 1).
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from tests import graphs


def brute_bc(n, src, dst):
    """BC(v) = sum over s != v != t of (#shortest s-t arc sequences through v)
    / (#shortest s-t arc sequences), counting parallel arcs separately."""
    adj = [[] for _ in range(n)]
    for u, v in zip(src.tolist(), dst.tolist()):
        adj[u].append(v)
    bc = [Fraction(0)] * n
    for s in range(n):
        # all walks by increasing length until every reachable t has its paths
        dist = {s: 0}
        frontier = [s]
        d = 0
        while frontier:
            nxt = []
            for u in frontier:
                for w in adj[u]:
                    if w not in dist:
                        dist[w] = d + 1
                        nxt.append(w)
            frontier = sorted(set(nxt))
            d += 1
        for t in range(n):
            if t == s or t not in dist:
                continue
            L = dist[t]
            paths = []

            def walk(path):
                u = path[-1]
                if len(path) - 1 == L:
                    if u == t:
                        paths.append(list(path))
                    return
                for w in adj[u]:
                    path.append(w)
                    walk(path)
                    path.pop()

            walk([s])
            total = len(paths)
            for v in range(n):
                if v in (s, t):
                    continue
                through = sum(1 for p in paths if v in p[1:-1])
                if through:
                    bc[v] += Fraction(through, total)
    return np.array([float(x) for x in bc])


def pagerank_solve(n, src, dst, d):
    """Fixed point of centrality.zig:924-958 on an all-live graph:
    (I - d (P + dangling/n)) r = (1 - d)/n."""
    outdeg = np.bincount(src, minlength=n).astype(float)
    M = np.zeros((n, n))
    for u, v in zip(src.tolist(), dst.tolist()):
        M[v, u] += 1.0 / outdeg[u]
    dang = outdeg == 0
    M[:, dang] += 1.0 / n
    return np.linalg.solve(np.eye(n) - d * M, np.full(n, (1 - d) / n))


# --- reference test cases -------------------------------------------------

def test_reference_path_bridge(oracle_mod):
    # centrality_test.exs:12-24  A->B->C
    g = oracle_mod.Graph(3, *graphs.arcs([(0, 1), (1, 2)]))
    bc, ovf = g.betweenness()
    assert not ovf
    assert bc[1] > bc[0] and bc[1] > bc[2]
    assert bc[0] == 0.0 and bc[2] == 0.0
    assert bc[1] == 1.0


def test_reference_star(oracle_mod):
    # centrality_test.exs:26-39 undirected star, 3 leaves -> centre 3.0 after x0.5
    g = oracle_mod.Graph(4, *graphs.undirected([(0, 1), (0, 2), (0, 3)]))
    bc, _ = g.betweenness()
    assert bc[0] * 0.5 == 3.0
    assert list(bc[1:]) == [0.0, 0.0, 0.0]


def test_reference_empty_and_single(oracle_mod):
    # centrality_test.exs:41-51
    g = oracle_mod.Graph(0, *graphs.arcs([]))
    assert len(g.betweenness()[0]) == 0
    assert len(g.pagerank()[0]) == 0
    g = oracle_mod.Graph(1, *graphs.arcs([]))
    assert g.betweenness()[0].tolist() == [0.0]
    assert g.pagerank()[0].tolist() == [1.0]


def test_reference_pagerank_relations(oracle_mod):
    # centrality_test.exs:138-162
    g = oracle_mod.Graph(3, *graphs.arcs([(0, 1), (2, 1)]))
    pr, _, _ = g.pagerank()
    assert pr[1] > pr[0] and pr[1] > pr[2]
    g = oracle_mod.Graph(3, *graphs.arcs([(0, 1), (1, 2)]))
    assert g.pagerank(damping=0.85)[0][0] > 0.0
    assert g.pagerank(damping=0.99)[0][0] > 0.0
    # resource_graph_test.exs:86-101 star: centre highest
    g = oracle_mod.Graph(4, *graphs.undirected([(0, 1), (0, 2), (0, 3)]))
    pr, _, _ = g.pagerank()
    assert all(pr[0] > pr[i] for i in (1, 2, 3))


def test_reference_triangle(oracle_mod):
    # resource_graph_test.exs:57-127: undirected triangle, BC 0, PR 1/3
    g = oracle_mod.Graph(3, *graphs.undirected([(0, 1), (1, 2), (2, 0)]))
    assert g.betweenness()[0].tolist() == [0.0, 0.0, 0.0]
    pr, _, _ = g.pagerank()
    assert np.allclose(pr, 1 / 3, atol=1e-12)


def test_successor_order_multi_edge(oracle_mod):
    # array_graph.zig:416-445: n0->n1, n0->n2, n0->n1 yields n1, n2, n1
    g = oracle_mod.Graph(3, *graphs.arcs([(0, 1), (0, 2), (0, 1)]))
    assert g.successors(0).tolist() == [1, 2, 1]


def test_tombstones(oracle_mod):
    # array_graph.zig:469-507: removeNode(n1) removes its arcs both ways
    src, dst = graphs.arcs([(0, 1), (1, 2), (2, 0)])
    g = oracle_mod.Graph(3, src, dst, live_mask=np.array([1, 0, 1], np.uint8))
    assert g.node_count() == 2
    assert g.successors(2).tolist() == [0]
    assert g.successors(0).tolist() == []
    bc, _ = g.betweenness()
    assert bc.tolist() == [0.0, 0.0, 0.0]
    pr, _, _ = g.pagerank()
    assert pr[1] == 0.0
    assert abs(pr.sum() - 1.0) < 1e-12


# --- closed forms ---------------------------------------------------------

@pytest.mark.parametrize("k", [2, 3, 7, 20])
def test_directed_path_closed_form(oracle_mod, k):
    g = oracle_mod.Graph(k, *graphs.arcs([(i, i + 1) for i in range(k - 1)]))
    bc, _ = g.betweenness()
    assert bc.tolist() == [float(i * (k - 1 - i)) for i in range(k)]


@pytest.mark.parametrize("leaves", [1, 3, 10])
def test_star_closed_form(oracle_mod, leaves):
    g = oracle_mod.Graph(leaves + 1, *graphs.undirected([(0, i) for i in range(1, leaves + 1)]))
    bc, _ = g.betweenness()
    assert bc[0] * 0.5 == leaves * (leaves - 1) / 2


def test_cycle_and_complete(oracle_mod):
    n = 9
    g = oracle_mod.Graph(n, *graphs.arcs([(i, (i + 1) % n) for i in range(n)]))
    bc, _ = g.betweenness()
    assert np.all(bc == bc[0])
    # directed n-cycle: every vertex is interior to paths of length 2..n-1
    assert bc[0] == sum(L - 1 for L in range(1, n))
    pr, _, _ = g.pagerank(tolerance=1e-14, max_iterations=1000)
    assert np.allclose(pr, 1.0 / n, atol=1e-14)
    K = 6
    g = oracle_mod.Graph(K, *graphs.arcs([(u, v) for u in range(K) for v in range(K) if u != v]))
    assert g.betweenness()[0].tolist() == [0.0] * K


def test_grid_path_counts_binomial(oracle_mod):
    # full a x b undirected grid, source at the corner: sigma(r, c) = C(r+c, r)
    a, b = 9, 13
    pairs = []
    for r in range(a):
        for c in range(b):
            if c + 1 < b:
                pairs.append((r * b + c, r * b + c + 1))
            if r + 1 < a:
                pairs.append((r * b + c, (r + 1) * b + c))
    g = oracle_mod.Graph(a * b, *graphs.undirected(pairs))
    depth, su, sf, shift, arcs, ovf = g.bfs_paths(0)
    for r in range(a):
        for c in range(b):
            assert depth[r * b + c] == r + c
            assert su[r * b + c] == math.comb(r + c, r)
            assert sf[r * b + c] == float(math.comb(r + c, r))
    assert not ovf and np.all(shift == 0)


# --- brute force on tiny multigraphs ---------------------------------------

@pytest.mark.parametrize("seed", range(12))
def test_brute_force_multigraph(oracle_mod, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(2, 7))
    m = int(rng.integers(0, 14))
    n, src, dst = graphs.random_multigraph(rng, n, m)
    g = oracle_mod.Graph(n, src, dst)
    bc, ovf = g.betweenness()
    assert not ovf
    assert np.allclose(bc, brute_bc(n, src, dst), rtol=1e-12, atol=1e-12)


def test_multi_edge_doubles_paths(oracle_mod):
    # 0 =>(x2) 1 -> 2 and 0 -> 3 -> 2: sigma(2) = 3, BC(1) = 2/3, BC(3) = 1/3
    g = oracle_mod.Graph(4, *graphs.arcs([(0, 1), (0, 1), (1, 2), (0, 3), (3, 2)]))
    bc, _ = g.betweenness()
    assert np.allclose(bc, [0, 2 / 3, 0, 1 / 3], rtol=1e-15)
    depth, su, *_ = g.bfs_paths(0)
    assert su.tolist() == [1, 2, 3, 1]


def test_self_loops_ignored_by_bc_counted_by_pagerank(oracle_mod):
    src, dst = graphs.arcs([(0, 1), (1, 2), (1, 1)])
    g = oracle_mod.Graph(3, src, dst)
    assert g.betweenness()[0].tolist() == [0.0, 1.0, 0.0]
    pr, _, _ = g.pagerank(tolerance=1e-15, max_iterations=10000)
    assert np.allclose(pr, pagerank_solve(3, src, dst, 0.85), atol=1e-13)


# --- networkx and linear-algebra cross-checks ------------------------------

def test_networkx_simple_digraph(oracle_mod):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(7)
    n = 300
    pairs = set()
    while len(pairs) < 1500:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u != v:
            pairs.add((u, v))
    pairs = sorted(pairs)
    src, dst = graphs.arcs(pairs)
    g = oracle_mod.Graph(n, src, dst)
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    G.add_edges_from(pairs)
    bc, _ = g.betweenness()
    ref = nx.betweenness_centrality(G, normalized=False)
    assert np.allclose(bc, [ref[i] for i in range(n)], rtol=1e-10, atol=1e-9)
    pr, _, _ = g.pagerank(tolerance=1e-15, max_iterations=10000)
    ref = nx.pagerank(G, alpha=0.85, tol=1e-15, max_iter=10000)
    assert np.abs(pr - [ref[i] for i in range(n)]).sum() < 1e-10


@pytest.mark.parametrize("seed", range(4))
def test_pagerank_linear_solve(oracle_mod, seed):
    rng = np.random.default_rng(50 + seed)
    n, src, dst = graphs.random_multigraph(rng, 60, 150)
    g = oracle_mod.Graph(n, src, dst)
    pr, it, l1 = g.pagerank(tolerance=1e-15, max_iterations=100000)
    assert l1 < 1e-15
    assert np.abs(pr - pagerank_solve(n, src, dst, 0.85)).sum() < 1e-12


def test_pagerank_stop_rule(oracle_mod):
    # strict l1 < tol after the swap; max_iterations caps; returns the last iterate
    n, src, dst = graphs.random_multigraph(np.random.default_rng(3), 40, 100)
    g = oracle_mod.Graph(n, src, dst)
    pr5, it5, _ = g.pagerank(max_iterations=5, tolerance=0.0)
    assert it5 == 5
    _, it_tol, l1_tol = g.pagerank(tolerance=1e-6)
    _, it_prev, l1_prev = g.pagerank(tolerance=1e-6, max_iterations=it_tol - 1)
    assert l1_tol < 1e-6 <= l1_prev
    pr0, it0, _ = g.pagerank(max_iterations=0)
    assert it0 == 0 and np.allclose(pr0, 1.0 / n)


# --- overflowing path counts ----------------------------------------------

def _grid(a, b):
    pairs = []
    for r in range(a):
        for c in range(b):
            if c + 1 < b:
                pairs.append((r * b + c, r * b + c + 1))
            if r + 1 < a:
                pairs.append((r * b + c, (r + 1) * b + c))
    return a * b, *graphs.undirected(pairs)


def _bigint_sigma(n, src, dst, s):
    adj = [[] for _ in range(n)]
    for u, v in zip(src.tolist(), dst.tolist()):
        adj[u].append(v)
    dist = [-1] * n
    sig = [0] * n
    dist[s] = 0
    sig[s] = 1
    order = [s]
    i = 0
    while i < len(order):
        u = order[i]
        i += 1
        for w in adj[u]:
            if dist[w] < 0:
                dist[w] = dist[u] + 1
                order.append(w)
            if dist[w] == dist[u] + 1:
                sig[w] += sig[u]
    return dist, sig, order, adj


def test_scaled_sigma_exact_ratios_beyond_u64(oracle_mod):
    # 40x40 grid from a corner: counts reach C(78,39) ~ 2.6e22 > 2^64.  The
    # u64 mode wraps (and says so); the FP64 mode's delta matches a delta
    # computed from exact big-integer ratios.
    n, src, dst = _grid(40, 40)
    g = oracle_mod.Graph(n, src, dst)
    _, ovf = g.betweenness(sources=[0])
    assert ovf
    bc, _ = g.betweenness(sources=[0], mode=oracle_mod.SIGMA_SCALED)
    dist, sig, order, adj = _bigint_sigma(n, src, dst, 0)
    delta = [0.0] * n
    for w in reversed(order):
        for v in adj[w]:
            if dist[v] == dist[w] + 1:
                delta[w] += float(Fraction(sig[w], sig[v])) * (1.0 + delta[v])
    ref = np.array(delta)
    ref[0] = 0.0
    assert np.allclose(bc, ref, rtol=1e-12, atol=0)


def test_scaled_sigma_shift_beyond_fp64_range(oracle_mod):
    # 300x300 grid from a corner: counts reach C(598,299) ~ 1e178 > 2^512, so
    # shifts are taken; stored * 2^(sum of shifts) equals the big integer.
    n, src, dst = _grid(300, 300)
    g = oracle_mod.Graph(n, src, dst)
    depth, _, sf, shift, _, _ = g.bfs_paths(0)
    assert shift.max() > 0
    cum = np.concatenate([[0], np.cumsum(shift.astype(np.int64))])
    dist, sig, _, _ = _bigint_sigma(n, src, dst, 0)
    for v in [0, 1, 299, 300 * 150 + 150, 300 * 299 + 299, 300 * 200 + 17]:
        L = depth[v]
        log2_true = math.log2(sig[v])
        log2_stored = math.log2(sf[v]) + cum[L]
        assert abs(log2_true - log2_stored) < 1e-9 * max(1.0, log2_true)
    # dependency checksum: sum_v delta_s(v) = sum_t (dist(s,t) - 1) over t != s
    bc, _ = g.betweenness(sources=[0], mode=oracle_mod.SIGMA_SCALED)
    checksum = sum(d - 1 for d in dist if d > 0)
    assert abs(bc.sum() - checksum) <= 1e-9 * checksum
    # the corner source is symmetric under transpose
    B = bc.reshape(300, 300)
    assert np.allclose(B, B.T, rtol=1e-9, atol=1e-6)


def test_exp_sigma_matches_bigint_and_scaled(oracle_mod):
    # per-vertex exponents: the 40x40 exact-ratio case and the 300x300 case
    # where both FP64 modes are valid
    O = oracle_mod
    n, src, dst = _grid(40, 40)
    g = O.Graph(n, src, dst)
    a, _ = g.betweenness(sources=[0], mode=O.SIGMA_EXP)
    b, _ = g.betweenness(sources=[0], mode=O.SIGMA_SCALED)
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    depth, mant, ex = g.bfs_paths_exp(0)
    dist, sig, _, _ = _bigint_sigma(n, src, dst, 0)
    for v in range(0, n, 37):
        got = mant[v] * 2.0 ** ex[v]
        if sig[v] < 2 ** 53:
            assert got == float(sig[v])  # exact
        else:
            assert abs(got - float(sig[v])) <= 1e-13 * float(sig[v])  # rounded sums
        assert 1.0 <= mant[v] < 2.0
    n, src, dst = _grid(300, 300)
    g = O.Graph(n, src, dst)
    a, _ = g.betweenness(sources=[0, 299, 45150], mode=O.SIGMA_EXP)
    b, _ = g.betweenness(sources=[0, 299, 45150], mode=O.SIGMA_SCALED)
    assert np.allclose(a, b, rtol=1e-12, atol=0)


def test_exp_sigma_beyond_one_level_fp64_range(oracle_mod):
    # 1600x1600 grid from a corner: a level's counts run from 1 (on the
    # grid's edges) to ~2^(level - log2 level), and the shifts taken for the
    # largest add up past 2^1074, so the per-level mode loses the small
    # counts (and goes non-finite) while per-vertex exponents keep them all
    O = oracle_mod
    R = 1600
    n, src, dst = _grid(R, R)
    g = O.Graph(n, src, dst)
    sc, _ = g.betweenness(sources=[0], mode=O.SIGMA_SCALED)
    assert not np.all(np.isfinite(sc))
    bc, _ = g.betweenness(sources=[0], mode=O.SIGMA_EXP)
    assert np.all(np.isfinite(bc))
    depth, mant, ex = g.bfs_paths_exp(0)
    for r, c in [(0, 0), (0, R - 1), (1, 1), (R // 2, R // 2), (R - 1, R - 1), (700, 13), (3, 1590), (1599, 2)]:
        v = r * R + c
        log2_true = math.log2(math.comb(r + c, r))  # lattice paths from the corner
        assert abs(math.log2(mant[v]) + int(ex[v]) - log2_true) < 1e-12 * max(1.0, log2_true)
    checksum = int((depth.astype(np.int64) - 1).clip(0).sum())
    assert abs(bc.sum() - checksum) <= 1e-9 * checksum
    B = bc.reshape(R, R)
    assert np.allclose(B, B.T, rtol=1e-9, atol=1e-6)


# --- the parity configuration ------------------------------------------------

def test_er10k_checksum_and_threads(oracle_mod):
    from zog_amd import workloads
    n, src, dst = workloads.er10k()
    g = oracle_mod.Graph(n, src, dst)
    srcs = np.arange(0, n, 37, dtype=np.uint32)
    bc1, ovf, arcs = g.betweenness(sources=srcs, return_arcs=True)
    assert not ovf
    bc4, _ = g.betweenness(sources=srcs, threads=4)
    assert np.allclose(bc1, bc4, rtol=1e-12, atol=1e-9)
    checksum = 0
    for s in srcs[:20]:
        depth, *_ = g.bfs_paths(int(s))
        checksum += int(np.where(depth != oracle_mod.NONE_U32, depth.astype(np.int64) - 1, 0).clip(0).sum())
    bc20, _ = g.betweenness(sources=srcs[:20])
    assert abs(bc20.sum() - checksum) <= 1e-9 * checksum


# --- weighted Brandes (centrality.zig:529-667, betweennessF64 :1355-1357) ---

def brute_bc_weighted(n, src, dst, w):
    """BC(v) from every simple s-t arc sequence whose weight sum (in path
    order, as Dijkstra adds) is the smallest, parallel arcs separately."""
    adj = [[] for _ in range(n)]
    for u, v, x in zip(src.tolist(), dst.tolist(), w.tolist()):
        adj[u].append((v, x))
    bc = [Fraction(0)] * n
    for s in range(n):
        best = {}
        paths = {}

        def walk(u, total, seen, inner):
            for v, x in adj[u]:
                if v in seen:
                    continue
                t = total + x
                if v not in best or t < best[v]:
                    best[v] = t
                    paths[v] = [inner]
                elif t == best[v]:
                    paths[v].append(inner)
                walk(v, t, seen | {v}, inner + [v])

        walk(s, 0.0, {s}, [])
        for t, plist in paths.items():
            if t == s:
                continue
            # keep only the shortest ones (walk recorded every improvement)
            good = [p for p in plist]
            k = len(good)
            for p in good:
                for v in p:
                    bc[v] += Fraction(1, k)
    return [float(x) for x in bc]


def test_weighted_reference_cases(oracle_mod):
    O = oracle_mod
    # centrality_test.exs:55-65 and centrality.zig:1436-1454: bridge on a chain
    g = O.Graph(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    g.set_weights([1.0, 1.0])
    bc = g.betweenness_f64()
    assert bc[1] > bc[0] and bc[1] > bc[2]
    # centrality_test.exs:67-79: the 10.0 shortcut is longer than A->B->C
    g = O.Graph(3, np.array([0, 1, 0], np.uint32), np.array([1, 2, 2], np.uint32))
    g.set_weights([1.0, 1.0, 10.0])
    assert g.betweenness_f64().tolist() == [0.0, 1.0, 0.0]
    g.set_weights([1.0, 1.0, 1.5])
    assert g.betweenness_f64().tolist() == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("seed", range(6))
def test_weighted_brute_force(oracle_mod, seed):
    # integer-valued weights so that ties (equal sums) are exact
    O = oracle_mod
    rng = np.random.default_rng(500 + seed)
    n = 7
    m = 16
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.integers(1, 4, len(src)).astype(np.float64)
    g = O.Graph(n, src, dst)
    g.set_weights(w)
    got = g.betweenness_f64()
    ref = brute_bc_weighted(n, src, dst, w)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (got, ref)


def test_weighted_unit_weights_equal_unweighted(oracle_mod):
    # all weights 1.0: the same shortest paths (parallel arcs included)
    O = oracle_mod
    rng = np.random.default_rng(9)
    n, src, dst = graphs.random_multigraph(rng, 200, 1500)
    g = O.Graph(n, src, dst)
    g.set_weights(np.ones(len(src)))
    ref, _ = g.betweenness()
    got = g.betweenness_f64()
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_weighted_networkx(oracle_mod):
    nx = pytest.importorskip("networkx")
    O = oracle_mod
    rng = np.random.default_rng(13)
    n = 120
    src = rng.integers(0, n, 900).astype(np.uint32)
    dst = rng.integers(0, n, 900).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    _, first = np.unique(src.astype(np.int64) * n + dst, return_index=True)
    first.sort()
    src, dst = src[first], dst[first]
    w = rng.uniform(0.1, 3.0, len(src))
    g = O.Graph(n, src, dst)
    g.set_weights(w)
    got = g.betweenness_f64()
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    for a, b, x in zip(src.tolist(), dst.tolist(), w.tolist()):
        G.add_edge(a, b, weight=x)
    ref = nx.betweenness_centrality(G, normalized=False, weight="weight")
    assert np.allclose(got, [ref[v] for v in range(n)], rtol=1e-12, atol=1e-9)


# --- eigenvector / Katz / alpha (centrality.zig:1050-1333) -------------------

def _dense_in(n, src, dst, live=None):
    """A^T as a dense matrix over the live nodes: row v holds the number of
    arcs u -> v (parallel arcs counted, self-loops kept)."""
    M = np.zeros((n, n))
    for u, v in zip(src.tolist(), dst.tolist()):
        if live is None or (live[u] and live[v]):
            M[v, u] += 1.0
    return M


def test_power_reference_cases(oracle_mod):
    O = oracle_mod
    # centrality.zig:1478-1498 / centrality_test.exs:166-178: a 3-cycle
    g = O.Graph(3, np.array([0, 1, 2], np.uint32), np.array([1, 2, 0], np.uint32))
    ev, it, oc = g.eigenvector(100, 1e-4)
    assert abs(ev[0] - ev[1]) < 0.01 and abs(ev[1] - ev[2]) < 0.01
    assert oc == O.Graph.EIG_CAP and it == 100  # a 3-cycle rotates: no convergence, no period 2
    # centrality_test.exs:180-192: undirected star, the hub scores highest
    s, d = graphs.undirected([(0, 1), (0, 2), (0, 3)])
    ev, _, _ = O.Graph(4, s, d).eigenvector()
    assert ev[0] > ev[1] and ev[0] > ev[2] and ev[0] > ev[3]
    # centrality_test.exs:196-206: B (1) receives from A (0) and C (2)
    g = O.Graph(3, np.array([0, 2], np.uint32), np.array([1, 1], np.uint32))
    k, _, _ = g.katz(0.1, 1.0)
    assert k[1] > k[0] and k[1] > k[2]
    # centrality_test.exs:208-217 and centrality.zig:1501-1518: at least beta
    g = O.Graph(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    k, _, _ = g.katz(0.1, 1.0, 100, 1e-4)
    assert np.all(k >= 1.0)
    # centrality.zig:1520-1538: path 0 -> 1 -> 2, alpha 0.5, start 1.0
    a, it, _ = g.alpha_centrality(0.5, 1.0, 100, 1e-4)
    assert a[2] >= a[0]
    assert a.tolist() == [0.0, 0.0, 0.0] and it == 4  # alpha^k A^k x0 dies after 3 steps


@pytest.mark.parametrize("seed", range(3))
def test_katz_alpha_match_dense_iterates(oracle_mod, seed):
    # the k-th iterate of x <- alpha A^T x (+ beta) from a dense matrix;
    # tolerance -1 never stops early
    O = oracle_mod
    rng = np.random.default_rng(60 + seed)
    n, src, dst = graphs.random_multigraph(rng, 40, 160)
    live = (rng.random(n) > 0.1).astype(np.uint8)
    g = O.Graph(n, src, dst, live)
    M = _dense_in(n, src, dst, live)
    lv = live.astype(bool)
    for k in (0, 1, 2, 7):
        x = np.where(lv, 1.0, 0.0)
        for _ in range(k):
            x = np.where(lv, 0.05 * (M @ x) + 1.0, 0.0)
        got, it, _ = g.katz(0.05, 1.0, k, -1.0)
        assert it == k and np.allclose(got, x, rtol=1e-13, atol=0)
        y = np.where(lv, 0.7, 0.0)
        for _ in range(k):
            y = np.where(lv, 0.3 * (M @ y), 0.0)
        got, it, _ = g.alpha_centrality(0.3, 0.7, k, -1.0)
        assert it == k and np.allclose(got, y, rtol=1e-13, atol=0)
    # converged Katz is the resolvent beta (I - alpha A^T)^-1 1
    rho = max(abs(np.linalg.eigvals(M)))
    alpha = 0.5 / max(rho, 1.0)
    got, it, l1 = g.katz(alpha, 1.0, 10_000, 1e-13)
    idx = np.nonzero(lv)[0]
    ref = np.linalg.solve(np.eye(len(idx)) - alpha * M[np.ix_(idx, idx)], np.ones(len(idx)))
    assert l1 < 1e-13 and np.allclose(got[idx], ref, rtol=1e-11)


def test_eigenvector_closed_forms(oracle_mod):
    O = oracle_mod
    # strongly connected and aperiodic (a 5-cycle with a chord and a
    # self-loop): the Perron vector of A^T, unit L2 norm
    pairs = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2), (3, 3)]
    s, d = graphs.arcs(pairs)
    ev, it, oc = O.Graph(5, s, d).eigenvector(1000, 1e-12)
    M = _dense_in(5, s, d)
    w, V = np.linalg.eig(M)
    p = np.real(V[:, np.argmax(np.real(w))])
    p = np.abs(p) / np.linalg.norm(p)
    assert oc == O.Graph.EIG_CONVERGED and np.allclose(ev, p, atol=1e-10)
    # undirected star with k leaves: period-2 iterates, detected as an
    # oscillation; the normalised average is the Perron vector
    k = 5
    s, d = graphs.undirected([(0, i) for i in range(1, k + 1)])
    ev, it, oc = O.Graph(k + 1, s, d).eigenvector(1000, 1e-9)
    assert oc == O.Graph.EIG_OSCILLATING
    assert np.allclose(ev, [1 / np.sqrt(2)] + [1 / np.sqrt(2 * k)] * k, atol=1e-9)
    # a DAG reaches the zero vector: uniform 1/sqrt(n)
    s, d = graphs.arcs([(0, 1), (1, 2), (0, 2), (2, 3)])
    ev, it, oc = O.Graph(4, s, d).eigenvector()
    assert oc == O.Graph.EIG_ZERO and it == 4 and np.allclose(ev, 0.5)
    # one node: 1.0; no iterations: the start vector 1 + i/1000 over live nodes
    assert O.Graph(1, s[:0], d[:0]).eigenvector()[0].tolist() == [1.0]
    live = np.array([1, 0, 1, 1], np.uint8)
    ev, it, _ = O.Graph(4, s, d, live).eigenvector(0, 1e-4)
    assert it == 0 and ev.tolist() == [1.0, 0.0, 1.001, 1.002]
