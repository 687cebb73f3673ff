"""Eigenvector, Katz and alpha centrality (priv/zog/src/centrality.zig:
1050-1201, :1212-1267, :1278-1333) on the GPU against the oracle: the
reference's own cases, every way the eigenvector iteration can end, random
multigraphs with tombstones and an R-MAT graph, iteration counts included."""
import sys

import numpy as np
import pytest

from tests import graphs
from zog_amd import _lib
from zog_amd import generators as gen
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu

# FP64 with compensated row sums against the reference's plain sums
RTOL = 1e-10


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make(n, src, dst, live=None):
    return rg.new(SoA.from_arrays("directed", n, src, dst), live_mask=live)


def og(n, src, dst, live=None):
    from oracle import oracle as O
    return O.Graph(n, src, dst, live)


def close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.allclose(a, b, rtol=RTOL, atol=1e-14 * max(1.0, float(np.abs(b).max())))


def test_reference_cases():
    say("power reference cases")
    # centrality_test.exs:166-178: a directed 3-cycle
    g = rg.new(SoA.from_list("directed", [("A", "B", 1.0), ("B", "C", 1.0), ("C", "A", 1.0)]))
    s = rg.eigenvector(g)
    assert abs(s["A"] - s["B"]) < 0.01 and abs(s["B"] - s["C"]) < 0.01
    rg.destroy(g)
    # centrality_test.exs:180-192: the hub of an undirected star
    b = SoA.from_list("undirected", [("hub", "A", 1.0), ("hub", "B", 1.0), ("hub", "C", 1.0)])
    g = rg.new(b)
    s = rg.eigenvector(g)
    assert all(s["hub"] > s[x] for x in "ABC")
    rg.destroy(g)
    # centrality_test.exs:196-217
    g = rg.new(SoA.from_list("directed", [("A", "B", 1.0), ("C", "B", 1.0)]))
    s = rg.katz(g)
    assert s["B"] > s["A"] and s["B"] > s["C"]
    s = rg.alpha_centrality(g)  # centrality_test.exs:221-232
    assert len(s) == 3 and all(isinstance(v, float) for v in s.values())
    rg.destroy(g)
    g = rg.new(SoA.from_list("directed", [("A", "B", 1.0)]))
    s = rg.katz(g, beta=1.0)
    assert s["A"] >= 1.0 and s["B"] >= 1.0
    rg.destroy(g)
    # centrality.zig:1520-1538: path 0 -> 1 -> 2
    g = make(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    info = {}
    a = rg.alpha_centrality(g, 0.5, 1.0, raw=True, info=info)
    assert a[2] >= a[0] and a == [0.0, 0.0, 0.0] and info["iterations"] == 4
    rg.destroy(g)


def _eig_case(n, pairs, max_it, tol, live=None):
    s, d = graphs.arcs(pairs) if pairs else (np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    ref, it, oc = og(n, s, d, live).eigenvector(max_it, tol)
    g = make(n, s, d, live)
    info = {}
    got = np.array(rg.eigenvector(g, max_it, tol, raw=True, info=info))
    rg.destroy(g)
    return got, ref, it, oc, info


@pytest.mark.parametrize("case", ["converged", "oscillating", "zero", "cap", "trivial", "start"])
def test_eigenvector_outcomes(case):
    say("eigenvector", case)
    from oracle import oracle as O
    if case == "converged":
        got, ref, it, oc, info = _eig_case(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2), (3, 3)], 1000, 1e-12)
        want = O.Graph.EIG_CONVERGED
    elif case == "oscillating":
        pairs = [(0, i) for i in range(1, 6)] + [(i, 0) for i in range(1, 6)]
        got, ref, it, oc, info = _eig_case(6, pairs, 1000, 1e-9)
        want = O.Graph.EIG_OSCILLATING
    elif case == "zero":
        got, ref, it, oc, info = _eig_case(4, [(0, 1), (1, 2), (0, 2), (2, 3)], 100, 1e-4)
        want = O.Graph.EIG_ZERO
    elif case == "cap":
        got, ref, it, oc, info = _eig_case(3, [(0, 1), (1, 2), (2, 0)], 37, 1e-4)
        want = O.Graph.EIG_CAP
    elif case == "trivial":
        got, ref, it, oc, info = _eig_case(3, [], 100, 1e-4, live=np.array([0, 1, 0], np.uint8))
        want = O.Graph.EIG_TRIVIAL
    else:
        got, ref, it, oc, info = _eig_case(4, [(0, 1), (2, 3)], 0, 1e-4, live=np.array([1, 0, 1, 1], np.uint8))
        want = O.Graph.EIG_CAP
        assert got.tolist() == [1.0, 0.0, 1.001, 1.002]
    assert oc == want and info["outcome"] == want
    assert info["iterations"] == it, (info["iterations"], it)
    assert close(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("seed", range(3))
def test_random_multigraphs(seed):
    say("power multigraph", seed)
    rng = np.random.default_rng(300 + seed)
    n, s, d = graphs.random_multigraph(rng, 3000, 30000)
    live = (rng.random(n) > 0.05).astype(np.uint8)
    o = og(n, s, d, live)
    g = make(n, s, d, live)
    for alpha, beta in ((0.05, 1.0), (0.01, 0.3)):
        ref, it, l1 = o.katz(alpha, beta, 100, 1e-10)
        info = {}
        got = np.array(rg.katz(g, alpha, beta, 100, 1e-10, raw=True, info=info))
        assert info["iterations"] == it and close(got, ref), (it, info["iterations"], np.abs(got - ref).max())
    ref, it, l1 = o.alpha_centrality(0.08, 2.0, 60, 1e-12)
    info = {}
    got = np.array(rg.alpha_centrality(g, 0.08, 2.0, 60, 1e-12, raw=True, info=info))
    assert info["iterations"] == it and close(got, ref)
    ref, it, oc = o.eigenvector(300, 1e-9)
    info = {}
    got = np.array(rg.eigenvector(g, 300, 1e-9, raw=True, info=info))
    assert info["outcome"] == oc and info["iterations"] == it and close(got, ref), (oc, it, info)
    assert np.all(got[live == 0] == 0.0)
    rg.destroy(g)


def test_rmat():
    # hub rows long enough for the PageRank kernel's segmented rows
    say("power rmat")
    n, s, d = gen.rmat(15, 1 << 19, 44)
    o = og(n, s, d)
    g = make(n, s, d)
    ref, it, oc = o.eigenvector(200, 1e-10)
    info = {}
    got = np.array(rg.eigenvector(g, 200, 1e-10, raw=True, info=info))
    assert info["outcome"] == oc and info["iterations"] == it and close(got, ref), (oc, it, info)
    ref, it, _ = o.katz(0.002, 1.0, 100, 1e-9)
    info = {}
    got = np.array(rg.katz(g, 0.002, 1.0, 100, 1e-9, raw=True, info=info))
    assert info["iterations"] == it and close(got, ref)
    rg.destroy(g)


def test_errors():
    res = rg.BatchResource([(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))] * 2)
    out = np.zeros(res.n_cap)
    inf = _lib.ZgPrInfo()
    L = _lib.lib()
    rc = L.zg_katz(res.handle, 0.1, 1.0, 10, 1e-4, out.ctypes.data, None)
    assert rc == _lib.ZG_ERR_INVALID_ARG
    rc = L.zg_eigenvector(res.handle, 10, 1e-4, out.ctypes.data, inf)
    assert rc == _lib.ZG_ERR_INVALID_ARG
    res.release()
    g = make(0, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert rg.katz(g, raw=True) == [] and rg.eigenvector(g, raw=True) == []
    rg.destroy(g)


def test_node_degrees_core_numbers_and_copy_in_out_entry_points():
    # resource_graph.ex:2130-2136 / :562-599 and :2171-2192; Zog.Centrality
    # (centrality.ex:191-312) builds a resource per call on the same kernels
    say("degrees, cores, Zog.Centrality")
    from zog_amd import centrality as C
    rng = np.random.default_rng(31)
    n, s, d = graphs.random_multigraph(rng, 500, 4000)
    live = (rng.random(n) > 0.05).astype(np.uint8)
    g = make(n, s, d, live)
    deg = rg.node_degrees(g)
    ref = np.zeros(n, np.int64)
    for a, b in zip(s.tolist(), d.tolist()):
        if live[a] and live[b]:
            ref[a] += 1
    assert deg == ref.tolist()
    cores = rg.core_numbers(g, raw=True)
    assert cores == og(n, s, d, live).core_numbers().tolist()
    assert rg.core_numbers(g) == {i: c for i, c in enumerate(cores)}
    rg.destroy(g)
    # labelled builder: scores keyed by label, x0.5 for the undirected BC
    b = SoA.from_list("undirected", [("a", "b", 1.0), ("b", "c", 2.0), ("c", "d", 1.0), ("a", "c", 5.0)])
    bc = C.betweenness_unweighted(b)
    g = rg.new(b)
    assert bc == rg.betweenness_unweighted(g)
    assert C.betweenness_f64(b) == rg.betweenness_f64(g)
    assert C.pagerank(b) == rg.pagerank(g)
    assert C.eigenvector(b) == rg.eigenvector(g)
    assert C.katz(b) == rg.katz(g)
    assert C.alpha_centrality(b, alpha=0.2) == rg.alpha_centrality(g, alpha=0.2)
    assert rg.node_degrees(g) == [2, 2, 3, 1]
    rg.destroy(g)
    # the weighted path through the copy-in/out entry point: a -> c costs 3
    # via b, less than the direct 5.0, so b carries (a, c) and (c, a)
    w = C.betweenness_f64(b)
    assert w["b"] > 0.0 and w["a"] == 0.0
