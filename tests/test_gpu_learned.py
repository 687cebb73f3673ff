"""Learned pivot selection on the GPU against the oracles: core numbers,
node features, weighted Brandes and per-source dependency totals, the BF16
network's forward pass and gradient, training, and what the trained network
gains over uniform pivots."""
import sys

import numpy as np
import pytest

from tests import graphs
from zog_amd import generators as gen
from zog_amd import pivots as PV
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make(n, src, dst, live=None):
    return rg.new(SoA.from_arrays("directed", n, src, dst), live_mask=live)


def test_core_numbers_parity():
    from oracle import oracle as O
    say("core numbers")
    # the reference's C4-with-tail case (connectivity.zig:103-130)
    src = np.array([0, 1, 1, 2, 2, 3, 3, 0, 0, 4], np.uint32)
    dst = np.array([1, 0, 2, 1, 3, 2, 0, 3, 4, 0], np.uint32)
    g = make(5, src, dst)
    assert PV.core_numbers(g).tolist() == [2, 2, 2, 2, 1]
    rg.destroy(g)
    rng = np.random.default_rng(4)
    for n, m in ((500, 4000), (3000, 40000)):
        n, s, d = graphs.random_multigraph(rng, n, m)
        live = (rng.random(n) > 0.03).astype(np.uint8)
        g = make(n, s, d, live)
        assert np.array_equal(PV.core_numbers(g), O.Graph(n, s, d, live).core_numbers())
        rg.destroy(g)
    n, s, d = gen.rmat(14, 1 << 18, 8)
    g = make(n, s, d)
    assert np.array_equal(PV.core_numbers(g), O.Graph(n, s, d).core_numbers())
    rg.destroy(g)


def test_node_features_parity():
    from oracle import mlp as M
    from oracle import oracle as O
    say("features")
    rng = np.random.default_rng(6)
    n, s, d = gen.rmat(13, 1 << 16, 12)
    live = (rng.random(n) > 0.02).astype(np.uint8)
    og = O.Graph(n, s, d, live)
    pr, _, _ = og.pagerank(0.85, 100, 1e-6)
    ref = M.features(n, s, d, live.astype(bool), pr, og.core_numbers())
    g = make(n, s, d, live)
    got = PV.node_features(g).cpu().numpy()
    rg.destroy(g)
    assert np.abs(got - ref).max() < 2e-3, np.abs(got - ref).max(0)


@pytest.mark.parametrize("engine", ["lanes", "queue"])
def test_weighted_bc_and_dependency_totals(engine):
    from oracle import oracle as O
    say("weighted bc", engine)
    rng = np.random.default_rng(9)
    n, s, d = gen.rmat(11, 1 << 14, 5)
    s = np.concatenate([s, s[:300]])  # parallel arcs
    d = np.concatenate([d, d[:300]])
    og = O.Graph(n, s, d)
    sources = rng.integers(0, n, 150).astype(np.uint32)
    sources[:5] = sources[5]  # repeated sources count repeatedly
    weights = rng.uniform(0.1, 3.0, len(sources))
    ref_bc = np.zeros(n)
    ref_tot = np.zeros(len(sources))
    for i, x in enumerate(sources.tolist()):
        b, _ = og.betweenness(sources=np.array([x], np.uint32))
        ref_bc += weights[i] * b
        depth, *_ = og.bfs_paths(x)
        r = depth != O.NONE_U32
        ref_tot[i] = float(np.sum(depth[r].astype(np.int64) - 1) + 1)  # the source itself has depth 0
    g = make(n, s, d)
    rg.set_engine(g, engine)
    bc, tot = PV.bc_weighted(g, sources, weights, want_totals=True)
    assert np.array_equal(tot, ref_tot)
    scale = max(1.0, np.abs(ref_bc).max())
    assert np.all(np.abs(bc - ref_bc) <= 1e-9 * np.maximum(np.abs(ref_bc), 1e-9 * scale))
    # all live sources, totals only: sum of totals = sum of BC
    _, tot_all = PV.bc_weighted(g, None, None, want_bc=False, want_totals=True)
    full, _ = og.betweenness()
    assert abs(tot_all.sum() - full.sum()) <= 1e-9 * full.sum()
    rg.destroy(g)


def _rand_params(rng, scale=0.25):
    from oracle import mlp as M
    return rng.normal(0, scale, M.P).astype(np.float32)


def test_mlp_forward_parity():
    import torch
    from oracle import mlp as M
    say("mlp forward")
    rng = np.random.default_rng(1)
    sc = PV.Scorer(seed=3)
    for n in (1, 31, 1000, 70001):
        p = _rand_params(rng)
        sc.params = p
        x = rng.normal(0, 1, (n, 8)).astype(np.float32)
        y = sc.forward(torch.from_numpy(x).cuda()).cpu().numpy()
        ref, _ = M.forward(p, x, bf16_ops=True)
        err = np.abs(y - ref).max()
        assert err <= 1e-3 * max(1.0, np.abs(ref).max()), (n, err)
    sc.close()


def test_mlp_gradient_parity():
    import torch
    from oracle import mlp as M
    say("mlp gradient")
    rng = np.random.default_rng(2)
    sc = PV.Scorer(seed=4)
    blocks = [(0, 1024), (1024, 1152), (1152, 17536), (17536, 17664), (17664, 17792), (17792, 17793)]
    for n in (37, 128, 4000, 65536):
        p = _rand_params(rng)
        sc.params = p
        x = rng.normal(0, 1, (n, 8)).astype(np.float32)
        t = rng.normal(0, 1, n).astype(np.float32)
        loss, g = sc.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
        ref_loss, ref = M.gradient(p, x, t, bf16_ops=True)
        assert abs(loss - ref_loss) <= 1e-5 * ref_loss
        for a, b in blocks:
            num = np.linalg.norm(g[a:b] - ref[a:b])
            den = np.linalg.norm(ref[a:b])
            assert num <= 2e-3 * den + 1e-7, (n, a, num, den)
    sc.close()


def test_mlp_training_reduces_loss():
    import torch
    from oracle import mlp as M
    say("mlp training")
    rng = np.random.default_rng(5)
    teacher = _rand_params(rng, 0.2)
    x = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    t, _ = M.forward(teacher, x, bf16_ops=False)
    sc = PV.Scorer(seed=11)
    xt = torch.from_numpy(x).cuda()
    tt = torch.from_numpy(t.astype(np.float32)).cuda()
    l0, _ = sc.gradient(xt, tt)
    losses = sc.fit(xt, tt, epochs=30, batch=4096, lr=2e-3)
    assert losses[-1] < 0.1 * l0, (l0, losses[-1])
    # the Adam update equals the oracle's on one step from fresh moments
    sc.params = teacher
    g_loss, g = sc.gradient(xt[:4096].contiguous(), tt[:4096].contiguous())
    sc.adam(lr=1e-2)
    step = sc.params - teacher
    ref = -1e-2 * (g / (np.abs(g) + 1e-8))  # bias-corrected first step: m/sqrt(v) = sign(g)
    assert np.allclose(step, ref, atol=2e-6)
    sc.close()


def test_learned_pivots_gain_over_uniform():
    # train the scorer on small R-MAT graphs (labels from the GPU sweep), then
    # at a budget of n/64 pivots on an R-MAT 13 graph, learned pivots over the
    # folded frame must beat (a) uniform pivots over the sources that have a
    # successor and (b) uniform pivots over the same folded frame -- what the
    # trained network adds -- on the top-1% overlap with exact Brandes: the
    # mean paired difference over 12 draws, less two standard errors, above 0
    from scipy.stats import spearmanr
    say("learned pivots")
    X, T = PV.rmat_training_set(scales=(10, 11, 12), per_scale=4, seed=77, workers=2)
    sc = PV.Scorer(seed=5)
    p = sc.params
    p[-1] = float(T.mean())
    sc.params = p
    losses = sc.fit(X, T, epochs=40, batch=8192, lr=2e-3)
    assert losses[-1] < losses[0]
    n, s, d = gen.rmat(13, 16 << 13, 4242)
    g = make(n, s, d)
    exact, tot = PV.bc_weighted(g, None, None, want_bc=True, want_totals=True)
    pred, live_ids = PV.predict_totals(g, sc)
    rho = spearmanr(pred, tot).statistic
    frame = PV.sampling_frame(g)
    nonsink = np.nonzero(np.bincount(s, minlength=n) > 0)[0]
    k = n // 64
    rng = np.random.default_rng(3)
    ov = {"learned_folded": [], "uniform_nonsink": [], "uniform_folded": []}
    for _ in range(12):
        est, m = PV.sampled_betweenness_folded(g, k, rng, pred, frame)
        assert abs(m - k) <= 1
        ov["learned_folded"].append(PV.top_overlap(est, exact))
        est, _ = PV.sampled_betweenness(g, k, rng, None, nonsink)
        ov["uniform_nonsink"].append(PV.top_overlap(est, exact))
        est, _ = PV.sampled_betweenness_folded(g, k, rng, None, frame)
        ov["uniform_folded"].append(PV.top_overlap(est, exact))
    say("spearman", rho, {m: round(float(np.mean(v)), 4) for m, v in ov.items()})
    assert rho > 0.8
    learned = np.array(ov["learned_folded"])
    for base in ("uniform_nonsink", "uniform_folded"):
        diff = learned - np.array(ov[base])
        assert diff.mean() - 2 * diff.std(ddof=1) / np.sqrt(len(diff)) > 0, (base, diff)
    # the whole frame falls back to exact Brandes
    est, m = PV.sampled_betweenness_folded(g, len(frame[0]), rng, pred, frame)
    assert m == len(frame[0]) and np.allclose(est, exact, rtol=1e-12, atol=1e-9)
    rg.destroy(g)
    sc.close()


def _fold_graph(rng):
    # an R-MAT core plus chains and leaves: single-successor vertices into
    # hubs (folded), into other single-successor vertices (chains), into
    # sinks, with parallel arcs and self-loops on them
    n, s, d = gen.rmat(11, 1 << 14, 77)
    extra_s, extra_d = [], []
    for v in range(n - 300, n):
        k = int(rng.integers(0, 4))
        if k == 0:
            continue                                    # sink (or whatever R-MAT gave it)
        t = int(rng.integers(0, n - 300))
        extra_s += [v] * k
        extra_d += [t] * k                              # one successor, parallel arcs
        if rng.random() < 0.3:
            extra_s.append(v)
            extra_d.append(v)                           # and a self-loop
    s = np.concatenate([s, np.array(extra_s, np.uint32)])
    d = np.concatenate([d, np.array(extra_d, np.uint32)])
    return n, s, d


def _fold_map_numpy(n, s, d, live):
    keep = live[s] & live[d] & (s != d)
    pairs = np.unique(np.stack([s[keep], d[keep]], 1), axis=0)
    ndist = np.bincount(pairs[:, 0], minlength=n)
    cls = np.minimum(ndist, 2).astype(np.uint8)
    cls[~live] = 0
    succ = np.full(n, 0xFFFFFFFF, np.uint32)
    one = pairs[cls[pairs[:, 0]] == 1]
    succ[one[:, 0]] = one[:, 1]
    into = np.full(n, 0xFFFFFFFF, np.uint32)
    f = (cls == 1) & (cls[np.minimum(succ, n - 1)] == 2)
    into[f] = succ[f]
    cnt = np.bincount(into[f], minlength=n).astype(np.uint32)
    return cls, succ, into, cnt


def test_fold_map_matches_the_arc_lists():
    say("fold map")
    rng = np.random.default_rng(12)
    n, s, d = _fold_graph(rng)
    live = rng.random(n) > 0.02
    g = make(n, s, d, live.astype(np.uint8))
    got = PV.fold_map(g)
    ref = _fold_map_numpy(n, s, d, live)
    for a, b, name in zip(got, ref, ("class", "successor", "folded into", "count")):
        assert np.array_equal(a, b), name
    assert int(got[3].sum()) == int(np.count_nonzero(got[2] != 0xFFFFFFFF)) > 0
    rg.destroy(g)


def test_folded_sweep_against_the_oracle_groups():
    # each source's result = its own sweep plus the sweeps of the sources
    # folded into it (summed by the oracle source by source), weighted
    from oracle import oracle as O
    say("folded sweep")
    rng = np.random.default_rng(13)
    n, s, d = _fold_graph(rng)
    og = O.Graph(n, s, d)
    g = make(n, s, d)
    cls, succ, into, cnt = PV.fold_map(g)
    frame, folded = PV.sampling_frame(g)
    members = {}
    for v in np.nonzero(into != PV.NONE)[0].tolist():
        members.setdefault(int(into[v]), []).append(v)
    # the hubs with the most folded members, and random frame vertices, one repeated
    hubs = np.argsort(-cnt.astype(np.int64))[:20]
    pick = np.concatenate([hubs, rng.choice(frame, 60, replace=False), hubs[:2]]).astype(np.uint32)
    w = rng.uniform(0.5, 3.0, len(pick))
    ref = np.zeros(n)
    for m, wi in zip(pick.tolist(), w.tolist()):
        b, _ = og.betweenness(sources=np.array([m] + members.get(m, []), np.uint32))
        ref += wi * b
    got, _ = PV.bc_folded(g, pick, w)
    scale = max(1.0, np.abs(ref).max())
    assert np.all(np.abs(got - ref) <= 1e-9 * np.maximum(np.abs(ref), 1e-9 * scale))
    # a folded vertex is counted by its group, not as a source of its own
    with pytest.raises(rg._lib.ZogError):
        PV.bc_folded(g, np.array([np.nonzero(into != PV.NONE)[0][0]], np.uint32))
    rg.destroy(g)


def test_folded_full_frame_is_exact_brandes():
    # every frame vertex once with weight 1 covers every source exactly once
    from oracle import oracle as O
    say("folded full frame")
    rng = np.random.default_rng(14)
    n, s, d = _fold_graph(rng)
    g = make(n, s, d)
    frame, folded = PV.sampling_frame(g)
    est, m = PV.sampled_betweenness_folded(g, len(frame), rng, None, (frame, folded))
    assert m == len(frame) < n
    ref, _ = O.Graph(n, s, d).betweenness()
    scale = max(1.0, np.abs(ref).max())
    assert np.all(np.abs(est - ref) <= 1e-9 * np.maximum(np.abs(ref), 1e-9 * scale))
    rg.destroy(g)
