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
    # train the scorer on small R-MAT graphs (labels from the GPU sweep),
    # then at the same source budget learned pivots must beat uniform pivots
    # on the top-1% overlap with exact Brandes, and rank sources well
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
    rng = np.random.default_rng(3)
    k = n // 2
    ov = {"uniform": [], "learned": []}
    for _ in range(6):
        est, m = PV.sampled_betweenness(g, k, rng, None, live_ids)
        ov["uniform"].append(PV.top_overlap(est, exact))
        est, m = PV.sampled_betweenness(g, k, rng, pred, live_ids)
        assert abs(m - k) <= 1
        ov["learned"].append(PV.top_overlap(est, exact))
    say("spearman", rho, "overlap uniform", np.mean(ov["uniform"]), "learned", np.mean(ov["learned"]))
    assert rho > 0.8
    assert np.mean(ov["learned"]) > np.mean(ov["uniform"])
    # the full budget falls back to exact Brandes
    est, m = PV.sampled_betweenness(g, n, rng, pred, live_ids)
    assert m == n and np.allclose(est, exact, rtol=1e-12, atol=0)
    rg.destroy(g)
    sc.close()
