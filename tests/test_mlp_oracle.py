"""Pins of the learned-part oracles (CPU): the network's gradient against
central finite differences, BF16 rounding against hand-worked values, core
numbers against the reference's own tests and networkx, features against
closed forms."""
import numpy as np
import pytest

from oracle import mlp as M
from zog_amd import generators as gen


def rand_params(rng, scale=0.3):
    return rng.normal(0, scale, M.P)


def test_param_layout():
    assert M.P == 17793
    p = np.arange(M.P, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = M.unpack(p)
    assert W1.shape == (8, 128) and W1[1, 0] == 128 and b1[0] == 1024
    assert W2[0, 1] == 1153 and b2[0] == 1152 + 16384 and w3[-1] == M.P - 2 and b3 == M.P - 1
    assert np.array_equal(M.pack(W1, b1, W2, b2, w3, b3), p)


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rand_params(rng)
    x = rng.normal(0, 1, (9, 8))
    t = rng.normal(0, 1, 9)
    loss, g = M.gradient(p, x, t, bf16_ops=False)
    idx = np.concatenate([rng.choice(M.P - 513, 300, replace=False),
                          np.arange(M.P - 513, M.P)])  # b2, w3, b3 and the tail of W2
    idx = np.unique(np.concatenate([idx, np.arange(0, 1152, 37)]))  # W1 and b1
    eps = 1e-6
    for i in idx:
        d = np.zeros(M.P)
        d[i] = eps
        lp, _ = M.gradient(p + d, x, t, bf16_ops=False)
        lm, _ = M.gradient(p - d, x, t, bf16_ops=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the loss itself
    y, _ = M.forward(p, x, bf16_ops=False)
    assert loss == pytest.approx(np.mean((y - t) ** 2), rel=1e-14)


def test_forward_closed_form():
    # one active path: x0 -> neuron 5 -> neuron 7 -> y
    W1 = np.zeros((8, 128)); W1[0, 5] = 2.0
    W2 = np.zeros((128, 128)); W2[5, 7] = 3.0
    w3 = np.zeros(128); w3[7] = 0.5
    p = M.pack(W1, np.zeros(128), W2, np.zeros(128), w3, 0.25)
    x = np.zeros((3, 8)); x[:, 0] = [1.0, -1.0, 0.5]
    y, _ = M.forward(p, x)
    assert np.allclose(y, [0.25 + 3.0, 0.25, 0.25 + 1.5])


def test_bf16_rounding():
    one = 1.0
    ulp = 2.0 ** -7
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4 * 3, -2.5, 3.0e38, 0.0])
    got = M.bf16(vals)
    # halfway cases go to even; 1 + 3ulp/2 -> 1 + 2ulp; 1 + 3ulp/4 -> 1 + ulp
    assert got.tolist() == [one, one, one + 2 * ulp, one + ulp, -2.5, float(M.bf16([3.0e38])[0]), 0.0]
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8


def test_core_numbers_reference_cases():
    from oracle import oracle as O
    # connectivity.zig:103-130 / connectivity_test.exs:7-31: C4 plus a tail
    src = np.array([0, 1, 1, 2, 2, 3, 3, 0, 0, 4], np.uint32)
    dst = np.array([1, 0, 2, 1, 3, 2, 0, 3, 4, 0], np.uint32)
    assert O.Graph(5, src, dst).core_numbers().tolist() == [2, 2, 2, 2, 1]
    # complete graph K5: every core number is 4; an isolated node is 0
    s, d = zip(*[(i, j) for i in range(5) for j in range(5) if i != j])
    c = O.Graph(6, np.array(s, np.uint32), np.array(d, np.uint32)).core_numbers()
    assert c.tolist() == [4, 4, 4, 4, 4, 0]


@pytest.mark.parametrize("seed", range(3))
def test_core_numbers_match_networkx(seed):
    nx = pytest.importorskip("networkx")
    from oracle import oracle as O
    from tests import graphs
    rng = np.random.default_rng(seed)
    n, src, dst = graphs.random_multigraph(rng, 400, 3000)  # parallel arcs and self-loops
    live = rng.random(n) > 0.05
    c = O.Graph(n, src, dst, live.astype(np.uint8)).core_numbers()
    G = nx.Graph()
    G.add_nodes_from(np.nonzero(live)[0].tolist())
    G.add_edges_from((int(a), int(b)) for a, b in zip(src, dst) if a != b and live[a] and live[b])
    ref = nx.core_number(G)
    for v in range(n):
        assert c[v] == (ref[v] if live[v] else 0)


def test_features_closed_form():
    # directed cycle: every node alike, so every standardised column is 0
    n = 6
    src = np.arange(n, dtype=np.uint32)
    dst = (src + 1) % n
    f = M.features(n, src, dst, None, np.full(n, 1.0 / n), np.full(n, 1))
    assert np.allclose(f[:, :6], 0.0) and np.all(f[:, 6] == 1.0) and np.all(f[:, 7] == 0.0)
    # a star 0 -> 1..4: raw columns known, standardisation over 5 nodes
    src = np.zeros(4, np.uint32)
    dst = np.arange(1, 5, dtype=np.uint32)
    f = M.features(5, src, dst, None, np.full(5, 0.2), np.ones(5))
    raw_out = np.log1p([4, 0, 0, 0, 0])
    z = (raw_out - raw_out.mean()) / raw_out.std()
    assert np.allclose(f[:, 2], z)
    # 2-hop: successors' out-degrees (leaves have none)
    assert np.allclose(f[:, 3], 0.0)
    n2, s2, d2 = gen.rmat(8, 1500, 2)
    f2 = M.features(n2, s2, d2, None, np.full(n2, 1.0 / n2), np.zeros(n2))
    assert np.allclose(f2[:, :6].mean(0), 0.0, atol=1e-12)
    sd = f2[:, :6].std(0)
    assert np.all((np.abs(sd - 1) < 1e-9) | (sd < 1e-12))


def test_inclusion_probabilities_and_systematic_sample():
    from zog_amd import pivots as PV
    rng = np.random.default_rng(0)
    score = np.concatenate([rng.exponential(1.0, 900), np.zeros(100), [1e6]])
    for k in (1, 10, 200, 1000):
        pi = PV.inclusion_probabilities(score, k, eps=0.05)
        assert abs(pi.sum() - k) < 1e-6 and np.all(pi <= 1.0) and np.all(pi > 0)
        if k >= 10:
            assert pi[-1] == 1.0  # the dominant score is taken for sure
        order = np.argsort(-score, kind="stable")
        picks = [PV.systematic_sample(pi, order, np.random.default_rng(r)) for r in range(200)]
        assert all(abs(len(p) - k) <= 1 and len(set(p.tolist())) == len(p) for p in picks)
        freq = np.bincount(np.concatenate(picks), minlength=len(score)) / 200
        assert np.abs(freq - pi).max() < 0.15  # empirical inclusion frequency tracks pi
    # Horvitz-Thompson: unbiased for a fixed vector of per-source values
    vals = rng.random(len(score))
    pi = PV.inclusion_probabilities(score, 300, eps=0.2)
    order = np.argsort(-score, kind="stable")
    ests = []
    for r in range(3000):
        p = PV.systematic_sample(pi, order, np.random.default_rng(10_000 + r))
        ests.append(np.sum(vals[p] / pi[p]))
    assert abs(np.mean(ests) - vals.sum()) < 0.01 * vals.sum()


def test_top_overlap_ties_and_uniform_pivots():
    from zog_amd import pivots as PV
    exact = np.array([5.0, 4.0, 4.0, 4.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0] * 10)
    # the top 1% (1 node) by estimate is any node tied at the exact maximum
    est = exact.copy()
    assert PV.top_overlap(est, exact, frac=0.01) == 1.0
    # top 4%: threshold 4.0; estimates that pick tied nodes in another order still hit
    est2 = exact + np.linspace(0, 0.1, len(exact))
    assert PV.top_overlap(est2, exact, frac=0.04) == 1.0
    assert PV.top_overlap(-exact, exact, frac=0.04) == 0.0
    src, w = PV.uniform_pivots(np.arange(100), 30, np.random.default_rng(1))
    assert len(set(src.tolist())) == 30 and np.allclose(w, 100 / 30)
