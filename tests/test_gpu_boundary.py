"""Guarantees of the C-ABI boundary on the GPU.

* Calls on one resource from several threads at once (the reference's NIFs
  run on dirty-CPU schedulers, so several Elixir processes may call BC and
  PageRank on the same resource concurrently, lib/zog/resource_graph.ex:48-54)
  give the results of the same calls made one after another; so do calls on
  two handles at once.
* A workspace larger than the device's memory returns ZG_ERR_GPU_OOM (the
  Python mirror raises ZogError, as Zigler raises a returned Zig error), and
  the handle stays usable at a smaller launch shape.
"""
import sys
import threading

import numpy as np
import pytest

from zog_amd import _lib
from zog_amd import generators as gen
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make(n, src, dst):
    return rg.new(SoA.from_arrays("directed", n, src, dst))


def close(a, b):
    # every sum of the library has a fixed order, so a repeated call gives
    # the same bits
    return bool(np.array_equal(a, b))


def test_concurrent_calls_on_one_handle():
    n, src, dst = gen.rmat(13, 1 << 17, 17)
    sources = gen.seeded_sources(n, src, dst, 300, 4)
    g = make(n, src, dst)
    ref_bc = rg.betweenness_unweighted_array(g, sources=sources)
    ref_all = rg.betweenness_unweighted_array(g)
    ref_pr = rg.pagerank_array(g, 0.85, 60, 0.0)
    ref_pr64 = rg.pagerank_array(g, 0.85, 40, 0.0, precision=64)
    jobs = [("bc", lambda: rg.betweenness_unweighted_array(g, sources=sources), ref_bc),
            ("bc_all", lambda: rg.betweenness_unweighted_array(g), ref_all),
            ("pr", lambda: rg.pagerank_array(g, 0.85, 60, 0.0), ref_pr),
            ("pr64", lambda: rg.pagerank_array(g, 0.85, 40, 0.0, precision=64), ref_pr64)]
    out, errors = [], []

    def worker(i):
        try:
            for r in range(4):
                name, fn, ref = jobs[(i + r) % len(jobs)]
                out.append((name, fn(), ref))
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    assert len(out) == 32
    for name, got, ref in out:
        assert np.array_equal(got, ref), name
    say("concurrent calls:", len(out), "results equal to the sequential ones")
    rg.destroy(g)


def test_concurrent_calls_on_two_handles():
    a = gen.rmat(12, 1 << 16, 21)
    b = gen.erdos_renyi_directed(3000, 8.0 / 2999, 22)
    ga, gb = make(*a), make(*b)
    refs = [rg.betweenness_unweighted_array(ga), rg.betweenness_unweighted_array(gb),
            rg.pagerank_array(ga, 0.85, 50, 0.0), rg.pagerank_array(gb, 0.85, 50, 0.0)]
    got = [None] * 4

    def run(k):
        for _ in range(3):
            if k == 0:
                got[0] = rg.betweenness_unweighted_array(ga)
            elif k == 1:
                got[1] = rg.betweenness_unweighted_array(gb)
            elif k == 2:
                got[2] = rg.pagerank_array(ga, 0.85, 50, 0.0)
            else:
                got[3] = rg.pagerank_array(gb, 0.85, 50, 0.0)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads)
    for k in range(4):
        assert np.array_equal(got[k], refs[k]), k
    rg.destroy(ga)
    rg.destroy(gb)


def test_gpu_oom_is_an_error_and_the_handle_stays_usable():
    # 2^24 node ids: 64 groups of the lane engine need 64 * 2^24 rows of 64
    # path counts (550 GB) -- more than the device holds
    from oracle import oracle as O
    V = 1 << 24
    rng = np.random.default_rng(8)
    src = rng.integers(0, 2000, 20000).astype(np.uint32)
    dst = rng.integers(0, 2000, 20000).astype(np.uint32)
    g = make(V, src, dst)
    rg.set_engine(g, "lanes")
    rg.set_batch_groups(g, 64)
    sources = np.arange(0, 2000, 100, dtype=np.uint32)
    with pytest.raises(_lib.ZogError) as e:
        rg.betweenness_unweighted_array(g, sources=sources)
    assert e.value.code == _lib.ZG_ERR_GPU_OOM
    say("OOM:", e.value)
    # the handle still works at one group per launch, and for PageRank
    rg.set_batch_groups(g, 1)
    got = rg.betweenness_unweighted_array(g, sources=sources)
    og = O.Graph(V, src, dst)
    ref, _ = og.betweenness(sources=sources)
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.all(np.abs(got - ref) <= 1e-9 * np.maximum(np.abs(ref), 1e-9 * scale))
    pr = rg.pagerank_array(g, 0.85, 30, 0.0)
    pref, _, _ = og.pagerank(0.85, 30, 0.0)
    assert np.abs(pr - pref).sum() <= 1e-6
    rg.destroy(g)


@pytest.mark.parametrize("engine", ["lanes", "queue", "weighted"])
def test_repeated_calls_give_the_same_bits(engine):
    # hubs with several successor chunks (lane engine), many sources per
    # batch adding into the same vertices (queue and weighted engines)
    rng = np.random.default_rng(40)
    n, src, dst = gen.rmat(13, 1 << 17, 41)
    hub = np.full(3000, 5, np.uint32)
    src = np.concatenate([src, hub, rng.integers(0, n, 2000).astype(np.uint32)])
    dst = np.concatenate([dst, rng.integers(0, n, 3000).astype(np.uint32), np.full(2000, 5, np.uint32)])
    sources = gen.seeded_sources(n, src, dst, 700, 6)
    if engine == "weighted":
        w = rng.uniform(0.1, 1.0, len(src))
        g = rg.new(SoA.from_arrays("directed", n, src, dst, w))
        runs = [rg.betweenness_f64_array(g, sources=sources) for _ in range(3)]
    else:
        g = make(n, src, dst)
        rg.set_engine(g, engine)
        runs = [rg.betweenness_unweighted_array(g, sources=sources) for _ in range(3)]
    assert runs[0].max() > 0
    for r in runs[1:]:
        assert np.array_equal(r, runs[0])
    rg.destroy(g)
