"""The allocation-free PageRank entry point (zg_pagerank_device_noalloc)
against zg_pagerank_device, through the C ABI: the same scores bit for bit
and the same iteration counts, on two calls in a row (the second on a smaller
graph, so the fixed scratch holds the first call's contents), in every
precision mode; device memory in use does not change across calls once warm;
a graph beyond the scratch is rejected with ZG_ERR_SCRATCH and leaves the
entry point usable."""
import ctypes
import sys

import numpy as np
import pytest

from zog_amd import _lib
from zog_amd import generators as gen
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def run(g, fn, damping, iters, tol, precision):
    import torch
    V = g.resource.n_cap
    out = torch.empty(max(V, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    info = _lib.ZgPrInfo()
    rc = getattr(_lib.lib(), fn)(g.resource.handle, damping, iters, tol, precision, ctypes.c_void_p(out.data_ptr()),
                                 None, ctypes.byref(info))
    return rc, out[:V].cpu().numpy(), info


def compare(g, damping, iters, tol, precision):
    rc0, ref, i0 = run(g, "zg_pagerank_device", damping, iters, tol, precision)
    rc1, got, i1 = run(g, "zg_pagerank_device_noalloc", damping, iters, tol, precision)
    assert rc0 == 0 and rc1 == 0, (rc0, rc1)
    assert np.array_equal(got, ref)
    assert (i1.iterations, i1.fp64_iterations, i1.last_l1) == (i0.iterations, i0.fp64_iterations, i0.last_l1)
    assert i1.residual_l1 == i0.residual_l1
    return got, i1


def make(n, src, dst):
    return rg.new(SoA.from_arrays("directed", n, src, dst))


def test_noalloc_pagerank_matches_the_workspace_entry_point():
    import torch
    big = make(*gen.rmat(18, 1 << 22, 31))          # 262,144 nodes, long rows (segment partials)
    small = make(*gen.erdos_renyi_directed(5000, 8.0 / 4999, 32))
    for prec in (_lib.ZG_PR_FP32, _lib.ZG_PR_FP32 | _lib.ZG_PR_SKIP_RESIDUAL, _lib.ZG_PR_FP64, _lib.ZG_PR_MIXED):
        say("noalloc precision", prec)
        got, info = compare(big, 0.85, 100, 1e-6 if prec != _lib.ZG_PR_MIXED else 1e-10, prec)
        assert info.iterations > 1
        compare(small, 0.85, 100, 1e-6, prec)   # second call, smaller graph
    # fixed iteration count, damping 0.99, and n = 1 / n = 0 graphs
    compare(small, 0.99, 37, 0.0, _lib.ZG_PR_FP32)
    one = make(*gen.erdos_renyi_directed(1, 0.0, 1))
    got, info = compare(one, 0.85, 100, 1e-6, _lib.ZG_PR_FP32)
    assert got.tolist() == [1.0]
    empty = rg.new(SoA.from_arrays("directed", 3, np.zeros(0, np.uint32), np.zeros(0, np.uint32)),
                   live_mask=np.zeros(3, np.uint8))
    got, info = compare(empty, 0.85, 100, 1e-6, _lib.ZG_PR_FP32)
    assert got.tolist() == [0.0, 0.0, 0.0]
    # no device memory taken once warm
    run(big, "zg_pagerank_device_noalloc", 0.85, 100, 1e-6, _lib.ZG_PR_MIXED)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        rc, _, _ = run(big, "zg_pagerank_device_noalloc", 0.85, 100, 1e-6, _lib.ZG_PR_MIXED)
        assert rc == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    for g in (big, small, one, empty):
        rg.destroy(g)


def test_noalloc_pagerank_rejects_graphs_beyond_the_scratch():
    # 3 M node ids: 48 bytes each with the residual check (144 MB) pass the
    # 64 MiB scratch; FP32 without it (16 bytes each, 48 MB) fits
    V = 3_000_000
    rng = np.random.default_rng(2)
    src = rng.integers(0, V, 200000).astype(np.uint32)
    dst = rng.integers(0, V, 200000).astype(np.uint32)
    g = make(V, src, dst)
    rc, _, _ = run(g, "zg_pagerank_device_noalloc", 0.85, 20, 0.0, _lib.ZG_PR_FP32)
    assert rc == _lib.ZG_ERR_SCRATCH
    with pytest.raises(_lib.ZogError):
        _lib.check(rc, "zg_pagerank_device_noalloc")
    compare(g, 0.85, 20, 0.0, _lib.ZG_PR_FP32 | _lib.ZG_PR_SKIP_RESIDUAL)
    small = make(*gen.erdos_renyi_directed(2000, 8.0 / 1999, 5))
    compare(small, 0.85, 50, 1e-6, _lib.ZG_PR_FP64)
    rg.destroy(g)
    rg.destroy(small)
