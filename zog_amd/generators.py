"""Seeded synthetic graphs for the parity tests and the benchmark.

Every generator returns ``(n, src, dst)``: node count and u32 arc arrays in
SoA insertion order (an undirected edge contributes (u, v) then (v, u), as
``Zog.SoA.add_edge`` does, lib/zog/soa.ex:110-128).  All randomness comes from
numpy's PCG64 seeded explicitly (the Barabasi-Albert generator, written in C
in csrc/gen.c, uses xoshiro256** seeded by SplitMix64), so a (generator,
parameters, seed) triple always yields the same arrays.

  * ``erdos_renyi_directed``: G(n, p) by the Batagelj-Brandes geometric-skip
    method over the n(n-1) ordered pairs, the method of
    lib/zog/generator.ex:97-120 (same index -> (source, target) map).
  * ``rmat``: R-MAT (Chakrabarti, Zhan, Faloutsos 2004) with quadrant
    probabilities a, b, c, d; arcs are drawn until the requested number of
    distinct arcs exist; self-loops are kept.
  * ``grid_2d``: row-major ids, right then down edge per cell as in
    lib/zog/generator.ex:277-305, optionally with an exact number of edges
    removed uniformly without replacement.
  * ``barabasi_albert``: preferential attachment as in
    lib/zog/generator.ex:131-198 (clique of m, degree-weighted target list,
    m distinct targets per new node, ascending); ``ba_batch`` draws the
    node counts of a batch log-uniformly.
"""
from __future__ import annotations

import math

import numpy as np

SEED_BASE = 0x20C00000


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed)))


def erdos_renyi_directed(n: int, p: float, seed: int):
    rng = _rng(seed)
    if n < 2 or p <= 0.0:
        return n, np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    max_idx = n * (n - 1)
    if p >= 1.0:
        idx = np.arange(max_idx, dtype=np.int64)
    else:
        ln_p = math.log(1.0 - p)
        chunks = []
        last = -1
        expected = int(max_idx * p) + 16
        while last < max_idx:
            r = rng.random(max(expected // 4, 1024))
            skip = np.floor(np.log1p(-r) / ln_p).astype(np.int64)
            pos = last + np.cumsum(skip + 1)
            chunks.append(pos)
            last = int(pos[-1])
        idx = np.concatenate(chunks)
        idx = idx[idx < max_idx]
    source = idx // (n - 1)
    rem = idx % (n - 1)
    target = np.where(rem < source, rem, rem + 1)
    return n, source.astype(np.uint32), target.astype(np.uint32)


def rmat(scale: int, n_arcs: int, seed: int, a=0.57, b=0.19, c=0.19):
    """R-MAT on 2^scale nodes: n_arcs distinct arcs in first-drawn order
    (self-loops kept), generated natively."""
    import ctypes
    L = _genlib()
    n = 1 << scale
    src = np.empty(n_arcs, dtype=np.uint32)
    dst = np.empty(n_arcs, dtype=np.uint32)
    got = L.zgen_rmat(scale, n_arcs, int(seed) & ((1 << 64) - 1), a, b, c,
                      ctypes.c_void_p(src.ctypes.data), ctypes.c_void_p(dst.ctypes.data))
    if got != n_arcs:
        raise ValueError("invalid R-MAT parameters")
    return n, src, dst


def grid_2d(rows: int, cols: int, seed: int = 0, delete: int = 0):
    """Undirected grid; `delete` edges removed uniformly without replacement."""
    r = np.arange(rows, dtype=np.int64)[:, None]
    c = np.arange(cols, dtype=np.int64)[None, :]
    u = (r * cols + c)
    right_ok = np.broadcast_to(c + 1 < cols, (rows, cols))
    down_ok = np.broadcast_to(r + 1 < rows, (rows, cols))
    # per cell: right edge first, then down edge
    eu = np.stack([u, u], axis=-1).reshape(-1)
    ev = np.stack([u + 1, u + cols], axis=-1).reshape(-1)
    ok = np.stack([right_ok, down_ok], axis=-1).reshape(-1)
    eu, ev = eu[ok], ev[ok]
    if delete:
        rng = _rng(seed)
        drop = rng.choice(len(eu), size=int(delete), replace=False)
        keep = np.ones(len(eu), dtype=bool)
        keep[drop] = False
        eu, ev = eu[keep], ev[keep]
    src = np.stack([eu, ev], axis=-1).reshape(-1).astype(np.uint32)
    dst = np.stack([ev, eu], axis=-1).reshape(-1).astype(np.uint32)
    return rows * cols, src, dst


_gen_lib = None


def _genlib():
    global _gen_lib
    if _gen_lib is None:
        import ctypes
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libzog_gen.so")
        L = ctypes.CDLL(path)
        L.zgen_ba_arcs.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.zgen_ba_arcs.restype = ctypes.c_uint64
        L.zgen_barabasi_albert.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.zgen_barabasi_albert.restype = ctypes.c_uint64
        L.zgen_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.zgen_rmat.restype = ctypes.c_uint64
        _gen_lib = L
    return _gen_lib


def barabasi_albert(n: int, m: int, seed: int):
    """Undirected BA graph (both arcs per edge), generated natively."""
    if not (n > m >= 1):
        raise ValueError("need n > m >= 1")
    L = _genlib()
    k = int(L.zgen_ba_arcs(n, m))
    src = np.empty(k, dtype=np.uint32)
    dst = np.empty(k, dtype=np.uint32)
    got = int(L.zgen_barabasi_albert(n, m, int(seed) & ((1 << 64) - 1), src.ctypes.data, dst.ctypes.data))
    if got != k:
        raise RuntimeError("barabasi_albert generation failed")
    return n, src, dst


def ba_batch(n_graphs: int, m: int, lo: int, hi: int, seed: int):
    """n_graphs BA graphs with node counts floor(exp(U(ln lo, ln hi))), seeded."""
    rng = _rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), n_graphs))).astype(np.int64)
    sizes = np.maximum(sizes, m + 1)
    return [barabasi_albert(int(sz), m, seed * 1000003 + i) for i, sz in enumerate(sizes)]


def seeded_sources(n: int, src: np.ndarray, dst: np.ndarray, k: int, seed: int) -> np.ndarray:
    """First k ids of a seeded permutation of the non-isolated nodes."""
    deg = np.bincount(src, minlength=n) + np.bincount(dst, minlength=n)
    cand = np.nonzero(deg > 0)[0].astype(np.uint32)
    perm = _rng(seed).permutation(len(cand))
    return cand[perm[: min(k, len(cand))]]
