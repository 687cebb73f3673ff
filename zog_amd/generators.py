"""Seeded synthetic graphs for the parity tests and the benchmark.

Every generator returns ``(n, src, dst)``: node count and u32 arc arrays in
SoA insertion order (an undirected edge contributes (u, v) then (v, u), as
``Zog.SoA.add_edge`` does, lib/zog/soa.ex:110-128).  All randomness comes from
numpy's PCG64 seeded explicitly, so a (generator, parameters, seed) triple
always yields the same arrays.

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
    m distinct targets per new node, ascending).
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
    rng = _rng(seed)
    n = 1 << scale
    seen_keys = []
    n_unique = 0
    uniq = np.zeros(0, dtype=np.uint64)
    first_pos = np.zeros(0, dtype=np.int64)
    offset = 0
    all_keys = []
    while n_unique < n_arcs:
        M = max(int((n_arcs - n_unique) * 1.15), 1 << 16)
        src = np.zeros(M, dtype=np.uint64)
        dst = np.zeros(M, dtype=np.uint64)
        for lvl in range(scale):
            u = rng.random(M)
            sbit = (u >= a + b).astype(np.uint64)
            dbit = (((u >= a) & (u < a + b)) | (u >= a + b + c)).astype(np.uint64)
            sh = np.uint64(scale - 1 - lvl)
            src |= sbit << sh
            dst |= dbit << sh
        keys = (src << np.uint64(32)) | dst
        all_keys.append(keys)
        offset += M
        cat = np.concatenate(all_keys)
        uniq, first_pos = np.unique(cat, return_index=True)
        n_unique = len(uniq)
        all_keys = [cat]
    order = np.sort(first_pos)[:n_arcs]
    keys = all_keys[0][order]
    src = (keys >> np.uint64(32)).astype(np.uint32)
    dst = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
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


def barabasi_albert(n: int, m: int, seed: int):
    """Undirected BA graph (both arcs per edge)."""
    if not (n > m >= 1):
        raise ValueError("need n > m >= 1")
    rng = _rng(seed)
    eu, ev = [], []
    for uu in range(m - 1):
        for vv in range(uu + 1, m):
            eu.append(uu)
            ev.append(vv)
    cap = m * (m - 1) + (n - m) * 2 * m + 1
    deg = np.empty(cap, dtype=np.int64)
    if m > 1:
        base = np.repeat(np.arange(m), m - 1)
    else:
        base = np.zeros(1, dtype=np.int64)
    deg[: len(base)] = base
    size = len(base)
    # draw uniform positions in bulk; rejection of repeats follows the method
    buf = rng.random(1 << 16)
    bi = 0
    for uu in range(m, n):
        chosen = []
        while len(chosen) < m:
            if bi == len(buf):
                buf = rng.random(1 << 16)
                bi = 0
            node = int(deg[int(buf[bi] * size)])
            bi += 1
            if node not in chosen:
                chosen.append(node)
        chosen.sort()
        for vv in chosen:
            eu.append(uu)
            ev.append(vv)
        add = [uu] + chosen + [uu] * (m - 1)
        deg[size: size + len(add)] = add
        size += len(add)
    eu = np.asarray(eu, dtype=np.int64)
    ev = np.asarray(ev, dtype=np.int64)
    src = np.stack([eu, ev], axis=-1).reshape(-1).astype(np.uint32)
    dst = np.stack([ev, eu], axis=-1).reshape(-1).astype(np.uint32)
    return n, src, dst


def seeded_sources(n: int, src: np.ndarray, dst: np.ndarray, k: int, seed: int) -> np.ndarray:
    """First k ids of a seeded permutation of the non-isolated nodes."""
    deg = np.bincount(src, minlength=n) + np.bincount(dst, minlength=n)
    cand = np.nonzero(deg > 0)[0].astype(np.uint32)
    perm = _rng(seed).permutation(len(cand))
    return cand[perm[: min(k, len(cand))]]
