"""The named synthetic workloads (sizes and seeds fixed here once).

  er10k      directed Erdos-Renyi G(n=10,000, p=8/(n-1)), all sources
  rmat20     R-MAT scale 20, 2^24 distinct arcs (a=.57 b=.19 c=.19 d=.05),
             4,096 seeded sources
  grid2048   undirected 2048x2048 grid, 5% of edges (419,226) removed,
             1,024 seeded sources
  rmat20w    rmat20's graph and sources with seeded arc weights uniform in
             (0, 1] (weighted Brandes)
"""
from __future__ import annotations

from . import generators as gen

SEEDS = {"er10k": gen.SEED_BASE + 1, "rmat20": gen.SEED_BASE + 2, "ba512": gen.SEED_BASE + 3,
         "grid2048": gen.SEED_BASE + 4, "rmat22": gen.SEED_BASE + 5, "rmat20w": gen.SEED_BASE + 6}
SOURCE_SEED_OFFSET = 0x1000


def er10k():
    n = 10_000
    return gen.erdos_renyi_directed(n, 8.0 / (n - 1), SEEDS["er10k"])


def rmat(scale: int, arcs: int, seed: int):
    return gen.rmat(scale, arcs, seed)


def rmat20():
    return gen.rmat(20, 1 << 24, SEEDS["rmat20"])


def rmat20_sources(n, src, dst, k: int = 4096):
    return gen.seeded_sources(n, src, dst, k, SEEDS["rmat20"] + SOURCE_SEED_OFFSET)


def arc_weights(m: int, seed: int):
    """f64[m] uniform in (0, 1]."""
    import numpy as np
    return 1.0 - np.random.default_rng(seed).random(m)


def rmat20w_weights(m: int):
    return arc_weights(m, SEEDS["rmat20w"])


def grid2048():
    return gen.grid_2d(2048, 2048, SEEDS["grid2048"], delete=419_226)


def grid2048_sources(n, src, dst, k: int = 1024):
    return gen.seeded_sources(n, src, dst, k, SEEDS["grid2048"] + SOURCE_SEED_OFFSET)
