"""Autotuner of the Brandes sweep's launch shape.

The candidates are the knobs the library exposes per handle:
  * batch width: 64-source groups swept by one set of launches
    (zg_set_batch_groups, 1..64),
  * the push/pull switch factor alpha (a level pushes from its frontier when
    the frontier's successor arcs times alpha are below the arcs of the
    group's span; ZG_TUNE_PUSH_ALPHA),
  * workgroups per group of the pull kernel (ZG_TUNE_PULL_GRID),
  * the source grouping (zg_set_source_order: by distance to the hub core, or
    by discovery rank in a search from it).
Each candidate sweeps a fixed sample of the sources (the same sample for all)
on the device; the tuner keeps the one with the most traversed arcs per
kernel millisecond (info.kernel_ms is HIP-event time of the sweep) and leaves
it set on the handle.  Results are the same for every candidate up to FP64
summation order.
"""
from __future__ import annotations

import ctypes
import itertools
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .resource_graph import _ptr


def _set(handle, groups, alpha, pull_grid, order):
    L = _lib.lib()
    rc = L.zg_set_source_order(handle, int(order))
    if rc < 0:
        _lib.check(rc, "zg_set_source_order")
    rc = L.zg_set_batch_groups(handle, int(groups))
    if rc < 0:
        _lib.check(rc, "zg_set_batch_groups")
    _lib.check(L.zg_set_tuning(handle, _lib.ZG_TUNE_PUSH_ALPHA, float(alpha)), "zg_set_tuning")
    _lib.check(L.zg_set_tuning(handle, _lib.ZG_TUNE_PULL_GRID, float(pull_grid)), "zg_set_tuning")


def _sweep(handle, n_cap, sources):
    out = np.empty(n_cap, dtype=np.float64)
    inf = _lib.ZgBcInfo()
    rc = _lib.lib().zg_bc_unweighted(handle, _ptr(sources), len(sources), _ptr(out), ctypes.byref(inf))
    _lib.check(rc, "zg_bc_unweighted")
    return inf


def tune(resource, sample: np.ndarray, groups: Sequence[int] = (4,), alphas: Sequence[float] = (4.0, 8.0, 16.0),
         pull_grids: Sequence[int] = (4096, 16384, 0), orders: Sequence[int] = (1,), repeats: int = 2,
         max_state_bytes: Optional[float] = None, span_vertices: Optional[int] = None, log=None) -> dict:
    """Times every (groups, alpha, pull_grid, order) candidate on `sample` (u32
    source ids of `resource`) and sets the fastest.  Candidates whose
    per-group state (about 1.6 KiB per vertex of a group's span, at most
    span_vertices, default all) would exceed max_state_bytes are skipped.  Returns {"best": {...}, "trials": [...]}."""
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    h = resource.handle
    n_cap = resource.n_cap
    trials = []
    best = None
    for gr, al, pg, od in itertools.product(groups, alphas, pull_grids, orders):
        if max_state_bytes is not None and gr * 1.6 * 1024 * (span_vertices or n_cap) > max_state_bytes:
            continue
        _set(h, gr, al, pg, od)
        _sweep(h, n_cap, sample)  # warm: workspace sized and first-touch done
        ms, arcs = [], 0
        for _ in range(repeats):
            inf = _sweep(h, n_cap, sample)
            ms.append(inf.kernel_ms)
            arcs = inf.arcs_traversed
        t = float(np.median(ms))
        rate = arcs / (t * 1e-3) / 1e9 if t > 0 else 0.0
        row = {"groups": gr, "alpha": al, "pull_grid": pg, "order": od, "kernel_ms": round(t, 3),
               "gteps": round(rate, 2)}
        trials.append(row)
        if log:
            log(f"[autotune] {row}")
        if best is None or rate > best["gteps"]:
            best = row
    if best is None:
        raise ValueError("no candidate fits")
    _set(h, best["groups"], best["alpha"], best["pull_grid"], best["order"])
    return {"best": best, "trials": trials}


def batch_sample(offsets: np.ndarray, per_graph: int = 2048, graphs: int = 3) -> np.ndarray:
    """Sources for tuning a batched resource: the first per_graph sources of
    the `graphs` largest graphs (the large graphs carry most of the work,
    which grows with n^2)."""
    sizes = np.diff(offsets)
    pick = np.argsort(-sizes, kind="stable")[:graphs]
    out = [np.arange(offsets[i], offsets[i] + min(per_graph, sizes[i]), dtype=np.uint32) for i in sorted(pick)]
    return np.concatenate(out) if out else np.zeros(0, np.uint32)
