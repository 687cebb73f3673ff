"""ctypes front-end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, always as the checker or the timed CPU
baseline, never as the product path.

The oracle restates priv/zog/src/centrality.zig:694-798 (unweighted Brandes,
ArrayGraph branch) and :871-965 (PageRank, ArrayGraph branch) over a
restatement of priv/zog/src/models/array_graph.zig (linked-list successors in
reverse insertion order, tombstones).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SIGMA_U64 = 0
SIGMA_SCALED = 1
SIGMA_EXP = 2
NONE_U32 = 0xFFFFFFFF

_lib = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.zo_graph_create.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.zo_graph_create.restype = ctypes.c_int
        L.zo_graph_destroy.argtypes = [P]
        L.zo_graph_destroy.restype = None
        L.zo_node_count.argtypes = [P]
        L.zo_node_count.restype = ctypes.c_uint32
        L.zo_successors.argtypes = [P, ctypes.c_uint32, P, ctypes.c_uint64]
        L.zo_successors.restype = ctypes.c_uint64
        L.zo_bc_unweighted.argtypes = [P, P, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, P, P]
        L.zo_bc_unweighted.restype = ctypes.c_int
        L.zo_bfs_paths.argtypes = [P, ctypes.c_uint32, P, P, P, P, P, P, P]
        L.zo_bfs_paths.restype = ctypes.c_int
        L.zo_bfs_paths_exp.argtypes = [P, ctypes.c_uint32, P, P, P]
        L.zo_bfs_paths_exp.restype = ctypes.c_int
        L.zo_pagerank.argtypes = [P, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, P, P, P]
        L.zo_pagerank_mt.argtypes = [P, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int, P, P, P]
        L.zo_pagerank_mt.restype = ctypes.c_int
        L.zo_core_numbers.argtypes = [P, P]
        L.zo_core_numbers.restype = ctypes.c_int
        L.zo_graph_set_weights.argtypes = [P, P, ctypes.c_uint64]
        L.zo_graph_set_weights.restype = ctypes.c_int
        L.zo_bc_f64.argtypes = [P, P, ctypes.c_uint64, P]
        L.zo_eigenvector.argtypes = [P, ctypes.c_uint64, ctypes.c_double, P, P, P]
        L.zo_katz_alpha.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_uint64, ctypes.c_double, P, P, P]
        L.zo_bc_f64.restype = ctypes.c_int
        L.zo_pagerank.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


class Graph:
    """ArrayGraph(void, f64) restated; built like resource_graph.ex:120-129."""

    def __init__(self, n_cap, src, dst, live_mask=None):
        self.n_cap = int(n_cap)
        self.src = np.ascontiguousarray(src, dtype=np.uint32)
        self.dst = np.ascontiguousarray(dst, dtype=np.uint32)
        self.live = None if live_mask is None else np.ascontiguousarray(live_mask, dtype=np.uint8)
        h = ctypes.c_void_p()
        rc = lib().zo_graph_create(self.n_cap, _ptr(self.live), _ptr(self.src), _ptr(self.dst),
                                   len(self.src), ctypes.byref(h))
        if rc != 0:
            raise OracleError(f"zo_graph_create failed: {rc}")
        self._h = h

    def close(self):
        if self._h:
            lib().zo_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def node_count(self) -> int:
        return int(lib().zo_node_count(self._h))

    def successors(self, v: int) -> np.ndarray:
        k = int(lib().zo_successors(self._h, v, None, 0))
        out = np.empty(k, dtype=np.uint32)
        lib().zo_successors(self._h, v, _ptr(out), k)
        return out

    def betweenness(self, sources=None, mode=SIGMA_U64, threads=1, return_arcs=False):
        """Unnormalised directed-sum BC, f64[n_cap]; returns (bc, overflowed[, arcs])."""
        out = np.zeros(self.n_cap, dtype=np.float64)
        src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
        nsrc = 0 if src is None else len(src)
        arcs = None
        if return_arcs:
            arcs = np.zeros(nsrc if src is not None else self.node_count(), dtype=np.uint64)
        rc = lib().zo_bc_unweighted(self._h, _ptr(src), nsrc, mode, threads, _ptr(out), _ptr(arcs))
        if rc < 0:
            raise OracleError(f"zo_bc_unweighted failed: {rc}")
        if return_arcs:
            return out, rc == 1, arcs
        return out, rc == 1

    def bfs_paths(self, s: int):
        """(depth u32[V] (NONE_U32 = unreached), sigma u64[V], sigma_scaled f64[V],
        shifts i32[levels], arcs_scanned, overflowed)."""
        V = self.n_cap
        depth = np.empty(V, dtype=np.uint32)
        su = np.empty(V, dtype=np.uint64)
        sf = np.empty(V, dtype=np.float64)
        shift = np.zeros(V + 1, dtype=np.int32)
        arcs = ctypes.c_uint64()
        levels = ctypes.c_uint32()
        ovf = ctypes.c_int()
        rc = lib().zo_bfs_paths(self._h, s, _ptr(depth), _ptr(su), _ptr(sf), _ptr(shift),
                                ctypes.byref(arcs), ctypes.byref(levels), ctypes.byref(ovf))
        if rc != 0:
            raise OracleError(f"zo_bfs_paths failed: {rc}")
        return depth, su, sf, shift[: levels.value].copy(), int(arcs.value), bool(ovf.value)

    def bfs_paths_exp(self, s: int):
        """(depth u32[V], count mantissa f64[V] in [1, 2), exponent i32[V]) of
        the per-vertex-exponent mode: sigma = mantissa * 2^exponent."""
        V = self.n_cap
        depth = np.empty(V, dtype=np.uint32)
        mant = np.empty(V, dtype=np.float64)
        ex = np.empty(V, dtype=np.int32)
        rc = lib().zo_bfs_paths_exp(self._h, s, _ptr(depth), _ptr(mant), _ptr(ex))
        if rc != 0:
            raise OracleError(f"zo_bfs_paths_exp failed: {rc}")
        return depth, mant, ex

    def pagerank(self, damping=0.85, max_iterations=100, tolerance=1e-4, threads=1):
        """FP64 PageRank, f64[n_cap]; returns (ranks, iterations, last_l1).
        threads > 1 splits each iteration's rows over threads (same bits)."""
        out = np.zeros(self.n_cap, dtype=np.float64)
        it = ctypes.c_uint64()
        l1 = ctypes.c_double()
        rc = lib().zo_pagerank_mt(self._h, damping, max_iterations, tolerance, int(threads), _ptr(out),
                                  ctypes.byref(it), ctypes.byref(l1))
        if rc != 0:
            raise OracleError(f"zo_pagerank failed: {rc}")
        return out, int(it.value), float(l1.value)

    def core_numbers(self) -> np.ndarray:
        """u32[n_cap] core numbers of the undirected simple view (0 for non-live)."""
        out = np.zeros(self.n_cap, dtype=np.uint32)
        rc = lib().zo_core_numbers(self._h, _ptr(out))
        if rc != 0:
            raise OracleError(f"zo_core_numbers failed: {rc}")
        return out

    # outcomes of eigenvector(): how the reference's loop ended
    EIG_CONVERGED, EIG_OSCILLATING, EIG_ZERO, EIG_CAP, EIG_TRIVIAL = range(5)

    def eigenvector(self, max_iterations=100, tolerance=1e-4):
        """Eigenvector centrality f64[n_cap]; returns (scores, iterations, outcome)."""
        out = np.zeros(self.n_cap, dtype=np.float64)
        it = ctypes.c_uint64()
        oc = ctypes.c_int()
        rc = lib().zo_eigenvector(self._h, max_iterations, tolerance, _ptr(out), ctypes.byref(it), ctypes.byref(oc))
        if rc != 0:
            raise OracleError(f"zo_eigenvector failed: {rc}")
        return out, int(it.value), int(oc.value)

    def katz(self, alpha=0.1, beta=1.0, max_iterations=100, tolerance=1e-4):
        """Katz centrality f64[n_cap]; returns (scores, iterations, last_l1)."""
        return self._katz_alpha(alpha, beta, True, beta, max_iterations, tolerance)

    def alpha_centrality(self, alpha=0.5, initial=1.0, max_iterations=100, tolerance=1e-4):
        """Alpha centrality f64[n_cap]; returns (scores, iterations, last_l1)."""
        return self._katz_alpha(alpha, 0.0, False, initial, max_iterations, tolerance)

    def _katz_alpha(self, alpha, beta, has_beta, start, max_iterations, tolerance):
        out = np.zeros(self.n_cap, dtype=np.float64)
        it = ctypes.c_uint64()
        l1 = ctypes.c_double()
        rc = lib().zo_katz_alpha(self._h, alpha, beta, int(has_beta), start, max_iterations, tolerance, _ptr(out),
                                 ctypes.byref(it), ctypes.byref(l1))
        if rc != 0:
            raise OracleError(f"zo_katz_alpha failed: {rc}")
        return out, int(it.value), float(l1.value)

    def set_weights(self, weights):
        """Edge data f64[m] in arc order."""
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        rc = lib().zo_graph_set_weights(self._h, _ptr(self.weights), len(self.weights))
        if rc != 0:
            raise OracleError(f"zo_graph_set_weights failed: {rc}")

    def betweenness_f64(self, sources=None):
        """Unnormalised weighted BC f64[n_cap] (betweennessF64)."""
        out = np.zeros(self.n_cap, dtype=np.float64)
        src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
        rc = lib().zo_bc_f64(self._h, _ptr(src), 0 if src is None else len(src), _ptr(out))
        if rc != 0:
            raise OracleError(f"zo_bc_f64 failed: {rc}")
        return out
