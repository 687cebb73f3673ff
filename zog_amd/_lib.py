"""ctypes binding of libzog_hip.so (the C-ABI declared in include/zog_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C zog_amd/csrc``).  There is no fallback: if the shared library is
missing, importing the product path raises immediately.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzog_hip.so")

ZG_OK = 0
ZG_ERR_INVALID_ARG = -1
ZG_ERR_NODE_NOT_FOUND = -2
ZG_ERR_OOM = -3
ZG_ERR_GPU_OOM = -4
ZG_ERR_HIP = -5
ZG_ERR_NO_DEVICE = -6
ZG_ERR_TOO_LARGE = -7
ZG_ERR_INVALID_HANDLE = -8
ZG_ERR_IO = -9
ZG_ERR_SCRATCH = -10

ZG_PR_FP32 = 32
ZG_PR_FP64 = 64
ZG_PR_MIXED = 96
ZG_PR_SKIP_RESIDUAL = 0x100
ZG_TUNE_PUSH_ALPHA = 1
ZG_TUNE_ENGINE = 2
ZG_ENGINE_AUTO = 0
ZG_ENGINE_LANES = 1
ZG_ENGINE_QUEUE = 2
ZG_TUNE_PULL_GRID = 3
ZG_TUNE_GRAPH = 4
ZG_TUNE_PR_CACHE = 5
ZG_TUNE_PR_GRID = 6
ZG_TUNE_PR_BLOCK = 12
ZG_FLOOR_LDS_COPY = 0x100
ZG_TUNE_WEIGHTED_BATCH = 7
ZG_TUNE_WEIGHTED_DELTA = 8
ZG_TUNE_WEIGHTED_GRID = 9
ZG_TUNE_COUNT_TREE_ARCS = 10
ZG_TUNE_QUEUE_ORDER = 11
ZG_TUNE_LANE_ORDER = 13
# work tables of zg_graph_table
(ZG_TABLE_ICH, ZG_TABLE_ICH_FIRST, ZG_TABLE_OCH, ZG_TABLE_OCH_FIRST, ZG_TABLE_OHUB, ZG_TABLE_PR_ROWS,
 ZG_TABLE_PR_DESC, ZG_TABLE_PR_LARGE, ZG_TABLE_PR_SEG, ZG_TABLE_PR_COL, ZG_TABLE_PR_BINS) = range(11)
ZG_EIG_CONVERGED, ZG_EIG_OSCILLATING, ZG_EIG_ZERO, ZG_EIG_CAP, ZG_EIG_TRIVIAL = range(5)

# Every symbol include/zog_hip.h declares.
EXPORTED = (
    "zg_graph_create",
    "zg_graph_create_weighted",
    "zg_graph_create_batch",
    "zg_set_build_threshold",
    "zg_graph_csr",
    "zg_graph_table",
    "zg_edgelist_read",
    "zg_edgelist_arcs",
    "zg_edgelist_labels",
    "zg_edgelist_free",
    "zg_graph_components",
    "zg_graph_destroy",
    "zg_graph_info",
    "zg_bc_unweighted",
    "zg_bc_unweighted_device",
    "zg_bc_f64",
    "zg_pagerank",
    "zg_pagerank_device",
    "zg_pagerank_device_noalloc",
    "zg_eigenvector",
    "zg_katz",
    "zg_alpha_centrality",
    "zg_gather_floor",
    "zg_bc_paths",
    "zg_set_batch_groups",
    "zg_set_source_order",
    "zg_set_tuning",
    "zg_set_profiling",
    "zg_get_profile",
    "zg_strerror",
    "zg_bc_weighted",
    "zg_fold_map",
    "zg_bc_folded",
    "zg_core_numbers",
    "zg_node_degrees",
    "zg_node_features",
    "zg_mlp_create",
    "zg_mlp_destroy",
    "zg_mlp_get_params",
    "zg_mlp_set_params",
    "zg_mlp_forward",
    "zg_mlp_gradient",
    "zg_mlp_adam",
)

ZG_NFEAT = 8
ZG_MLP_HIDDEN = 128
ZG_MLP_PARAMS = 17793
ZG_MLP_MAX_BATCH = 65536

PROF_SLOTS = ("bc_forward", "bc_finalize", "bc_backward", "bc_hub", "bc_setup", "pr_iter", "bc_push", "bc_dag")


class ZgBcInfo(ctypes.Structure):
    _fields_ = [
        ("n_sources", ctypes.c_uint64),
        ("arcs_traversed", ctypes.c_uint64),
        ("nodes_reached", ctypes.c_uint64),
        ("max_levels", ctypes.c_uint32),
        ("batches", ctypes.c_uint32),
        ("push_levels", ctypes.c_uint32),
        ("pull_levels", ctypes.c_uint32),
        ("kernel_ms", ctypes.c_double),
        ("arcs_relaxed", ctypes.c_uint64),
        ("row_visits", ctypes.c_uint64),
        ("row_arcs", ctypes.c_uint64),
        ("tree_arcs", ctypes.c_uint64),
        ("group_slots", ctypes.c_uint64),
        ("span_balance", ctypes.c_double),
    ]


class ZgPrInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint64),
        ("last_l1", ctypes.c_double),
        ("residual_l1", ctypes.c_double),
        ("error_bound_l1", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("outcome", ctypes.c_int32),
        ("fp64_iterations", ctypes.c_uint64),
    ]


class ZgProfile(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 8), ("launches", ctypes.c_uint64 * 8)]


class ZogError(RuntimeError):
    """A non-zero status from the C-ABI (the NIF would raise it to Elixir)."""

    def __init__(self, code: int, where: str = ""):
        self.code = code
        msg = strerror(code)
        super().__init__(f"{where}: {msg} ({code})" if where else f"{msg} ({code})")


class NodeNotFoundError(ZogError):
    pass


_lib = None


def _init_torch_first():
    """PyTorch-ROCm ships its own HIP runtime next to the one this library
    links (/opt/rocm).  Both sit on the same HSA runtime, so device pointers
    pass between them, but torch's runtime finds no GPU if it initialises
    after ours; when torch is present it is initialised first.  Work does
    not order across the two runtimes' queues: the Python wrappers
    synchronise at every hand-over, and only stream 0 (the null stream) is
    passed from torch to this library."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    """Load libzog_hip.so; raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    _init_torch_first()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.zg_graph_create.argtypes = [ctypes.c_uint32, P, P, P, ctypes.c_uint64, ctypes.POINTER(P)]
    L.zg_graph_create.restype = ctypes.c_int
    L.zg_graph_create_weighted.argtypes = [ctypes.c_uint32, P, P, P, P, ctypes.c_uint64, ctypes.POINTER(P)]
    L.zg_graph_create_weighted.restype = ctypes.c_int
    L.zg_set_build_threshold.argtypes = [ctypes.c_uint64]
    L.zg_set_build_threshold.restype = ctypes.c_int
    L.zg_edgelist_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    L.zg_edgelist_read.restype = ctypes.c_int
    L.zg_edgelist_arcs.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P),
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)]
    L.zg_edgelist_arcs.restype = ctypes.c_int
    L.zg_edgelist_labels.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(ctypes.c_uint32)]
    L.zg_edgelist_labels.restype = ctypes.c_int
    L.zg_edgelist_free.argtypes = [P]
    L.zg_edgelist_free.restype = None
    L.zg_graph_csr.argtypes = [P, P, P, P, P, P]
    L.zg_graph_csr.restype = ctypes.c_int
    L.zg_graph_table.argtypes = [P, ctypes.c_int, P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.zg_graph_table.restype = ctypes.c_int
    L.zg_graph_create_batch.argtypes = [ctypes.c_uint32, P, P, P, P, ctypes.POINTER(P)]
    L.zg_graph_create_batch.restype = ctypes.c_int
    L.zg_graph_components.argtypes = [P, P]
    L.zg_graph_components.restype = ctypes.c_int
    L.zg_graph_destroy.argtypes = [P]
    L.zg_graph_destroy.restype = None
    L.zg_graph_info.argtypes = [P, P, P, P]
    L.zg_graph_info.restype = ctypes.c_int
    L.zg_bc_unweighted.argtypes = [P, P, ctypes.c_uint64, P, ctypes.POINTER(ZgBcInfo)]
    L.zg_bc_unweighted.restype = ctypes.c_int
    L.zg_bc_unweighted_device.argtypes = [P, P, ctypes.c_uint64, P, P, ctypes.POINTER(ZgBcInfo)]
    L.zg_bc_unweighted_device.restype = ctypes.c_int
    L.zg_bc_f64.argtypes = [P, P, ctypes.c_uint64, P, ctypes.POINTER(ZgBcInfo)]
    L.zg_bc_f64.restype = ctypes.c_int
    L.zg_pagerank.argtypes = [P, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int, P,
                              ctypes.POINTER(ZgPrInfo)]
    L.zg_pagerank.restype = ctypes.c_int
    L.zg_pagerank_device.argtypes = [P, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                     P, P, ctypes.POINTER(ZgPrInfo)]
    L.zg_pagerank_device.restype = ctypes.c_int
    L.zg_pagerank_device_noalloc.argtypes = [P, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                             P, P, ctypes.POINTER(ZgPrInfo)]
    L.zg_pagerank_device_noalloc.restype = ctypes.c_int
    L.zg_eigenvector.argtypes = [P, ctypes.c_uint64, ctypes.c_double, P, ctypes.POINTER(ZgPrInfo)]
    L.zg_eigenvector.restype = ctypes.c_int
    L.zg_katz.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, P,
                          ctypes.POINTER(ZgPrInfo)]
    L.zg_katz.restype = ctypes.c_int
    L.zg_alpha_centrality.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, P,
                                      ctypes.POINTER(ZgPrInfo)]
    L.zg_alpha_centrality.restype = ctypes.c_int
    L.zg_gather_floor.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.zg_gather_floor.restype = ctypes.c_int
    L.zg_bc_paths.argtypes = [P, P, ctypes.c_uint64, P, P, P, ctypes.c_uint32, P]
    L.zg_bc_paths.restype = ctypes.c_int
    L.zg_set_batch_groups.argtypes = [P, ctypes.c_int]
    L.zg_set_batch_groups.restype = ctypes.c_int
    L.zg_set_tuning.argtypes = [P, ctypes.c_int, ctypes.c_double]
    L.zg_set_tuning.restype = ctypes.c_int
    L.zg_bc_weighted.argtypes = [P, P, P, ctypes.c_uint64, P, P, ctypes.POINTER(ZgBcInfo)]
    L.zg_bc_weighted.restype = ctypes.c_int
    L.zg_fold_map.argtypes = [P, P, P, P, P]
    L.zg_fold_map.restype = ctypes.c_int
    L.zg_bc_folded.argtypes = [P, P, P, ctypes.c_uint64, P, P, ctypes.POINTER(ZgBcInfo)]
    L.zg_bc_folded.restype = ctypes.c_int
    L.zg_node_degrees.argtypes = [P, P]
    L.zg_node_degrees.restype = ctypes.c_int
    L.zg_core_numbers.argtypes = [P, P]
    L.zg_core_numbers.restype = ctypes.c_int
    L.zg_node_features.argtypes = [P, P, P]
    L.zg_node_features.restype = ctypes.c_int
    L.zg_mlp_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(P)]
    L.zg_mlp_create.restype = ctypes.c_int
    L.zg_mlp_destroy.argtypes = [P]
    L.zg_mlp_destroy.restype = None
    L.zg_mlp_get_params.argtypes = [P, P]
    L.zg_mlp_get_params.restype = ctypes.c_int
    L.zg_mlp_set_params.argtypes = [P, P]
    L.zg_mlp_set_params.restype = ctypes.c_int
    L.zg_mlp_forward.argtypes = [P, P, ctypes.c_uint64, P, P]
    L.zg_mlp_forward.restype = ctypes.c_int
    L.zg_mlp_gradient.argtypes = [P, P, P, ctypes.c_uint64, P, P, P]
    L.zg_mlp_gradient.restype = ctypes.c_int
    L.zg_mlp_adam.argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, P]
    L.zg_mlp_adam.restype = ctypes.c_int
    L.zg_set_source_order.argtypes = [P, ctypes.c_int]
    L.zg_set_source_order.restype = ctypes.c_int
    L.zg_set_profiling.argtypes = [P, ctypes.c_int]
    L.zg_set_profiling.restype = ctypes.c_int
    L.zg_get_profile.argtypes = [P, ctypes.POINTER(ZgProfile)]
    L.zg_get_profile.restype = ctypes.c_int
    L.zg_strerror.argtypes = [ctypes.c_int]
    L.zg_strerror.restype = ctypes.c_char_p
    _lib = L
    return L


def strerror(code: int) -> str:
    try:
        return lib().zg_strerror(code).decode()
    except OSError:
        return f"status {code}"


def check(code: int, where: str = "") -> None:
    if code == ZG_OK:
        return
    if code == ZG_ERR_NODE_NOT_FOUND:
        raise NodeNotFoundError(code, where)
    raise ZogError(code, where)
