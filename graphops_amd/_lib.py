"""ctypes binding of libgraphops_amd.so (the C ABI in include/graphops_amd.h).

The shared library is built in-tree by graphops_amd/csrc/Makefile.  There is
no fallback: if the library is missing, or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraphops_amd.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

GOPS_OK = 0
GOPS_INVALID_PARAMETER = 1
GOPS_INDEX_OUT_OF_BOUNDS = 2
GOPS_DIMENSION_MISMATCH = 3
GOPS_DEVICE_ERROR = 4
GOPS_OUT_OF_MEMORY = 5
GOPS_NOT_SUPPORTED = 6

WEIGHT_NONE, WEIGHT_F64, WEIGHT_F32, WEIGHT_BF16 = 0, 1, 2, 3
OPT_PRECISION, OPT_ANCHOR_PERIOD, OPT_NORMALIZE, OPT_TIMING, OPT_PPR_TILE = 1, 2, 3, 4, 5
OPT_ANCHOR_DROP = 6
OPT_F32_BELOW, OPT_PPR_F32_BELOW, OPT_PPR_ANCHOR_PERIOD = 7, 8, 9
OPT_GRAPH_BUILD = 10
OPT_TILE_F32_BELOW = 11
OPT_DELTA_SWEEP = 12
OPT_LAUNCH = 13
OPT_ANCHOR_SWEEP = 14
OPT_BLOCKED_BIN_RUNS = 15
OPT_CENTRALITY_SWEEP = 16
CENTRALITY_SWEEP_AUTO, CENTRALITY_SWEEP_PULL, CENTRALITY_SWEEP_BLOCKED = 0, 1, 2
ANCHOR_SWEEP_AUTO = 0
ANCHOR_SWEEP_PULL = 1
LAUNCH_AUTO, LAUNCH_QUEUED, LAUNCH_GRAPH = 0, 1, 2
DELTA_SWEEP_AUTO, DELTA_SWEEP_PULL, DELTA_SWEEP_BLOCKED = 0, 1, 2
BUILD_DEVICE, BUILD_HOST = 0, 1
PRECISION_MIXED, PRECISION_FP64 = 0, 1
GEN_UNIT, GEN_UNIFORM, GEN_LOGNORMAL = 0, 1, 2
ROUND_F64, ROUND_F32, ROUND_BF16 = 0, 1, 2


class PageRankConfigC(ctypes.Structure):
    _fields_ = [("damping", ctypes.c_double), ("max_iterations", ctypes.c_uint64),
                ("tolerance", ctypes.c_double)]


class KatzConfigC(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("max_iterations", ctypes.c_uint64), ("tolerance", ctypes.c_double)]


class RunInfoC(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("diff_l1", ctypes.c_double),
                ("converged", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RunStatsC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("nnz", ctypes.c_uint64),
                ("sweeps_fp32", ctypes.c_uint64), ("sweeps_fp64", ctypes.c_uint64),
                ("device_ms", ctypes.c_double), ("fp32_sweep_ms", ctypes.c_double),
                ("fp64_sweep_ms", ctypes.c_double), ("other_ms", ctypes.c_double),
                ("sweeps_blocked", ctypes.c_uint64), ("blocked_fallback", ctypes.c_int32),
                ("graph_replay", ctypes.c_int32), ("blocked_bins", ctypes.c_int32),
                ("blocked_split_bins", ctypes.c_int32)]


class RowPlanC(ctypes.Structure):
    _fields_ = [("row_min", ctypes.c_int32), ("hub_chunk", ctypes.c_int32),
                ("row_iters", ctypes.c_int32), ("candidates", ctypes.c_int32),
                ("default_us", ctypes.c_double), ("tuned_us", ctypes.c_double)]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("max_slots", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_double),
                ("tile", ctypes.c_uint32), ("topk", ctypes.c_uint32)]


class ServerResultC(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("info", RunInfoC),
                ("submit_ms", ctypes.c_double), ("finish_ms", ctypes.c_double)]


# (name, restype, argtypes) of every symbol include/graphops_amd.h declares
SIGNATURES = [
    ("gops_context_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("gops_context_destroy", None, [_vp]),
    ("gops_last_error", ctypes.c_char_p, [_vp]),
    ("gops_context_set_option", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    ("gops_last_run_stats", ctypes.c_int, [_vp, ctypes.POINTER(RunStatsC)]),
    ("gops_graph_create", ctypes.c_int,
     [_vp, ctypes.c_uint64, _u64p, _u32p, _f64p, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("gops_graph_from_dense", ctypes.c_int,
     [_vp, ctypes.c_uint64, _f64p, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("gops_graph_destroy", None, [_vp]),
    ("gops_graph_info", ctypes.c_int, [_vp, _u64p, _u64p]),
    ("gops_graph_export", ctypes.c_int, [_vp, _i32p, _i32p, _vp, _u32p, _f64p, _f64p]),
    ("gops_pagerank_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(PageRankConfigC), _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_pagerank_checked_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(PageRankConfigC), _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_pagerank_weighted_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(PageRankConfigC), _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_pagerank_weighted_checked_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(PageRankConfigC), _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_personalized_pagerank_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(PageRankConfigC), _f64p, ctypes.c_uint64, _f64p,
      ctypes.POINTER(RunInfoC)]),
    ("gops_personalized_pagerank_checked_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(PageRankConfigC), _f64p, ctypes.c_uint64, _f64p,
      ctypes.POINTER(RunInfoC)]),
    ("gops_gather_probe", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _f64p]),
    ("gops_top_k", ctypes.c_int,
     [_vp, _f64p, ctypes.c_uint64, ctypes.c_uint32, _u32p, _f64p, _u32p]),
    ("gops_result_top_k", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _f64p, _u32p]),
    ("gops_result_top_k_noalloc", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _f64p, _u32p]),
    ("gops_katz_centrality_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(KatzConfigC), _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_katz_centrality_checked_run", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(KatzConfigC), _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_eigenvector_centrality_run", ctypes.c_int,
     [_vp, _vp, ctypes.c_uint64, ctypes.c_double, _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_hits", ctypes.c_int,
     [_vp, _vp, ctypes.c_uint64, ctypes.c_double, _f64p, _f64p, ctypes.POINTER(RunInfoC)]),
    ("gops_ppr_batch_run", ctypes.c_int,
     [_vp, _vp, ctypes.c_uint32, ctypes.POINTER(PageRankConfigC), _u64p, _u32p, _f64p, _f64p,
      ctypes.POINTER(RunInfoC)]),
    ("gops_ppr_tile_info", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _f64p]),
    ("gops_graph_tune_rows", ctypes.c_int, [_vp, _vp, ctypes.POINTER(RowPlanC)]),
    ("gops_graph_set_rows", ctypes.c_int, [_vp, _vp, ctypes.POINTER(RowPlanC)]),
    ("gops_server_create", ctypes.c_int,
     [_vp, _vp, ctypes.POINTER(ServerConfigC), ctypes.POINTER(_vp)]),
    ("gops_server_destroy", None, [_vp]),
    ("gops_server_submit", ctypes.c_int,
     [_vp, ctypes.POINTER(PageRankConfigC), _u32p, ctypes.c_uint64, _f64p, _u64p]),
    ("gops_server_step", ctypes.c_int, [_vp, ctypes.c_uint32, _u32p]),
    ("gops_server_poll", ctypes.c_int,
     [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ServerResultC), _f64p, _u32p, _f64p]),
    ("gops_server_stats", ctypes.c_int, [_vp, _u64p, _u64p, _u64p, _u32p]),
    ("gops_gen_rmat", ctypes.c_int,
     [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
      ctypes.POINTER(_vp)]),
    ("gops_gen_erdos_renyi", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("gops_gen_seed_sets", ctypes.c_int,
     [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64,
      _u64p, _u32p, ctypes.c_uint64]),
    ("gops_edges_view", ctypes.c_int,
     [_vp, _u64p, _u64p, ctypes.POINTER(_u64p), ctypes.POINTER(_u32p), ctypes.POINTER(_f64p)]),
    ("gops_edges_destroy", None, [_vp]),
]

_lib = None


def lib():
    """Load libgraphops_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"graphops_amd: HIP library not built ({LIB_PATH}); run "
                "`make -C graphops_amd/csrc` or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a, t):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)
