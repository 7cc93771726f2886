"""ctypes front end of the f64 parity oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's cpu_baseline leg, always as the checker (or the timed CPU
baseline), never as part of the product path.  The product package
graphops_amd never imports this module.

Each function restates the reference function named in its docstring
(crate graphops v0.4.1); see oracle.c for the line-level citations.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)


def build() -> None:
    """Compile oracle.c with its Makefile (gcc, -ffp-contract=off)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_validate_config.argtypes = [ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                          ctypes.c_char_p, ctypes.c_size_t]
        L.orc_check_weights.argtypes = [ctypes.c_uint64, _f64p, ctypes.c_char_p, ctypes.c_size_t]
        L.orc_check_personalization.argtypes = [ctypes.c_uint64, _f64p, ctypes.c_uint64,
                                                ctypes.c_char_p, ctypes.c_size_t]
        L.orc_dense_neighbors.argtypes = [ctypes.c_uint64, _f64p, _u64p, _u32p, _f64p]
        L.orc_dense_neighbors.restype = ctypes.c_uint64
        run_tail = [ctypes.c_double, ctypes.c_uint64, ctypes.c_double, _f64p, _u64p, _f64p, _i32p]
        L.orc_pagerank_run.argtypes = [ctypes.c_uint64, _u64p, _u32p] + run_tail
        L.orc_pagerank_weighted_run.argtypes = [ctypes.c_uint64, _u64p, _u32p, _f64p] + run_tail
        L.orc_ppr_run.argtypes = [ctypes.c_uint64, _u64p, _u32p, _f64p] + run_tail
        L.orc_out_wsum.argtypes = [ctypes.c_uint64, _u64p, _f64p, _f64p]
        L.orc_pull_run.argtypes = [ctypes.c_int, ctypes.c_uint64, _u64p, _u32p, _f64p, _f64p, _f64p,
                                   ctypes.c_double, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                   _f64p, _u64p, _f64p, _i32p]
        L.orc_pull_run_ld.argtypes = [ctypes.c_int, ctypes.c_uint64, _u64p, _u32p, _f64p, _f64p,
                                      ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_int, _f64p, _u64p, _f64p, _i32p]
        L.orc_push_sweep.argtypes = [ctypes.c_int, ctypes.c_uint64, _u64p, _u32p, _f64p, _f64p,
                                     ctypes.c_double, _f64p, _f64p]
        L.orc_push_sweep.restype = ctypes.c_double
        L.orc_push_find_edge.argtypes = [ctypes.c_uint64, _u32p, _u64p, _u32p, _f64p, _f64p,
                                         ctypes.c_double, _f64p, _f64p]
        L.orc_transpose.argtypes = [ctypes.c_uint64, _u64p, _u32p, _f64p, _u64p, _u32p, _f64p]
        L.orc_top_k.argtypes = [ctypes.c_uint64, _f64p, ctypes.c_uint64, _u64p, _f64p]
        L.orc_top_k.restype = ctypes.c_uint64
        L.orc_validate_katz.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                        ctypes.c_double, ctypes.c_char_p, ctypes.c_size_t]
        L.orc_katz_run.argtypes = [ctypes.c_uint64, _u64p, _u32p, ctypes.c_double] + run_tail
        L.orc_eigenvector_run.argtypes = [ctypes.c_uint64, _u64p, _u32p, ctypes.c_uint64,
                                          ctypes.c_double, _f64p, _u64p, _f64p, _i32p]
        L.orc_hits.argtypes = [ctypes.c_uint64, _u64p, _u32p, ctypes.c_uint64, ctypes.c_double,
                               _f64p, _f64p, _u64p, _f64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)


@dataclass
class Run:
    """PageRankRun, pagerank.rs:14-29."""

    scores: np.ndarray
    iterations: int
    diff_l1: float
    converged: bool


class OracleError(ValueError):
    """Error::InvalidParameter, lib.rs:117-118 ("invalid parameter: {0}")."""


def _msg_call(fn, *args):
    buf = ctypes.create_string_buffer(256)
    rc = fn(*args, buf, 256)
    if rc != 0:
        raise OracleError("invalid parameter: " + buf.value.decode())


def validate_config(damping: float, max_iterations: int, tolerance: float) -> None:
    """PageRankConfig::validate, pagerank.rs:63-85."""
    _msg_call(lib().orc_validate_config, float(damping), int(max_iterations), float(tolerance))


def check_weights(ew: np.ndarray) -> None:
    """Weight checks of pagerank_weighted_checked, pagerank.rs:300-314."""
    ew = np.ascontiguousarray(ew, dtype=np.float64)
    _msg_call(lib().orc_check_weights, ew.size, _p(ew, _f64p))


def check_personalization(n: int, p: np.ndarray) -> None:
    """Personalization checks of personalized_pagerank_checked, ppr.rs:17-42."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    _msg_call(lib().orc_check_personalization, int(n), _p(p, _f64p), p.size)


def dense_neighbors(mat: np.ndarray):
    """AdjacencyMatrix::neighbors / edge_weight, graph.rs:122-142."""
    mat = np.ascontiguousarray(mat, dtype=np.float64)
    n = mat.shape[0]
    off = np.zeros(n + 1, dtype=np.uint64)
    m = lib().orc_dense_neighbors(n, _p(mat, _f64p), _p(off, _u64p), None, None)
    nb = np.zeros(max(m, 1), dtype=np.uint32)
    ew = np.zeros(max(m, 1), dtype=np.float64)
    lib().orc_dense_neighbors(n, _p(mat, _f64p), _p(off, _u64p), _p(nb, _u32p), _p(ew, _f64p))
    return off, nb[:m], ew[:m]


def _tail(n):
    scores = np.zeros(max(n, 1), dtype=np.float64)
    it = np.zeros(1, dtype=np.uint64)
    diff = np.zeros(1, dtype=np.float64)
    conv = np.zeros(1, dtype=np.int32)
    return scores, it, diff, conv


def _run(scores, it, diff, conv, n):
    return Run(scores[:n].copy(), int(it[0]), float(diff[0]), bool(conv[0]))


def pagerank_run(off, nb, damping=0.85, max_iterations=100, tolerance=1e-6) -> Run:
    """pagerank_run, pagerank.rs:122-187 (neighbour lists in adapter order)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    n = off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_pagerank_run(n, _p(off, _u64p), _p(nb, _u32p), damping, max_iterations, tolerance,
                           _p(s, _f64p), _p(it, _u64p), _p(d, _f64p), _p(c, _i32p))
    return _run(s, it, d, c, n)


def pagerank_weighted_run(off, nb, ew, damping=0.85, max_iterations=100, tolerance=1e-6) -> Run:
    """pagerank_weighted_run, pagerank.rs:210-289; ew[j] = edge_weight(u, nb[j])."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    ew = np.ascontiguousarray(ew, dtype=np.float64)
    n = off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_pagerank_weighted_run(n, _p(off, _u64p), _p(nb, _u32p), _p(ew, _f64p), damping,
                                    max_iterations, tolerance, _p(s, _f64p), _p(it, _u64p),
                                    _p(d, _f64p), _p(c, _i32p))
    return _run(s, it, d, c, n)


def personalized_pagerank_run(off, nb, p, damping=0.85, max_iterations=100,
                              tolerance=1e-6) -> Run:
    """personalized_pagerank_run, ppr.rs:90-156."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    p = np.ascontiguousarray(p, dtype=np.float64)
    n = off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_ppr_run(n, _p(off, _u64p), _p(nb, _u32p), _p(p, _f64p), damping, max_iterations,
                      tolerance, _p(s, _f64p), _p(it, _u64p), _p(d, _f64p), _p(c, _i32p))
    return _run(s, it, d, c, n)


def out_wsum(off, ew):
    """out_wsum, pagerank.rs:227-234."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    ew = np.ascontiguousarray(ew, dtype=np.float64)
    n = off.size - 1
    ws = np.zeros(max(n, 1), dtype=np.float64)
    lib().orc_out_wsum(n, _p(off, _u64p), _p(ew, _f64p), _p(ws, _f64p))
    return ws[:n]


def transpose(off, nb, ew=None):
    """In-edge lists sorted by ascending source (the pull order)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    n = off.size - 1
    m = int(off[-1])
    in_off = np.zeros(n + 1, dtype=np.uint64)
    in_src = np.zeros(max(m, 1), dtype=np.uint32)
    in_w = np.zeros(max(m, 1), dtype=np.float64) if ew is not None else None
    ewc = np.ascontiguousarray(ew, dtype=np.float64) if ew is not None else None
    lib().orc_transpose(n, _p(off, _u64p), _p(nb, _u32p), _p(ewc, _f64p), _p(in_off, _u64p),
                        _p(in_src, _u32p), _p(in_w, _f64p))
    return in_off, in_src[:m], (in_w[:m] if in_w is not None else None)


def pull_run(kind, in_off, in_src, scale, w=None, p=None, damping=0.85, max_iterations=100,
             tolerance=1e-6, threads=0) -> Run:
    """Pull restatement (OpenMP over rows) of pagerank_run (kind 0),
    pagerank_weighted_run (kind 1) or personalized_pagerank_run (kind 2);
    bit-identical to the push forms.  scale = out-degree (kinds 0, 2) or
    out_wsum (kind 1) per node, as f64."""
    in_off = np.ascontiguousarray(in_off, dtype=np.uint64)
    in_src = np.ascontiguousarray(in_src, dtype=np.uint32)
    scale = np.ascontiguousarray(scale, dtype=np.float64)
    wc = np.ascontiguousarray(w, dtype=np.float64) if w is not None else None
    pc = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
    n = in_off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_pull_run(kind, n, _p(in_off, _u64p), _p(in_src, _u32p), _p(wc, _f64p),
                       _p(scale, _f64p), _p(pc, _f64p), damping, max_iterations, tolerance,
                       threads, _p(s, _f64p), _p(it, _u64p), _p(d, _f64p), _p(c, _i32p))
    return _run(s, it, d, c, n)


def pull_run_ld(kind, in_off, in_src, w=None, p=None, damping=0.85, max_iterations=100,
                tolerance=1e-6, threads=0) -> Run:
    """The iteration of pagerank_run (kind 0), pagerank_weighted_run (kind 1)
    or personalized_pagerank_run (kind 2) carried in extended precision (long
    double): a stand-in for the exact iteration at sizes fractions cannot
    handle.  Out-degrees / out-weight sums are formed from the in-edges;
    w = max(edge_weight, 0) per in-edge (kind 1).  Scores and the residual
    come back rounded to f64."""
    in_off = np.ascontiguousarray(in_off, dtype=np.uint64)
    in_src = np.ascontiguousarray(in_src, dtype=np.uint32)
    wc = np.ascontiguousarray(w, dtype=np.float64) if w is not None else None
    pc = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
    n = in_off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_pull_run_ld(kind, n, _p(in_off, _u64p), _p(in_src, _u32p), _p(wc, _f64p),
                          _p(pc, _f64p), damping, max_iterations, tolerance, threads,
                          _p(s, _f64p), _p(it, _u64p), _p(d, _f64p), _p(c, _i32p))
    return _run(s, it, d, c, n)


def push_sweep(off, nb, cur, damping=0.85, ew=None, ws=None):
    """One sweep of the reference's push loop (pagerank.rs:142-173 or
    239-274) from `cur`; returns (new vector, L1 residual)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    cur = np.ascontiguousarray(cur, dtype=np.float64)
    n = off.size - 1
    out = np.zeros(n, dtype=np.float64)
    weighted = ew is not None
    ewc = np.ascontiguousarray(ew, dtype=np.float64) if weighted else None
    wsc = np.ascontiguousarray(ws, dtype=np.float64) if weighted else None
    r = lib().orc_push_sweep(int(weighted), n, _p(off, _u64p), _p(nb, _u32p), _p(ewc, _f64p),
                             _p(wsc, _f64p), damping, _p(cur, _f64p), _p(out, _f64p))
    return out, float(r)


def push_find_edge(us, off, nb, ew, ws, cur, damping=0.85, out=None):
    """The weighted sweep's scatter for the sources `us`, each entry's weight
    looked up by a find_edge scan (graph.rs:205-211); adds into out."""
    us = np.ascontiguousarray(us, dtype=np.uint32)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    ew = np.ascontiguousarray(ew, dtype=np.float64)
    ws = np.ascontiguousarray(ws, dtype=np.float64)
    cur = np.ascontiguousarray(cur, dtype=np.float64)
    n = off.size - 1
    if out is None:
        out = np.zeros(n, dtype=np.float64)
    lib().orc_push_find_edge(us.size, _p(us, _u32p), _p(off, _u64p), _p(nb, _u32p), _p(ew, _f64p),
                             _p(ws, _f64p), damping, _p(cur, _f64p), _p(out, _f64p))
    return out


def validate_katz(alpha: float, beta: float, max_iterations: int, tolerance: float) -> None:
    """KatzConfig::validate, katz.rs:57-84."""
    _msg_call(lib().orc_validate_katz, float(alpha), float(beta), int(max_iterations),
              float(tolerance))


def katz_run(off, nb, alpha=0.1, beta=1.0, max_iterations=100, tolerance=1e-6) -> Run:
    """katz_centrality_run, katz.rs:136-201 (KatzRun has PageRankRun's fields)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    n = off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_katz_run(n, _p(off, _u64p), _p(nb, _u32p), float(alpha), float(beta),
                       int(max_iterations), float(tolerance), _p(s, _f64p), _p(it, _u64p),
                       _p(d, _f64p), _p(c, _i32p))
    return _run(s, it, d, c, n)


def eigenvector_run(off, nb, max_iterations=100, tolerance=1e-6) -> Run:
    """eigenvector_centrality_run, eigenvector.rs:49-125."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    n = off.size - 1
    s, it, d, c = _tail(n)
    lib().orc_eigenvector_run(n, _p(off, _u64p), _p(nb, _u32p), int(max_iterations),
                              float(tolerance), _p(s, _f64p), _p(it, _u64p), _p(d, _f64p),
                              _p(c, _i32p))
    return _run(s, it, d, c, n)


@dataclass
class HitsRun:
    """hits, centrality.rs:114-166: (hub, authority), plus the loop count and
    the last L2 changes, which the reference does not return."""

    hub: np.ndarray
    auth: np.ndarray
    iterations: int
    auth_diff: float
    hub_diff: float


def hits(off, nb, max_iter=100, tol=1e-6) -> HitsRun:
    """hits, centrality.rs:114-166."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    nb = np.ascontiguousarray(nb, dtype=np.uint32)
    n = off.size - 1
    hub = np.zeros(max(n, 1))
    auth = np.zeros(max(n, 1))
    it = np.zeros(1, dtype=np.uint64)
    ad = np.zeros(1)
    hd = np.zeros(1)
    lib().orc_hits(n, _p(off, _u64p), _p(nb, _u32p), int(max_iter), float(tol), _p(hub, _f64p),
                   _p(auth, _f64p), _p(it, _u64p), _p(ad, _f64p), _p(hd, _f64p))
    return HitsRun(hub[:n].copy(), auth[:n].copy(), int(it[0]), float(ad[0]), float(hd[0]))


def top_k(scores, k):
    """top_k, topk.rs:17-42: [(index, score)], descending; equal scores by
    ascending index (the reference's unstable sort leaves that open)."""
    s = np.ascontiguousarray(scores, dtype=np.float64)
    k = int(k)
    ids = np.zeros(max(min(k, s.size), 1), dtype=np.uint64)
    vals = np.zeros(max(min(k, s.size), 1), dtype=np.float64)
    cnt = lib().orc_top_k(s.size, _p(s, _f64p), min(k, s.size), _p(ids, _u64p), _p(vals, _f64p))
    return [(int(i), float(v)) for i, v in zip(ids[:cnt], vals[:cnt])]
