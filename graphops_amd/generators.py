"""Synthetic inputs of the benchmark configurations (host C++ generators).

The reference crate has no graph generator; these are new and documented in
include/graphops_amd.h (gops_gen_*).  Every draw is a pure function of
(seed, counter), so results do not depend on the thread count.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


class Edges:
    """Host neighbour lists (adapter order) owned by the library.

    offsets[n+1] (uint64), targets[m] (uint32), weights[m] (float64 or None)
    are zero-copy numpy views valid while this object lives."""

    def __init__(self, handle):
        self._h = handle
        n = ctypes.c_uint64()
        m = ctypes.c_uint64()
        off = L._u64p()
        tg = L._u32p()
        w = L._f64p()
        rc = L.lib().gops_edges_view(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(off),
                                     ctypes.byref(tg), ctypes.byref(w))
        if rc != 0:
            raise RuntimeError(f"gops_edges_view failed ({rc})")
        self.n = int(n.value)
        self.m = int(m.value)
        self.offsets = np.ctypeslib.as_array(off, shape=(self.n + 1,))
        self.targets = (np.ctypeslib.as_array(tg, shape=(self.m,)) if self.m
                        else np.zeros(0, dtype=np.uint32))
        self.weights = (np.ctypeslib.as_array(w, shape=(self.m,))
                        if (self.m and bool(w)) else None)

    def out_degree(self):
        return np.diff(self.offsets).astype(np.int64)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib().gops_edges_destroy(h)
            self._h = None


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0, weights="unit", p0=0.0,
         p1=1.0, rounding="f64") -> Edges:
    """R-MAT graph, self-loops and duplicates removed.  weights: "unit",
    "uniform" (in [p0, p1)) or "lognormal" (mu=p0, sigma=p1); rounding of the
    drawn weights: "f64", "f32" or "bf16"."""
    kind = {"unit": L.GEN_UNIT, "uniform": L.GEN_UNIFORM, "lognormal": L.GEN_LOGNORMAL}[weights]
    rnd = {"f64": L.ROUND_F64, "f32": L.ROUND_F32, "bf16": L.ROUND_BF16}[rounding]
    h = ctypes.c_void_p()
    rc = L.lib().gops_gen_rmat(scale, edge_factor, a, b, c, seed, kind, p0, p1, rnd,
                               ctypes.byref(h))
    if rc != 0:
        raise ValueError(f"gops_gen_rmat failed ({rc})")
    return Edges(h)


def erdos_renyi(n, p, seed) -> Edges:
    """Directed G(n, p) over ordered pairs u != v (unit weights)."""
    h = ctypes.c_void_p()
    rc = L.lib().gops_gen_erdos_renyi(n, p, seed, ctypes.byref(h))
    if rc != 0:
        raise ValueError(f"gops_gen_erdos_renyi failed ({rc})")
    return Edges(h)


def seed_sets(edges: Edges, num_queries, min_seeds, max_seeds, log_uniform=False, seed=0):
    """Seed sets drawn proportional to out-degree; returns (offsets, seeds)."""
    off = np.zeros(num_queries + 1, dtype=np.uint64)
    cap = num_queries * max_seeds
    seeds = np.zeros(max(cap, 1), dtype=np.uint32)
    rc = L.lib().gops_gen_seed_sets(edges._h, num_queries, min_seeds, max_seeds,
                                    int(bool(log_uniform)), seed, L.ptr(off, L._u64p),
                                    L.ptr(seeds, L._u32p), cap)
    if rc != 0:
        raise ValueError(f"gops_gen_seed_sets failed ({rc})")
    return off, seeds[: int(off[-1])].copy()
