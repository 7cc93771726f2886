"""graphops_amd -- graphops' PageRank family on AMD MI355X (gfx950).

Python mirror of the reference's operator API (crate graphops v0.4.1,
src/pagerank.rs and src/ppr.rs; Katz, eigenvector and HITS centrality from
src/katz.rs, src/eigenvector.rs and src/centrality.rs): same function names, the same
PageRankConfig / PageRankRun, the same checked/unchecked split and error
messages.  Every call goes through the C ABI of libgraphops_amd.so
(include/graphops_amd.h) to hand-written HIP kernels; there is no CPU
fallback.

Adapters (src/graph.rs):
  AdjacencyMatrix(matrix)   graph.rs:122-142 (edge iff entry > 0.0)
  DiGraph                   the petgraph adapter, graph.rs:144-158, 199-212
  any object with node_count() / neighbors(u) [/ edge_weight(u, v)]
  DeviceGraph               a graph already resident on the GPU
"""
from __future__ import annotations

import atexit
import ctypes
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import generators  # noqa: F401

__all__ = [
    "PageRankConfig", "PageRankRun", "GraphopsError", "InvalidParameter", "IndexOutOfBounds",
    "DimensionMismatch", "DeviceError", "AdjacencyMatrix", "DiGraph", "DeviceGraph", "Context",
    "default_context", "pagerank", "pagerank_run", "pagerank_checked", "pagerank_checked_run",
    "pagerank_weighted", "pagerank_weighted_run", "pagerank_weighted_checked",
    "pagerank_weighted_checked_run", "personalized_pagerank", "personalized_pagerank_run",
    "personalized_pagerank_checked", "personalized_pagerank_checked_run", "ppr_batch_run",
    "top_k", "result_top_k", "result_top_k_noalloc", "normalize", "Server", "graph_tune_rows", "graph_set_rows",
    "ppr_tile_info", "KatzConfig", "KatzRun", "katz_centrality",
    "katz_centrality_run", "katz_centrality_checked", "EigenvectorRun",
    "eigenvector_centrality", "eigenvector_centrality_run", "hits", "hits_run",
]


# ---------------------------------------------------------------- errors
class GraphopsError(Exception):
    """Base of the errors below (graphops::Error, lib.rs:108-123)."""


class InvalidParameter(GraphopsError, ValueError):
    """Error::InvalidParameter; str() is the reference's Display text."""


class IndexOutOfBounds(GraphopsError, IndexError):
    """Error::IndexOutOfBounds."""


class DimensionMismatch(GraphopsError, ValueError):
    """Error::DimensionMismatch."""


class DeviceError(GraphopsError, RuntimeError):
    """A HIP failure (no reference counterpart)."""


class OutOfMemory(DeviceError):
    pass


class NotSupported(GraphopsError):
    pass


_ERRORS = {
    L.GOPS_INVALID_PARAMETER: InvalidParameter,
    L.GOPS_INDEX_OUT_OF_BOUNDS: IndexOutOfBounds,
    L.GOPS_DIMENSION_MISMATCH: DimensionMismatch,
    L.GOPS_DEVICE_ERROR: DeviceError,
    L.GOPS_OUT_OF_MEMORY: OutOfMemory,
    L.GOPS_NOT_SUPPORTED: NotSupported,
}


def _check(rc, ctx=None, what=""):
    if rc == L.GOPS_OK:
        return
    msg = ""
    if ctx is not None:
        msg = (L.lib().gops_last_error(ctx) or b"").decode()
    raise _ERRORS.get(rc, GraphopsError)(msg or f"{what} failed with status {rc}")


# Objects holding device resources.  At interpreter exit they are closed in
# dependency order (servers, graphs, then contexts): finalizers alone may run
# in any order then, and a graph or server must not outlive its context.
_LIVE = {"server": weakref.WeakSet(), "graph": weakref.WeakSet(), "context": weakref.WeakSet()}


@atexit.register
def _close_all():
    for kind in ("server", "graph", "context"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


# ---------------------------------------------------------------- config/run
@dataclass
class PageRankConfig:
    """PageRankConfig, pagerank.rs:37-59 (defaults 0.85 / 100 / 1e-6)."""

    damping: float = 0.85
    max_iterations: int = 100
    tolerance: float = 1e-6

    def validate(self) -> None:
        """PageRankConfig::validate, pagerank.rs:63-85."""
        d, t = float(self.damping), float(self.tolerance)
        if not np.isfinite(d):
            raise InvalidParameter("invalid parameter: damping must be finite")
        if not (0.0 <= d <= 1.0):
            raise InvalidParameter("invalid parameter: damping must be in [0,1]")
        if int(self.max_iterations) == 0:
            raise InvalidParameter("invalid parameter: max_iterations must be > 0")
        if not np.isfinite(t) or t <= 0.0:
            raise InvalidParameter("invalid parameter: tolerance must be finite and > 0")

    def _c(self):
        return L.PageRankConfigC(float(self.damping), int(self.max_iterations),
                                 float(self.tolerance))


@dataclass
class PageRankRun:
    """PageRankRun, pagerank.rs:14-29."""

    scores: np.ndarray
    iterations: int
    diff_l1: float
    converged: bool
    stats: dict = field(default_factory=dict, repr=False)


# ---------------------------------------------------------------- context
class Context:
    """One HIP device + stream + the memory it allocates (gops_context)."""

    def __init__(self, device: int = 0, precision: str = "mixed", anchor_period: int = 32,
                 normalize: bool = False, timing: bool = False, graph_build: str = "device"):
        h = ctypes.c_void_p()
        rc = L.lib().gops_context_create(device, ctypes.byref(h))
        if rc != 0:
            raise DeviceError(f"gops_context_create(device={device}) failed with status {rc}; "
                              "a HIP-capable GPU is required")
        self._h = h
        self.device = device
        _LIVE["context"].add(self)
        self.set_precision(precision)
        self.set_option(L.OPT_ANCHOR_PERIOD, anchor_period)
        self.set_option(L.OPT_NORMALIZE, int(bool(normalize)))
        self.set_option(L.OPT_TIMING, int(bool(timing)))
        self.set_graph_build(graph_build)

    @property
    def handle(self):
        return self._h

    def set_option(self, opt, value):
        _check(L.lib().gops_context_set_option(self._h, opt, float(value)), self._h, "set_option")

    def set_precision(self, precision: str):
        self.set_option(L.OPT_PRECISION,
                        {"mixed": L.PRECISION_MIXED, "fp64": L.PRECISION_FP64}[precision])

    def set_graph_build(self, where: str):
        """Where graphs are preprocessed: "device" (default) or "host"."""
        self.set_option(L.OPT_GRAPH_BUILD,
                        {"device": L.BUILD_DEVICE, "host": L.BUILD_HOST}[where])

    def last_stats(self) -> dict:
        s = L.RunStatsC()
        _check(L.lib().gops_last_run_stats(self._h, ctypes.byref(s)), self._h, "stats")
        return {k: getattr(s, k) for k, _ in L.RunStatsC._fields_}

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gops_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---------------------------------------------------------------- adapters
class AdjacencyMatrix:
    """Row-major adjacency-matrix adapter, graph.rs:122-142: matrix[i][j] is
    the weight of edge i -> j; an entry is an edge iff > 0.0."""

    def __init__(self, matrix):
        self.matrix = np.ascontiguousarray(np.asarray(matrix, dtype=np.float64))
        if self.matrix.ndim != 2 or self.matrix.shape[0] != self.matrix.shape[1]:
            if self.matrix.size == 0:
                self.matrix = self.matrix.reshape(0, 0)
            else:
                raise ValueError("AdjacencyMatrix needs a square matrix")

    def node_count(self) -> int:
        return self.matrix.shape[0]

    def neighbors(self, node: int):
        return [int(j) for j in np.nonzero(self.matrix[node] > 0.0)[0]]

    def out_degree(self, node: int) -> int:
        return len(self.neighbors(node))

    def edge_weight(self, source: int, target: int) -> float:
        return float(self.matrix[source, target])


class DiGraph:
    """Edge-list graph with petgraph adapter semantics (graph.rs:144-158,
    199-212): neighbors(u) lists every out-edge, parallel edges repeated, most
    recently added first (petgraph 0.6 Graph keeps outgoing edges in a
    linked list with new edges at the head); edge_weight(u, v) is the weight
    of the first edge u -> v met in that order (find_edge), 0.0 if none."""

    def __init__(self, n: int = 0):
        self._out = [[] for _ in range(n)]  # (target, weight) in insertion order
        self._version = 0

    def add_node(self) -> int:
        self._out.append([])
        self._version += 1
        return len(self._out) - 1

    def add_edge(self, u: int, v: int, weight: float = 1.0) -> None:
        self._out[u].append((int(v), float(weight)))
        self._version += 1

    def node_count(self) -> int:
        return len(self._out)

    def neighbors(self, node: int):
        return [v for v, _ in reversed(self._out[node])]

    def out_degree(self, node: int) -> int:
        return len(self._out[node])

    def edge_weight(self, source: int, target: int) -> float:
        for v, w in reversed(self._out[source]):
            if v == target:
                return w
        return 0.0

    def _lists(self, weighted):
        n = len(self._out)
        off = np.zeros(n + 1, dtype=np.uint64)
        for u in range(n):
            off[u + 1] = off[u] + len(self._out[u])
        nb = np.zeros(int(off[-1]), dtype=np.uint32)
        ew = np.zeros(int(off[-1]), dtype=np.float64) if weighted else None
        for u in range(n):
            nbrs = self.neighbors(u)
            nb[off[u]:off[u + 1]] = nbrs
            if weighted:
                ew[off[u]:off[u + 1]] = [self.edge_weight(u, v) for v in nbrs]
        return off, nb, ew


_STORAGE = {"none": L.WEIGHT_NONE, "f64": L.WEIGHT_F64, "f32": L.WEIGHT_F32,
            "bf16": L.WEIGHT_BF16}


class DeviceGraph:
    """A graph resident on the GPU (gops_graph): the CSR of in-edges built
    once from the adapter's neighbour lists (pagerank.rs:136, 226-234)."""

    def __init__(self, handle, ctx: Context, weighted: bool, weight_storage: int = None):
        self._h = handle
        self.ctx = ctx
        self.weighted = weighted
        self.weight_storage = (weight_storage if weight_storage is not None
                               else (L.WEIGHT_F64 if weighted else L.WEIGHT_NONE))
        _LIVE["graph"].add(self)
        n = ctypes.c_uint64()
        m = ctypes.c_uint64()
        L.lib().gops_graph_info(handle, ctypes.byref(n), ctypes.byref(m))
        self.n = int(n.value)
        self.nnz = int(m.value)

    @property
    def handle(self):
        return self._h

    def node_count(self) -> int:
        return self.n

    def export(self) -> dict:
        """Host copies of the preprocessed graph (gops_graph_export): row_ptr,
        col, weights (storage dtype, bf16 as uint16), order, inv_deg, inv_ws."""
        n, m = self.n, self.nnz
        out = {"row_ptr": np.zeros(n + 1, np.int32), "col": np.zeros(m, np.int32),
               "order": np.zeros(n, np.uint32), "inv_deg": np.zeros(n)}
        w = ws = None
        if self.weighted:
            storage = self.weight_storage
            w = np.zeros(m, {L.WEIGHT_F64: np.float64, L.WEIGHT_F32: np.float32,
                             L.WEIGHT_BF16: np.uint16}[storage])
            ws = np.zeros(n)
        _check(L.lib().gops_graph_export(self._h, L.ptr(out["row_ptr"], L._i32p),
                                         L.ptr(out["col"], L._i32p),
                                         None if w is None else w.ctypes.data_as(L._vp),
                                         L.ptr(out["order"], L._u32p),
                                         L.ptr(out["inv_deg"], L._f64p), L.ptr(ws, L._f64p)),
               self.ctx.handle, "gops_graph_export")
        out["weights"], out["inv_ws"] = w, ws
        return out

    @classmethod
    def from_lists(cls, n, offsets, targets, edge_weights=None, weight_storage="f64",
                   ctx: Context = None):
        """Neighbour lists in adapter order: targets[offsets[u]:offsets[u+1]]
        = Graph::neighbors(u); edge_weights[j] = edge_weight(u, targets[j])."""
        ctx = ctx or default_context()
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        if off.size != n + 1:
            raise DimensionMismatch(f"dimension mismatch: {off.size} vs {n + 1}")
        storage = L.WEIGHT_NONE
        ew = None
        if edge_weights is not None:
            storage = _STORAGE[weight_storage]
            ew = np.ascontiguousarray(edge_weights, dtype=np.float64)
            if ew.size != tg.size:
                raise DimensionMismatch(f"dimension mismatch: {ew.size} vs {tg.size}")
        h = ctypes.c_void_p()
        rc = L.lib().gops_graph_create(ctx.handle, n, L.ptr(off, L._u64p), L.ptr(tg, L._u32p),
                                       L.ptr(ew, L._f64p), storage, ctypes.byref(h))
        _check(rc, ctx.handle, "gops_graph_create")
        return cls(h, ctx, storage != L.WEIGHT_NONE, storage)

    @classmethod
    def from_dense(cls, matrix, weighted=True, weight_storage="f64", ctx: Context = None):
        """AdjacencyMatrix semantics (graph.rs:122-142)."""
        ctx = ctx or default_context()
        mat = np.ascontiguousarray(np.asarray(matrix, dtype=np.float64))
        n = mat.shape[0] if mat.ndim == 2 else 0
        storage = _STORAGE[weight_storage] if weighted else L.WEIGHT_NONE
        h = ctypes.c_void_p()
        rc = L.lib().gops_graph_from_dense(ctx.handle, n, L.ptr(mat, L._f64p), storage,
                                           ctypes.byref(h))
        _check(rc, ctx.handle, "gops_graph_from_dense")
        return cls(h, ctx, storage != L.WEIGHT_NONE, storage)

    @classmethod
    def from_edges(cls, edges, weight_storage="f64", ctx: Context = None):
        """From a generators.Edges (weights used when present)."""
        return cls.from_lists(edges.n, edges.offsets, edges.targets, edges.weights,
                              weight_storage, ctx)

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gops_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _device_graph(graph, weighted: bool, ctx: Context = None) -> DeviceGraph:
    if isinstance(graph, DeviceGraph):
        return graph
    ctx = ctx or default_context()
    if isinstance(graph, AdjacencyMatrix):
        return DeviceGraph.from_dense(graph.matrix, weighted=weighted, ctx=ctx)
    if isinstance(graph, DiGraph):
        key = (id(ctx), weighted, graph._version)
        cache = getattr(graph, "_gops_cache", None)
        if cache is not None and cache[0] == key:
            return cache[1]
        off, nb, ew = graph._lists(weighted)
        dg = DeviceGraph.from_lists(graph.node_count(), off, nb, ew, ctx=ctx)
        graph._gops_cache = (key, dg)
        return dg
    # generic Graph / WeightedGraph adapter (graph.rs:26-39, 67-72)
    n = int(graph.node_count())
    lists = [list(graph.neighbors(u)) for u in range(n)]
    off = np.zeros(n + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(x) for x in lists]) if n else []
    nb = np.array([v for x in lists for v in x], dtype=np.uint32)
    ew = None
    if weighted:
        ew = np.array([graph.edge_weight(u, v) for u in range(n) for v in lists[u]],
                      dtype=np.float64)
    return DeviceGraph.from_lists(n, off, nb, ew, ctx=ctx)


# ---------------------------------------------------------------- operators
def _run(fn_name, graph, config, weighted, pers=None, checked=False, ctx=None,
         with_scores=True):
    config = config if config is not None else PageRankConfig()
    dg = _device_graph(graph, weighted, ctx)
    c = dg.ctx.handle
    n = dg.n
    scores = np.zeros(max(n, 1), dtype=np.float64) if with_scores else None
    info = L.RunInfoC()
    cfg = config._c()
    fn = getattr(L.lib(), fn_name)
    if pers is not None:
        p = np.ascontiguousarray(pers, dtype=np.float64)
        rc = fn(c, dg.handle, ctypes.byref(cfg), L.ptr(p, L._f64p), p.size,
                L.ptr(scores, L._f64p), ctypes.byref(info))
    else:
        rc = fn(c, dg.handle, ctypes.byref(cfg), L.ptr(scores, L._f64p), ctypes.byref(info))
    _check(rc, c, fn_name)
    return PageRankRun(scores[:n] if with_scores else None, int(info.iterations),
                       float(info.diff_l1), bool(info.converged), dg.ctx.last_stats())


def pagerank_run(graph, config: PageRankConfig = None, ctx: Context = None) -> PageRankRun:
    """pagerank_run, pagerank.rs:122-187."""
    return _run("gops_pagerank_run", graph, config, False, ctx=ctx)


def pagerank(graph, config: PageRankConfig = None, ctx: Context = None) -> np.ndarray:
    """pagerank, pagerank.rs:114-116."""
    return pagerank_run(graph, config, ctx).scores


def pagerank_checked_run(graph, config: PageRankConfig, ctx: Context = None) -> PageRankRun:
    """pagerank_checked_run, pagerank.rs:321-324."""
    return _run("gops_pagerank_checked_run", graph, config, False, ctx=ctx)


def pagerank_checked(graph, config: PageRankConfig, ctx: Context = None) -> np.ndarray:
    """pagerank_checked, pagerank.rs:91-94."""
    return pagerank_checked_run(graph, config, ctx).scores


def pagerank_weighted_run(graph, config: PageRankConfig = None,
                          ctx: Context = None) -> PageRankRun:
    """pagerank_weighted_run, pagerank.rs:210-289."""
    return _run("gops_pagerank_weighted_run", graph, config, True, ctx=ctx)


def pagerank_weighted(graph, config: PageRankConfig = None, ctx: Context = None) -> np.ndarray:
    """pagerank_weighted, pagerank.rs:201-203."""
    return pagerank_weighted_run(graph, config, ctx).scores


def pagerank_weighted_checked_run(graph, config: PageRankConfig,
                                  ctx: Context = None) -> PageRankRun:
    """pagerank_weighted_checked_run, pagerank.rs:328-336 (computes once)."""
    return _run("gops_pagerank_weighted_checked_run", graph, config, True, ctx=ctx)


def pagerank_weighted_checked(graph, config: PageRankConfig, ctx: Context = None) -> np.ndarray:
    """pagerank_weighted_checked, pagerank.rs:294-316."""
    return pagerank_weighted_checked_run(graph, config, ctx).scores


def personalized_pagerank_run(graph, config: PageRankConfig, personalization,
                              ctx: Context = None) -> PageRankRun:
    """personalized_pagerank_run, ppr.rs:90-156."""
    return _run("gops_personalized_pagerank_run", graph, config, False, pers=personalization,
                ctx=ctx)


def personalized_pagerank(graph, config: PageRankConfig, personalization,
                          ctx: Context = None) -> np.ndarray:
    """personalized_pagerank, ppr.rs:66-72."""
    return personalized_pagerank_run(graph, config, personalization, ctx).scores


def personalized_pagerank_checked_run(graph, config: PageRankConfig, personalization,
                                      ctx: Context = None) -> PageRankRun:
    """personalized_pagerank_checked_run, ppr.rs:77-85 (computes once)."""
    return _run("gops_personalized_pagerank_checked_run", graph, config, False,
                pers=personalization, ctx=ctx)


def personalized_pagerank_checked(graph, config: PageRankConfig, personalization,
                                  ctx: Context = None) -> np.ndarray:
    """personalized_pagerank_checked, ppr.rs:11-44."""
    return personalized_pagerank_checked_run(graph, config, personalization, ctx).scores


def ppr_batch_run(graph, configs, seed_offsets, seeds, seed_weights=None, ctx: Context = None,
                  with_scores=True, tile=None):
    """Many personalized_pagerank_run calls (ppr.rs:90-156) in one device run.
    Query q's personalization is 1.0 (or seed_weights[j]) on
    seeds[seed_offsets[q]:seed_offsets[q+1]] and 0 elsewhere.  Returns
    (scores[num_queries, n] or None, list of PageRankRun without scores)."""
    dg = _device_graph(graph, False, ctx)
    c = dg.ctx.handle
    if tile is not None:  # queries per tile; -1 = one query per device run
        dg.ctx.set_option(L.OPT_PPR_TILE, int(tile))
    q = len(seed_offsets) - 1
    if isinstance(configs, PageRankConfig):
        configs = [configs] * q
    cfgs = (L.PageRankConfigC * max(q, 1))(*[cf._c() for cf in configs])
    off = np.ascontiguousarray(seed_offsets, dtype=np.uint64)
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    sw = None if seed_weights is None else np.ascontiguousarray(seed_weights, dtype=np.float64)
    scores = np.zeros((q, dg.n), dtype=np.float64) if with_scores else None
    infos = (L.RunInfoC * max(q, 1))()
    rc = L.lib().gops_ppr_batch_run(c, dg.handle, q, cfgs, L.ptr(off, L._u64p),
                                    L.ptr(sd, L._u32p), L.ptr(sw, L._f64p),
                                    L.ptr(scores, L._f64p), infos)
    if tile is not None:
        dg.ctx.set_option(L.OPT_PPR_TILE, 0)
    _check(rc, c, "gops_ppr_batch_run")
    runs = [PageRankRun(None, int(infos[i].iterations), float(infos[i].diff_l1),
                        bool(infos[i].converged)) for i in range(q)]
    return scores, runs


def ppr_tile_info(graph, num_queries, ctx: Context = None):
    """(tile width a batch of num_queries queries would use, {width: measured
    ms per query-sweep}) -- gops_ppr_tile_info; runs the autotuner's trial on
    first use."""
    dg = _device_graph(graph, False, ctx)
    t = ctypes.c_uint32()
    cost = np.zeros(4, dtype=np.float64)
    rc = L.lib().gops_ppr_tile_info(dg.ctx.handle, dg.handle, int(num_queries), ctypes.byref(t),
                                    L.ptr(cost, L._f64p))
    _check(rc, dg.ctx.handle, "gops_ppr_tile_info")
    return int(t.value), {w: float(c) for w, c in zip((8, 16, 32, 64), cost)}


def graph_tune_rows(graph, ctx: Context = None) -> dict:
    """Autotune the pull sweeps' row plan of a device graph
    (gops_graph_tune_rows): the chosen cut-offs (row_min, hub_chunk,
    row_iters), the number of plans timed, and the f32 pull sweep's time
    with the default and the chosen cut-offs (us)."""
    dg = _device_graph(graph, False, ctx)
    r = L.RowPlanC()
    _check(L.lib().gops_graph_tune_rows(dg.ctx.handle, dg.handle, ctypes.byref(r)), dg.ctx.handle,
           "gops_graph_tune_rows")
    return {"row_min": r.row_min, "hub_chunk": r.hub_chunk, "row_iters": r.row_iters,
            "candidates": r.candidates, "default_us": r.default_us, "tuned_us": r.tuned_us}


def graph_set_rows(graph, row_min=4096, hub_chunk=65536, row_iters=8, ctx: Context = None):
    """Set the pull sweeps' row plan of a device graph (gops_graph_set_rows)."""
    dg = _device_graph(graph, False, ctx)
    r = L.RowPlanC(int(row_min), int(hub_chunk), int(row_iters), 0, 0.0, 0.0)
    _check(L.lib().gops_graph_set_rows(dg.ctx.handle, dg.handle, ctypes.byref(r)), dg.ctx.handle,
           "gops_graph_set_rows")


# ---------------------------------------------------------------- ranking
# ---------------------------------------------------------------- centrality
@dataclass
class KatzConfig:
    """KatzConfig, katz.rs:25-53 (defaults 0.1 / 1.0 / 100 / 1e-6)."""

    alpha: float = 0.1
    beta: float = 1.0
    max_iterations: int = 100
    tolerance: float = 1e-6

    def validate(self) -> None:
        """KatzConfig::validate, katz.rs:57-84."""
        a, b, t = float(self.alpha), float(self.beta), float(self.tolerance)
        if not np.isfinite(a) or a <= 0.0:
            raise InvalidParameter("invalid parameter: alpha must be finite and > 0")
        if a >= 1.0:
            raise InvalidParameter(
                "invalid parameter: alpha must be < 1 (required for convergence; tighter bound "
                "alpha < 1/spectral_radius(A) may be needed for large graphs)")
        if not np.isfinite(b):
            raise InvalidParameter("invalid parameter: beta must be finite")
        if int(self.max_iterations) == 0:
            raise InvalidParameter("invalid parameter: max_iterations must be > 0")
        if not np.isfinite(t) or t <= 0.0:
            raise InvalidParameter("invalid parameter: tolerance must be finite and > 0")

    def _c(self):
        return L.KatzConfigC(float(self.alpha), float(self.beta), int(self.max_iterations),
                             float(self.tolerance))


@dataclass
class KatzRun:
    """KatzRun, katz.rs:87-104."""

    scores: np.ndarray
    iterations: int
    diff_l1: float
    converged: bool
    stats: dict = field(default_factory=dict, repr=False)


@dataclass
class EigenvectorRun:
    """EigenvectorRun, eigenvector.rs:12-23."""

    scores: np.ndarray
    iterations: int
    diff_l1: float
    converged: bool
    stats: dict = field(default_factory=dict, repr=False)


def _katz(graph, config, checked, ctx):
    config = config if config is not None else KatzConfig()
    dg = _device_graph(graph, False, ctx)
    c = dg.ctx.handle
    scores = np.zeros(max(dg.n, 1))
    info = L.RunInfoC()
    cfg = config._c()
    name = "gops_katz_centrality_checked_run" if checked else "gops_katz_centrality_run"
    rc = getattr(L.lib(), name)(c, dg.handle, ctypes.byref(cfg), L.ptr(scores, L._f64p),
                                ctypes.byref(info))
    _check(rc, c, name)
    return KatzRun(scores[:dg.n], int(info.iterations), float(info.diff_l1),
                   bool(info.converged), dg.ctx.last_stats())


def katz_centrality_run(graph, config: KatzConfig = None, ctx: Context = None) -> KatzRun:
    """katz_centrality_run, katz.rs:136-201."""
    return _katz(graph, config, False, ctx)


def katz_centrality(graph, config: KatzConfig = None, ctx: Context = None) -> np.ndarray:
    """katz_centrality, katz.rs:131-133."""
    return _katz(graph, config, False, ctx).scores


def katz_centrality_checked(graph, config: KatzConfig, ctx: Context = None) -> np.ndarray:
    """katz_centrality_checked, katz.rs:207-210."""
    return _katz(graph, config, True, ctx).scores


def eigenvector_centrality_run(graph, max_iterations: int = 100, tolerance: float = 1e-6,
                               ctx: Context = None) -> EigenvectorRun:
    """eigenvector_centrality_run, eigenvector.rs:49-125."""
    dg = _device_graph(graph, False, ctx)
    c = dg.ctx.handle
    scores = np.zeros(max(dg.n, 1))
    info = L.RunInfoC()
    rc = L.lib().gops_eigenvector_centrality_run(c, dg.handle, int(max_iterations),
                                                 float(tolerance), L.ptr(scores, L._f64p),
                                                 ctypes.byref(info))
    _check(rc, c, "gops_eigenvector_centrality_run")
    return EigenvectorRun(scores[:dg.n], int(info.iterations), float(info.diff_l1),
                          bool(info.converged), dg.ctx.last_stats())


def eigenvector_centrality(graph, ctx: Context = None) -> np.ndarray:
    """eigenvector_centrality, eigenvector.rs:26-28 (100 iterations, 1e-6)."""
    return eigenvector_centrality_run(graph, 100, 1e-6, ctx).scores


def hits_run(graph, max_iter: int = 100, tol: float = 1e-6, ctx: Context = None):
    """hits (centrality.rs:114-166) with the iteration count, the last hub
    change and whether both changes ended below tol: (hub, authority, info)."""
    dg = _device_graph(graph, False, ctx)
    c = dg.ctx.handle
    hub = np.zeros(max(dg.n, 1))
    auth = np.zeros(max(dg.n, 1))
    info = L.RunInfoC()
    rc = L.lib().gops_hits(c, dg.handle, int(max_iter), float(tol), L.ptr(hub, L._f64p),
                           L.ptr(auth, L._f64p), ctypes.byref(info))
    _check(rc, c, "gops_hits")
    return hub[:dg.n], auth[:dg.n], {"iterations": int(info.iterations),
                                     "hub_diff": float(info.diff_l1),
                                     "converged": bool(info.converged)}


def hits(graph, max_iter: int, tol: float, ctx: Context = None):
    """hits, centrality.rs:114-166: (hub_scores, authority_scores)."""
    hub, auth, _ = hits_run(graph, max_iter, tol, ctx)
    return hub, auth


def top_k(scores, k, ctx: Context = None):
    """top_k, topk.rs:17-42, on the device (gops_top_k): the k largest
    finite, positive scores as (node, score), descending; at a tie on the
    k-th score the entries the reference's heap keeps, equal scores by
    ascending node."""
    ctx = ctx or default_context()
    s = np.ascontiguousarray(scores, dtype=np.float64)
    k = int(k)
    if k == 0 or s.size == 0:
        return []
    ids = np.zeros(k, dtype=np.uint32)
    vals = np.zeros(k, dtype=np.float64)
    cnt = ctypes.c_uint32()
    _check(L.lib().gops_top_k(ctx.handle, L.ptr(s, L._f64p), s.size, k, L.ptr(ids, L._u32p),
                              L.ptr(vals, L._f64p), ctypes.byref(cnt)), ctx.handle, "gops_top_k")
    return [(int(i), float(v)) for i, v in zip(ids[:cnt.value], vals[:cnt.value])]


def result_top_k(graph: "DeviceGraph", k):
    """top_k of the last PageRank-family run on graph, selected where the
    scores stay (gops_result_top_k): only the k pairs cross PCIe."""
    k = int(k)
    ids = np.zeros(max(k, 1), dtype=np.uint32)
    vals = np.zeros(max(k, 1), dtype=np.float64)
    cnt = ctypes.c_uint32()
    c = graph.ctx.handle
    _check(L.lib().gops_result_top_k(c, graph.handle, k, L.ptr(ids, L._u32p), L.ptr(vals, L._f64p),
                                     ctypes.byref(cnt)), c, "gops_result_top_k")
    return [(int(i), float(v)) for i, v in zip(ids[:cnt.value], vals[:cnt.value])]


def result_top_k_noalloc(graph: "DeviceGraph", k):
    """result_top_k without allocating device memory
    (gops_result_top_k_noalloc): the same pairs, k <= 4096, scratch in the
    library's file-scope device array (calls that overlap on one GPU race)."""
    k = int(k)
    ids = np.zeros(max(k, 1), dtype=np.uint32)
    vals = np.zeros(max(k, 1), dtype=np.float64)
    cnt = ctypes.c_uint32()
    c = graph.ctx.handle
    _check(L.lib().gops_result_top_k_noalloc(c, graph.handle, k, L.ptr(ids, L._u32p),
                                             L.ptr(vals, L._f64p), ctypes.byref(cnt)),
           c, "gops_result_top_k_noalloc")
    return [(int(i), float(v)) for i, v in zip(ids[:cnt.value], vals[:cnt.value])]


def normalize(scores):
    """normalize, topk.rs:54-61 (in place; no-op unless the sum is > 0)."""
    total = float(np.sum(scores))
    if total > 0.0:
        scores /= total
    return scores


# ---------------------------------------------------------------- serving
class Server:
    """Continuous-batching personalized PageRank server over a resident graph
    (gops_server_*).  Queries wait in a host queue and take a device slot as
    soon as one frees up; step() advances every tile with live queries."""

    def __init__(self, graph: DeviceGraph, tile: int = 16, max_slots: int = 0,
                 hbm_budget_bytes: float = 0.0, topk: int = 0):
        self.graph = graph
        self.ctx = graph.ctx
        cfg = L.ServerConfigC(int(max_slots), float(hbm_budget_bytes), int(tile), int(topk))
        h = ctypes.c_void_p()
        rc = L.lib().gops_server_create(self.ctx.handle, graph.handle, ctypes.byref(cfg),
                                        ctypes.byref(h))
        _check(rc, self.ctx.handle, "gops_server_create")
        self._h = h
        self.topk = int(topk)
        _LIVE["server"].add(self)

    def submit(self, config: PageRankConfig, seeds, seed_weights=None) -> int:
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        sw = None if seed_weights is None else np.ascontiguousarray(seed_weights, dtype=np.float64)
        t = ctypes.c_uint64()
        cfg = config._c()
        rc = L.lib().gops_server_submit(self._h, ctypes.byref(cfg), L.ptr(sd, L._u32p), sd.size,
                                        L.ptr(sw, L._f64p), ctypes.byref(t))
        _check(rc, self.ctx.handle, "gops_server_submit")
        return int(t.value)

    def step(self, sweeps: int = 4) -> int:
        act = ctypes.c_uint32()
        _check(L.lib().gops_server_step(self._h, int(sweeps), ctypes.byref(act)), self.ctx.handle,
               "gops_server_step")
        return int(act.value)

    def poll(self, with_scores=None):
        """One finished query or None: (ticket, PageRankRun, submit_ms,
        finish_ms, topk) -- topk is a list of (node, score) if enabled; a
        top-k server returns no score vector (PageRankRun.scores is None)."""
        if with_scores is None:
            with_scores = not self.topk
        ready = ctypes.c_int()
        res = L.ServerResultC()
        scores = np.zeros(self.graph.n, dtype=np.float64) if with_scores else None
        ids = np.zeros(max(self.topk, 1), dtype=np.uint32) if self.topk else None
        tsc = np.zeros(max(self.topk, 1), dtype=np.float64) if self.topk else None
        rc = L.lib().gops_server_poll(self._h, ctypes.byref(ready), ctypes.byref(res),
                                      L.ptr(scores, L._f64p), L.ptr(ids, L._u32p),
                                      L.ptr(tsc, L._f64p))
        _check(rc, self.ctx.handle, "gops_server_poll")
        if not ready.value:
            return None
        run = PageRankRun(scores, int(res.info.iterations), float(res.info.diff_l1),
                          bool(res.info.converged))
        top = None
        if self.topk:
            top = [(int(i), float(s)) for i, s in zip(ids, tsc) if i != 0xFFFFFFFF]
        return int(res.ticket), run, float(res.submit_ms), float(res.finish_ms), top

    def stats(self) -> dict:
        q = ctypes.c_uint64()
        a = ctypes.c_uint64()
        f = ctypes.c_uint64()
        s = ctypes.c_uint32()
        _check(L.lib().gops_server_stats(self._h, ctypes.byref(q), ctypes.byref(a),
                                         ctypes.byref(f), ctypes.byref(s)), self.ctx.handle, "stats")
        return {"queued": q.value, "active": a.value, "finished": f.value, "slots": s.value}

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gops_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
