"""The preprocessed graph built on the device against the host build and an
independent numpy statement of the same arrays (bit for bit).

The arrays (gops_graph_export): rows are nodes ordered by in-degree,
descending, the nodes without in- and out-edges after the other nodes
without in-edges, ties by id; each row lists the new ids of its sources,
ascending, duplicate edges in adapter order (the order in which the
reference's push loop adds them, pagerank.rs:154-162) with the weight
max(w, 0) (pagerank.rs:257); 1/out-degree counts duplicate entries
(pagerank.rs:137) and 1/out_wsum sums max(w, 0) in neighbour order
(pagerank.rs:227-234).
"""
import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx_dev():
    return ga.Context(0, graph_build="device")


@pytest.fixture(scope="module")
def ctx_host():
    return ga.Context(0, graph_build="host")


def bf16_bits(x32):
    u = np.asarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    rne = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    keep = (u >> 16) | np.where((u & 0x7FFFFF) != 0, np.uint64(0x40), np.uint64(0))
    return np.where(special, keep, rne).astype(np.uint16)


def numpy_csrt(n, off, nb, ew, storage):
    """The preprocessed arrays, restated with numpy sorts."""
    off = np.asarray(off, np.int64)
    nb = np.asarray(nb, np.int64)
    m = nb.size
    deg = np.diff(off)
    src = np.repeat(np.arange(n), deg)
    indeg = np.bincount(nb, minlength=n)
    isolated = (indeg == 0) & (np.diff(off) == 0)  # no in-, no out-edges: the tail
    order = np.lexsort((np.arange(n), isolated, -indeg))
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    row, col = rank[nb], rank[src]
    perm = np.lexsort((np.arange(m), col, row))
    out = {"row_ptr": np.concatenate([[0], np.cumsum(indeg[order])]).astype(np.int32),
           "col": col[perm].astype(np.int32), "order": order.astype(np.uint32)}
    inv_deg = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    out["inv_deg"] = inv_deg[order]
    if ew is not None:
        w = np.fmax(np.asarray(ew, np.float64), 0.0)
        ws = np.zeros(n)
        for u in range(n):
            s = 0.0
            for j in range(off[u], off[u + 1]):
                s += float(w[j])
            ws[u] = s
        inv_ws = np.zeros(n)
        nz = ws != 0.0
        inv_ws[nz] = 1.0 / ws[nz]
        out["inv_ws"] = inv_ws[order]
        wp = w[perm]
        out["weights"] = {"f64": wp, "f32": wp.astype(np.float32),
                          "bf16": bf16_bits(wp.astype(np.float32))}[storage]
    return out


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype.itemsize != b.dtype.itemsize:
        return False
    if a.size == 0:
        return True
    v = {8: np.uint64, 4: np.uint32, 2: np.uint16}[a.dtype.itemsize]
    return np.array_equal(a.view(v), b.view(v))


def check_build(ctx_dev, ctx_host, n, off, nb, ew=None, storage="f64"):
    want = numpy_csrt(n, off, nb, ew, storage)
    got = {}
    for where, cx in (("device", ctx_dev), ("host", ctx_host)):
        dg = ga.DeviceGraph.from_lists(n, off, nb, ew, storage, ctx=cx)
        got[where] = dg.export()
    for key, ref in want.items():
        for where in ("device", "host"):
            assert same_bits(got[where][key], ref), (where, key)
    return got


def random_multigraph(n, avg_deg, seed, weighted=True):
    """Neighbour lists with duplicates, self-loops and empty lists, in no
    particular order; weights log-normal with zeros, negatives, a NaN and
    an infinity among them."""
    rng = np.random.default_rng(seed)
    deg = rng.poisson(avg_deg, n)
    deg[rng.random(n) < 0.1] = 0
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    m = int(off[-1])
    # skewed targets: many in-degree ties and a few heavy rows
    nb = np.minimum((rng.pareto(1.2, m) * n / 50).astype(np.int64), n - 1)
    nb = rng.permutation(n)[nb].astype(np.uint32)
    # repeat some entries within a list (duplicates u -> v)
    for u in rng.choice(n, size=n // 10, replace=False):
        b, e = int(off[u]), int(off[u + 1])
        if e - b >= 2:
            nb[b + 1:e] = np.where(rng.random(e - b - 1) < 0.5, nb[b], nb[b + 1:e])
    ew = None
    if weighted:
        ew = rng.lognormal(0.0, 1.0, m)
        ew[rng.random(m) < 0.02] = 0.0
        ew[rng.random(m) < 0.01] *= -1.0
        if m > 10:
            ew[3] = np.nan
            ew[7] = np.inf
    return off, nb, ew


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
def test_device_build_weighted_multigraph(ctx_dev, ctx_host, storage):
    off, nb, ew = random_multigraph(3000, 20, seed=5)
    check_build(ctx_dev, ctx_host, 3000, off, nb, ew, storage)


@pytest.mark.parametrize("first_bad", ["none", "nan", "negative"])
def test_device_build_long_out_lists(ctx_dev, ctx_host, first_bad):
    """Out-lists longer than one wave (64 entries) are summed by a wave each:
    lengths around the 64-entry chunks, hubs next to each other, and the
    first invalid weight deep inside a long list."""
    rng = np.random.default_rng(11)
    n = 4000
    deg = rng.integers(0, 8, n)
    deg[:8] = [65, 64, 128, 129, 1000, 5000, 70000, 63]
    deg[2000] = 300
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    m = int(off[-1])
    nb = rng.integers(0, n, m).astype(np.uint32)
    ew = rng.lognormal(0.0, 2.0, m)  # a wide range: the summation order shows
    ew[rng.random(m) < 0.05] = 0.0
    if first_bad != "none":
        j = int(off[5]) + 4099  # list 5, chunk 64, lane 3
        ew[j] = np.nan if first_bad == "nan" else -0.5
        ew[int(off[6]) + 70] = -1.0 if first_bad == "nan" else np.inf
    check_build(ctx_dev, ctx_host, n, off, nb, ew, "f64")
    if first_bad != "none":
        msg = "finite" if first_bad == "nan" else "non-negative"
        for cx in (ctx_dev, ctx_host):
            dg = ga.DeviceGraph.from_lists(n, off, nb, ew, ctx=cx)
            with pytest.raises(ga.InvalidParameter, match=msg):
                ga.pagerank_weighted_checked(dg, ga.PageRankConfig(), ctx=cx)


def test_device_build_unweighted_multigraph(ctx_dev, ctx_host):
    off, nb, _ = random_multigraph(5000, 12, seed=6, weighted=False)
    check_build(ctx_dev, ctx_host, 5000, off, nb)


def test_device_build_hub_rows(ctx_dev, ctx_host):
    """A row above the hub chunk (65536 in-edges) and one between the LDS
    stage (4096) and the chunk: all three descriptor kinds in one plan."""
    n = 80000
    rng = np.random.default_rng(7)
    lists = []
    for u in range(n):
        t = [0] if u != 0 else []
        if u % 13 == 0:
            t.append(1)
        t.extend(rng.integers(0, n, size=int(rng.integers(0, 3))).tolist())
        lists.append(t)
    off = np.concatenate([[0], np.cumsum([len(t) for t in lists])]).astype(np.uint64)
    nb = np.array([v for t in lists for v in t], np.uint32)
    ew = rng.random(nb.size) + 0.05
    got = check_build(ctx_dev, ctx_host, n, off, nb, ew, "f32")
    rp = got["device"]["row_ptr"]
    assert rp[1] - rp[0] > 65536 and 4096 < rp[2] - rp[1] <= 65536
    cfg = ga.PageRankConfig(0.85, 100, 1e-10)
    res = []
    for cx in (ctx_dev, ctx_host):
        dg = ga.DeviceGraph.from_lists(n, off, nb, ew, "f32", ctx=cx)
        res.append(ga.pagerank_weighted_run(dg, cfg, ctx=cx))
    assert res[0].iterations == res[1].iterations
    assert np.array_equal(res[0].scores, res[1].scores)


def test_device_build_small_graphs(ctx_dev, ctx_host):
    check_build(ctx_dev, ctx_host, 0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    check_build(ctx_dev, ctx_host, 1, np.zeros(2, np.uint64), np.zeros(0, np.uint32))
    check_build(ctx_dev, ctx_host, 4, np.zeros(5, np.uint64), np.zeros(0, np.uint32),
                np.zeros(0), "f64")
    # a self-loop only, and a two-cycle with a parallel edge
    check_build(ctx_dev, ctx_host, 1, np.array([0, 1], np.uint64), np.array([0], np.uint32),
                np.array([2.0]), "bf16")
    check_build(ctx_dev, ctx_host, 2, np.array([0, 2, 3], np.uint64),
                np.array([1, 1, 0], np.uint32), np.array([3.0, 1.0, 0.5]), "f64")
    # every node dangling but one
    check_build(ctx_dev, ctx_host, 6, np.array([0, 0, 0, 5, 5, 5, 5], np.uint64),
                np.array([4, 0, 5, 0, 1], np.uint32))


def test_device_build_errors(ctx_dev, ctx_host):
    for cx in (ctx_dev, ctx_host):
        with pytest.raises(ga.IndexOutOfBounds, match="index out of bounds: 9"):
            ga.DeviceGraph.from_lists(3, np.array([0, 2, 3, 4], np.uint64),
                                      np.array([1, 9, 7, 2], np.uint32), ctx=cx)
        with pytest.raises(ga.InvalidParameter, match="non-decreasing"):
            ga.DeviceGraph.from_lists(2, np.array([0, 2, 1], np.uint64),
                                      np.array([1, 0], np.uint32), ctx=cx)
        # the first invalid weight in adapter order names the error
        # (pagerank.rs:300-314)
        for ws, msg in (([1.0, -1.0, np.nan], "non-negative"), ([1.0, np.inf, -2.0], "finite")):
            dg = ga.DeviceGraph.from_lists(2, np.array([0, 2, 3], np.uint64),
                                           np.array([1, 0, 0], np.uint32), np.array(ws), ctx=cx)
            with pytest.raises(ga.InvalidParameter, match=msg):
                ga.pagerank_weighted_checked(dg, ga.PageRankConfig(), ctx=cx)


def test_device_build_rmat_runs_identical(ctx_dev, ctx_host):
    """Same arrays, same plan: every run is bitwise the same."""
    e = generators.rmat(16, 16, seed=7, weights="lognormal", p0=0.0, p1=1.0, rounding="bf16")
    cfg = ga.PageRankConfig(0.85, 200, 1e-10)
    runs = []
    for cx in (ctx_dev, ctx_host):
        dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16", ctx=cx)
        runs.append((ga.pagerank_weighted_run(dg, cfg, ctx=cx), ga.pagerank_run(dg, cfg, ctx=cx)))
    for a, b in zip(runs[0], runs[1]):
        assert a.iterations == b.iterations and a.diff_l1 == b.diff_l1
        assert np.array_equal(a.scores, b.scores)
