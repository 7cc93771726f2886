"""GPU parity of the blocked f32 delta sweep (GOPS_OPT_DELTA_SWEEP =
GOPS_DELTA_SWEEP_BLOCKED): a scatter pass per chunk of sources and an
accumulate pass per bin of rows, fixed-point row sums.

  * against the f64 oracle with the bounds of the pull sweep: L1 within
    1e-2 * tolerance + 1e-12 for converged runs, iterations equal (or one
    apart when a residual sits within 1e-4 of the tolerance);
  * against the pull sweep on the same graph: the same iterations, L1 within
    2e-2 * tolerance + 1e-12 (both are that close to the oracle);
  * two runs give the same bits (integer sums: no dependence on scheduling);
  * the f64 anchor sweeps run blocked too (128-bit fixed-point row sums):
    against pull anchors on the same graph, the same iterations and L1
    within f64 rounding;
  * layouts that exercise every part: bins cut short by their rows' run
    bounds, and with bins of 4096 rows, bins split over several workgroups,
    several chunks and bins with n not a multiple
    of either, empty bins and rows, dangling nodes, every weight storage,
    personalization, and graphs without edges.
"""
from types import SimpleNamespace

import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import _lib as L
from graphops_amd import generators

pytestmark = pytest.mark.gpu


def make_ctx(mode):
    c = ga.Context(0, precision="mixed")
    c.set_option(L.OPT_DELTA_SWEEP, mode)
    return c


@pytest.fixture(scope="module")
def ctx_b():
    return make_ctx(L.DELTA_SWEEP_BLOCKED)


@pytest.fixture(scope="module")
def ctx_p():
    return make_ctx(L.DELTA_SWEEP_PULL)


def oracle_pull(oracle, e, kind, tol, max_it, p=None):
    in_off, in_src, in_w = oracle.transpose(e.offsets, e.targets, e.weights if kind == 1 else None)
    if kind == 1:
        scale = oracle.out_wsum(e.offsets, e.weights)
    else:
        scale = np.diff(e.offsets).astype(np.float64)
    return oracle.pull_run(kind, in_off, in_src, scale, w=in_w, p=p, damping=0.85,
                           max_iterations=max_it, tolerance=tol, threads=0)


def check(run, ref, tol):
    noise = 1e-15 + 4e-18 * ref.scores.size
    l1 = np.abs(run.scores - ref.scores).sum()
    assert run.converged == ref.converged
    assert l1 <= (1e-2 * tol + 1e-12 if ref.converged else 1e-7), l1
    if run.iterations != ref.iterations:
        assert abs(run.iterations - ref.iterations) == 1
        assert (abs(ref.diff_l1 - tol) <= 1e-4 * tol + noise
                or abs(run.diff_l1 - tol) <= 1e-4 * tol + noise)
    else:
        assert abs(run.diff_l1 - ref.diff_l1) <= 1e-4 * ref.diff_l1 + noise


def runs(ctx_b, ctx_p, e, kind, storage, cfg, p=None):
    out = []
    for cx in (ctx_b, ctx_p, ctx_b):
        dg = ga.DeviceGraph.from_edges(e, weight_storage=storage, ctx=cx)
        if kind == 2:
            out.append(ga.personalized_pagerank_run(dg, cfg, p, ctx=cx))
        elif kind == 1:
            out.append(ga.pagerank_weighted_run(dg, cfg, ctx=cx))
        else:
            out.append(ga.pagerank_run(dg, cfg, ctx=cx))
        dg.close()
    b, pl, b2 = out
    assert np.array_equal(b.scores, b2.scores) and b.diff_l1 == b2.diff_l1  # deterministic
    assert b.iterations == pl.iterations
    assert np.abs(b.scores - pl.scores).sum() <= 2e-2 * cfg.tolerance + 1e-12
    return b


@pytest.mark.parametrize("weights,storage", [("unit", "f64"), ("uniform", "f32"),
                                             ("lognormal", "bf16"), ("uniform", "f64")])
def test_rmat_weights(oracle, ctx_b, ctx_p, weights, storage):
    """R-MAT scale 16 (n = 65 536: 4 chunks, 16 bins)."""
    if weights == "unit":
        e = generators.rmat(16, 16, seed=7)
    else:
        e = generators.rmat(16, 16, seed=7, weights=weights,
                            p0=0.1 if weights == "uniform" else 0.0,
                            p1=1.0, rounding={"f32": "f32", "bf16": "bf16", "f64": "f64"}[storage])
    kind = 0 if weights == "unit" else 1
    tol = 1e-10
    ref = oracle_pull(oracle, e, kind, tol, 300)
    b = runs(ctx_b, ctx_p, e, kind, storage, ga.PageRankConfig(0.85, 300, tol))
    check(b, ref, tol)


def test_ragged_sizes_and_long_rows(oracle, ctx_b, ctx_p):
    """n = 70 001 (not a multiple of a chunk or a bin); node 5 has 140 000
    in-edges, node 9 has 70 000 out-edges, many rows and a whole bin without
    in-edges, dangling nodes."""
    n = 70001
    rng = np.random.default_rng(4)
    src = [np.arange(n, dtype=np.int64) % 60000, np.full(70000, 9),
           rng.integers(0, 60000, size=140000)]
    dst = [rng.integers(0, 50000, size=n), rng.integers(0, 60000, size=70000),
           np.full(140000, 5)]
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    dst[(dst >= 16384) & (dst < 20480)] = 3  # rows of one bin without in-edges
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    nb = dst.astype(np.uint32)
    w = rng.uniform(0.1, 1.0, size=nb.size).astype(np.float32).astype(np.float64)  # f32 storage
    e = SimpleNamespace(n=n, m=nb.size, offsets=off, targets=nb, weights=w)
    for kind, storage in ((0, "f64"), (1, "f32")):
        tol = 1e-10
        ref = oracle_pull(oracle, e, kind, tol, 300)
        b = runs(ctx_b, ctx_p, e, kind, storage, ga.PageRankConfig(0.85, 300, tol))
        check(b, ref, tol)


def test_personalized(oracle, ctx_b, ctx_p):
    e = generators.rmat(15, 16, seed=13)
    rng = np.random.default_rng(8)
    p = np.zeros(e.n)
    p[rng.choice(e.n, size=40, replace=False)] = rng.random(40)
    tol = 1e-9
    ref = oracle_pull(oracle, e, 2, tol, 300, p=p)
    b = runs(ctx_b, ctx_p, e, 2, "f64", ga.PageRankConfig(0.85, 300, tol), p=p)
    check(b, ref, tol)


def test_small_and_empty(oracle, ctx_b):
    """Graphs smaller than one bin, with no edges, and a single node."""
    for n, pairs in ((3, [(0, 1), (1, 2), (2, 0), (0, 2)]), (5, []), (1, [(0, 0)]),
                     (4097, [(i, (i * 7) % 4097) for i in range(4097)])):
        off = np.zeros(n + 1, np.uint64)
        for u, _ in pairs:
            off[u + 1] += 1
        off = np.cumsum(off).astype(np.uint64)
        nb = np.array([v for _, v in sorted(pairs)], dtype=np.uint32)
        dg = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx_b)
        r = ga.pagerank_run(dg, ga.PageRankConfig(0.85, 200, 1e-12), ctx=ctx_b)
        ref = oracle.pagerank_run(off, nb, 0.85, 200, 1e-12)
        assert r.converged == ref.converged
        assert np.abs(r.scores - ref.scores).sum() <= 1e-2 * 1e-12 + 1e-12, (n, r.scores)
        dg.close()


def test_default_below_2_21_nodes_is_the_pull(ctx_p):
    """GOPS_DELTA_SWEEP_AUTO (the default) pulls below 2^21 nodes: the same
    bits as a context that chose the pull; unknown values are refused.  (From
    2^21 nodes it takes the blocked sweep: test_gpu_full_size.py's scale-24
    run.)"""
    c = ga.Context(0, precision="mixed")
    e = generators.rmat(14, 16, seed=3)
    cfg = ga.PageRankConfig(0.85, 200, 1e-9)
    a = ga.pagerank_run(ga.DeviceGraph.from_edges(e, ctx=c), cfg, ctx=c)
    pl = ga.pagerank_run(ga.DeviceGraph.from_edges(e, ctx=ctx_p), cfg, ctx=ctx_p)
    assert a.iterations == pl.iterations and np.array_equal(a.scores, pl.scores)
    with pytest.raises(ga.InvalidParameter):
        c.set_option(L.OPT_DELTA_SWEEP, 7)


def test_nonfinite_weight_carries_through(oracle, ctx_b, ctx_p):
    """An unchecked weighted run with one +inf edge weight: the reference's
    share (d*x_u)*(w/ws_u) is inf/inf = NaN there (pagerank.rs:255-259) and
    the NaN spreads along the edges.  The blocked sweep's run sums carry the
    non-finite values through its integer row sums, so its NaNs land where the
    pull's and the oracle's do."""
    e0 = generators.rmat(14, 16, seed=21, weights="uniform", p0=0.1, p1=1.0, rounding="f32")
    w = e0.weights.copy()
    w[int(e0.offsets[7])] = np.inf  # the first out-edge of node 7
    e = SimpleNamespace(n=e0.n, m=e0.m, offsets=e0.offsets, targets=e0.targets, weights=w)
    cfg = ga.PageRankConfig(0.85, 40, 1e-10)
    got = {}
    for name, cx in (("blocked", ctx_b), ("pull", ctx_p)):
        dg = ga.DeviceGraph.from_edges(e, weight_storage="f32", ctx=cx)
        got[name] = ga.pagerank_weighted_run(dg, cfg, ctx=cx)
        if name == "blocked":
            assert cx.last_stats()["sweeps_blocked"] > 0
        dg.close()
    ref = oracle.pagerank_weighted_run(e.offsets, e.targets, w, 0.85, 40, 1e-10)
    b, pl = got["blocked"], got["pull"]
    assert np.isnan(ref.scores).any() and not ref.converged
    for r in (b, pl):
        assert not r.converged and r.iterations == 40 and np.isnan(r.diff_l1)
        assert np.array_equal(np.isnan(r.scores), np.isnan(ref.scores))
    fin = ~np.isnan(ref.scores)
    assert np.abs(b.scores[fin] - pl.scores[fin]).sum() <= 1e-9


@pytest.fixture(scope="module")
def ctx_bp():
    c = make_ctx(L.DELTA_SWEEP_BLOCKED)
    c.set_option(L.OPT_ANCHOR_SWEEP, L.ANCHOR_SWEEP_PULL)
    return c


def test_blocked_anchors_match_pull_anchors(oracle, ctx_b, ctx_bp):
    """Blocked f64 anchors (the default with blocked delta sweeps) against
    pull anchors: every sweep of the first context runs blocked, the
    anchors' row sums differ from the pull's only by f64 rounding, so the
    iterations are the same, the residuals agree as closely as f64 rounding
    of the iterate lets a residual near 1e-12 agree (1e-4 of it, or 1e-15)
    and the scores to 1e-13 in L1; two runs give the same bits.  The graph has split bins and
    ragged chunks (n = 70 001, a row of 140 000 in-edges)."""
    n = 70001
    rng = np.random.default_rng(12)
    src = np.concatenate([np.arange(n, dtype=np.int64) % 65000, rng.integers(0, 65000, 140000)])
    dst = np.concatenate([rng.integers(0, 60000, n), np.full(140000, 7)])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    w = bf16_exact(np.exp(rng.normal(0.0, 1.0, dst.size)))  # stored as bf16 without rounding
    e = SimpleNamespace(n=n, m=dst.size, offsets=off, targets=dst.astype(np.uint32), weights=w)
    tol = 1e-12
    cfg = ga.PageRankConfig(0.85, 300, tol)
    out = []
    for cx in (ctx_b, ctx_bp, ctx_b):
        dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16", ctx=cx)
        out.append((ga.pagerank_weighted_run(dg, cfg, ctx=cx), cx.last_stats()))
        dg.close()
    (b, sb), (pl, sp), (b2, _) = out
    assert sb["sweeps_fp64"] > 0 and sb["sweeps_blocked"] == b.iterations
    assert sp["sweeps_blocked"] == sp["sweeps_fp32"]
    assert np.array_equal(b.scores, b2.scores) and b.diff_l1 == b2.diff_l1
    assert b.iterations == pl.iterations and b.converged and pl.converged
    assert abs(b.diff_l1 - pl.diff_l1) <= 1e-4 * pl.diff_l1 + 1e-15
    assert np.abs(b.scores - pl.scores).sum() <= 1e-13
    check(b, oracle_pull(oracle, e, 1, tol, 300), tol)


def bf16_exact(w):
    """w with its f32 mantissa cut to bf16's 8 bits: values that bf16 holds
    exactly, whatever rounding the storage conversion applies."""
    u = np.asarray(w, dtype=np.float32).view(np.uint32) & np.uint32(0xFFFF0000)
    return u.view(np.float32).astype(np.float64)


def bin_runs(e, chunk=16384, bin_rows=4096):
    """Runs (a row's in-edges from one chunk of sources) per bin of rows, in
    the library's ids: nodes renumbered by in-degree, descending, ties by id."""
    src = np.repeat(np.arange(e.n, dtype=np.int64), np.diff(e.offsets.astype(np.int64)))
    dst = e.targets.astype(np.int64)
    indeg = np.bincount(dst, minlength=e.n)
    order = np.lexsort((np.arange(e.n), -indeg))
    rank = np.empty(e.n, np.int64)
    rank[order] = np.arange(e.n)
    pairs = np.unique(rank[dst] * (e.n // chunk + 1) + rank[src] // chunk)
    rows = pairs // (e.n // chunk + 1)
    return np.bincount(rows // bin_rows)


@pytest.mark.parametrize("bin_cap", [0, 65536])
def test_split_bins(oracle, ctx_b, ctx_p, bin_cap):
    """With bins of 4096 rows each (GOPS_OPT_BLOCKED_BIN_RUNS = 0), bins whose
    runs exceed one accumulate item (65 536 runs) are split over several
    workgroups whose partial sums the combine kernel adds, 64-bit in delta
    sweeps and 128-bit in anchors.  n = 2^19 + 3 (33 chunks): 4096 rows take an
    in-edge from 33 sources spread over the chunks, which puts 87 203 runs in
    the first bin (counted here in the library's ids); the tolerance makes the
    run take several anchors.  With a cap of 65 536 runs the layout cuts that
    bin short by its rows' run bounds, so no bin is split.  Against the
    oracle, the pull, and pull anchors."""
    n = (1 << 19) + 3
    rng = np.random.default_rng(19)
    hot = np.arange(4096)
    src = [(hot[:, None] + 16384 * np.arange(33)[None, :] + 7).ravel() % n,
           rng.integers(0, n, size=3 * n)]
    dst = [np.repeat(hot, 33), rng.integers(4096, n, size=3 * n)]
    src, dst = np.concatenate(src), np.concatenate(dst)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    w = bf16_exact(rng.uniform(0.1, 1.0, size=dst.size))
    e = SimpleNamespace(n=n, m=dst.size, offsets=off, targets=dst.astype(np.uint32), weights=w)
    runs_per_bin = bin_runs(e)
    assert runs_per_bin.max() > 65536  # at least two accumulate items in one bin
    tol = 1e-12
    ref = oracle_pull(oracle, e, 1, tol, 300)
    ctx_b.set_option(L.OPT_BLOCKED_BIN_RUNS, bin_cap)
    try:
        b = runs(ctx_b, ctx_p, e, 1, "bf16", ga.PageRankConfig(0.85, 300, tol))
        st = ctx_b.last_stats()
        dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16", ctx=ctx_b)
        ctx_b.set_option(L.OPT_ANCHOR_SWEEP, L.ANCHOR_SWEEP_PULL)
        try:
            pa = ga.pagerank_weighted_run(dg, ga.PageRankConfig(0.85, 300, tol), ctx=ctx_b)
        finally:
            ctx_b.set_option(L.OPT_ANCHOR_SWEEP, L.ANCHOR_SWEEP_AUTO)
        dg.close()
    finally:
        ctx_b.set_option(L.OPT_BLOCKED_BIN_RUNS, 131072)
    check(b, ref, tol)
    assert st["sweeps_fp64"] >= 3
    if bin_cap == 0:
        assert st["blocked_bins"] == (n + 4095) // 4096
        assert st["blocked_split_bins"] >= 1
    else:  # the first bin ends early; the rest are 4096 rows each up to a shorter tail
        assert st["blocked_bins"] >= (n + 4095) // 4096
        assert st["blocked_split_bins"] == 0
    assert pa.iterations == b.iterations
    assert np.abs(pa.scores - b.scores).sum() <= 1e-13


def test_sweep_kinds_with_f64_phase_and_timing(oracle, ctx_b):
    """With every sweep blocked the host launches only the kernels of each
    sweep's kind (f32 delta or f64 anchor), which it knows from the mode the
    previous sweep's finish wrote; with a leading all-f64 phase
    (GOPS_OPT_F32_BELOW > 0) an anchor may follow an anchor, so both kinds
    are launched while the phase lasts.  Both schedules, and the default with
    every sweep timed, match the oracle; the timed run gives the bits of the
    untimed one."""
    e = generators.rmat(16, 16, seed=21, weights="lognormal", p0=0.0, p1=1.0, rounding="bf16")
    tol = 1e-12
    cfg = ga.PageRankConfig(0.85, 300, tol)
    ref = oracle_pull(oracle, e, 1, tol, 300)
    dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16", ctx=ctx_b)
    try:
        base = ga.pagerank_weighted_run(dg, cfg, ctx=ctx_b)
        s0 = ctx_b.last_stats()
        check(base, ref, tol)
        assert s0["sweeps_blocked"] == base.iterations and s0["sweeps_fp64"] >= 2
        ctx_b.set_option(L.OPT_TIMING, 1)
        try:
            timed = ga.pagerank_weighted_run(dg, cfg, ctx=ctx_b)
        finally:
            ctx_b.set_option(L.OPT_TIMING, 0)
        assert np.array_equal(timed.scores, base.scores) and timed.diff_l1 == base.diff_l1
        ctx_b.set_option(L.OPT_F32_BELOW, 1e6)
        try:
            ph = ga.pagerank_weighted_run(dg, cfg, ctx=ctx_b)
            s1 = ctx_b.last_stats()
        finally:
            ctx_b.set_option(L.OPT_F32_BELOW, 0.0)
        check(ph, ref, tol)
        assert s1["sweeps_blocked"] == ph.iterations
        assert s1["sweeps_fp64"] > s0["sweeps_fp64"] + 2  # the leading phase ran
    finally:
        dg.close()
