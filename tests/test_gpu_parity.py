"""GPU parity: the HIP path (through the C ABI) against the f64 oracle and the
golden fixtures.

Tolerances (floating point; the reference is f64 throughout):
  * precision "fp64": every sweep in f64 -> iterations equal the oracle's,
    scores within 1e-12 (L1) of the oracle.
  * precision "mixed" (default): f32 delta sweeps with f64 anchors ->
    iterations equal the oracle's (unless the oracle's last residuals sit
    within 1e-4 relative of the tolerance), L1(scores - oracle) <=
    1e-2 * tolerance + 1e-12.
"""
import os
from fractions import Fraction
from types import SimpleNamespace

import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu

D = 0.85


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0, precision="mixed")


@pytest.fixture(scope="module")
def ctx64():
    return ga.Context(0, precision="fp64")


def frac(s):
    return float(Fraction(s))


def score_tol(ctxp, tol):
    return 1e-12 if ctxp == "fp64" else 1e-2 * tol + 1e-12


def same_iterations(got, ref, oracle_diffs=None):
    return got == ref


def ref_noise(n):
    """Absolute noise of the reference's own f64 residual: its sequential
    dangling sum (pagerank.rs:144-149) over ~n/2 nodes leaves the iterate's
    total mass off by O(n * eps64 * dsum) every sweep (measured ~1e-12 at
    n = 2^20), and sum |x_old - x_new| carries that noise."""
    return 1e-15 + 4e-18 * n


def check_run(run, ref, tol, prec, name=""):
    l1 = np.abs(run.scores - ref.scores).sum()
    if ref.converged or prec == "fp64":
        assert l1 <= score_tol(prec, tol), (name, prec, l1)
    else:
        # an unconverged mixed run ends with its last segment's f32 storage
        # rounding (relative ~6e-8 per sweep) not yet re-anchored
        assert l1 <= 1e-7, (name, prec, l1)
    assert run.converged == ref.converged, (name, prec)
    noise = ref_noise(ref.scores.size)
    if run.iterations != ref.iterations:
        # only a residual that lands within 1e-4 (relative) of the tolerance,
        # or within the reference's own noise of it, may flip the stopping sweep
        assert abs(run.iterations - ref.iterations) == 1, (name, run.iterations, ref.iterations)
        assert (abs(ref.diff_l1 - tol) <= 1e-4 * tol + noise
                or abs(run.diff_l1 - tol) <= 1e-4 * tol + noise)
    else:
        assert abs(run.diff_l1 - ref.diff_l1) <= 1e-4 * ref.diff_l1 + noise, (name, prec)


def test_closed_forms(golden, ctx, ctx64):
    for c in golden["closed_forms"]:
        x = np.array([frac(s) for s in c["fixed_point"]])
        for cx in (ctx, ctx64):
            g = ga.AdjacencyMatrix(c["matrix"])
            cfg = ga.PageRankConfig(0.85, 1000, 1e-12)
            if c["personalization"] is not None:
                r = ga.personalized_pagerank_run(g, cfg, c["personalization"], ctx=cx)
            elif c["weighted"]:
                r = ga.pagerank_weighted_run(g, cfg, ctx=cx)
            else:
                r = ga.pagerank_run(g, cfg, ctx=cx)
            assert r.converged, c["name"]
            assert np.abs(r.scores - x).sum() <= (D / (1 - D)) * 1e-12 + 1e-13, c["name"]


def test_exact_trajectories(golden, ctx, ctx64):
    for t in golden["trajectories"]:
        g = ga.AdjacencyMatrix(t["matrix"])
        cfg = ga.PageRankConfig(t["damping"], t["max_iterations"], t["tolerance"])
        for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
            if t["personalization"] is not None:
                r = ga.personalized_pagerank_run(g, cfg, t["personalization"], ctx=cx)
            elif t["weighted"]:
                r = ga.pagerank_weighted_run(g, cfg, ctx=cx)
            else:
                r = ga.pagerank_run(g, cfg, ctx=cx)
            assert r.iterations == t["iterations"], (t["name"], prec)
            assert r.converged == t["converged"], (t["name"], prec)
            rel = 1e-9 if prec == "fp64" else 1e-4
            assert abs(r.diff_l1 - t["diff_l1"]) <= rel * t["diff_l1"] + 1e-15, (t["name"], prec,
                                                                              r.diff_l1)
            l1 = np.abs(r.scores - np.array(t["scores"])).sum()
            bound = score_tol(prec, t["tolerance"]) if (t["converged"] or prec == "fp64") else 1e-7
            assert l1 <= bound, (t["name"], prec, l1)


def lists_from(nbrs):
    off = np.zeros(len(nbrs) + 1, dtype=np.uint64)
    for u, lst in enumerate(nbrs):
        off[u + 1] = off[u] + len(lst)
    return off, np.array([v for lst in nbrs for v in lst], dtype=np.uint32)


def test_dense_and_networkx_cases(golden, oracle, ctx, ctx64):
    for c in golden["dense_solves"] + golden["networkx"]:
        off, nb = lists_from(c["neighbors"])
        ew = (np.array([w for lst in c["edge_weights"] for w in lst])
              if c["edge_weights"] is not None else None)
        tol = 1e-10
        for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
            dg = ga.DeviceGraph.from_lists(c["n"], off, nb, ew, ctx=cx)
            cfg = ga.PageRankConfig(0.85, 500, tol)
            if c["kind"] == "ppr":
                r = ga.personalized_pagerank_run(dg, cfg, c["personalization"], ctx=cx)
                ref = oracle.personalized_pagerank_run(off, nb, c["personalization"], 0.85, 500,
                                                       tol)
            elif c["kind"] == "weighted":
                r = ga.pagerank_weighted_run(dg, cfg, ctx=cx)
                ref = oracle.pagerank_weighted_run(off, nb, ew, 0.85, 500, tol)
            else:
                r = ga.pagerank_run(dg, cfg, ctx=cx)
                ref = oracle.pagerank_run(off, nb, 0.85, 500, tol)
            check_run(r, ref, tol, prec, c["name"])
            x = np.array(c["fixed_point"])
            assert np.abs(r.scores - x).sum() <= (D / (1 - D)) * tol + 1e-11


def test_c1_dense_workload(golden, oracle, ctx, ctx64):
    """C1: G(2000, 0.005) seed 42 as a dense AdjacencyMatrix, tol 1e-10."""
    e = generators.erdos_renyi(2000, 0.005, 42)
    mat = np.zeros((2000, 2000))
    for u in range(2000):
        mat[u, e.targets[e.offsets[u]:e.offsets[u + 1]]] = 1.0
    off, nb, _ = oracle.dense_neighbors(mat)
    ref = oracle.pagerank_run(off, nb, 0.85, 100, 1e-10)
    x = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_er2000_fixed_point.npy"))
    for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
        r = ga.pagerank_run(ga.AdjacencyMatrix(mat), ga.PageRankConfig(0.85, 100, 1e-10), ctx=cx)
        check_run(r, ref, 1e-10, prec, "c1")
        assert np.abs(r.scores - x).sum() <= (D / (1 - D)) * 1e-10 + 1e-12


def rmat_case(scale, seed, weights, rounding, ef=16):
    if weights == "unit":
        return generators.rmat(scale, ef, seed=seed)
    if weights == "uniform":
        return generators.rmat(scale, ef, seed=seed, weights="uniform", p0=0.1, p1=1.0,
                               rounding=rounding)
    return generators.rmat(scale, ef, seed=seed, weights="lognormal", p0=0.0, p1=1.0,
                           rounding=rounding)


def oracle_pull(oracle, e, kind, tol, max_it, p=None, damping=0.85):
    in_off, in_src, in_w = oracle.transpose(e.offsets, e.targets, e.weights if kind == 1 else None)
    if kind == 1:
        scale = oracle.out_wsum(e.offsets, e.weights)
    else:
        scale = np.diff(e.offsets).astype(np.float64)
    return oracle.pull_run(kind, in_off, in_src, scale, w=in_w, p=p, damping=damping,
                           max_iterations=max_it, tolerance=tol, threads=0)


@pytest.mark.parametrize("weights,storage", [("unit", "none"), ("uniform", "f32"),
                                             ("lognormal", "bf16"), ("uniform", "f64")])
def test_rmat_scale16(oracle, ctx, ctx64, weights, storage):
    rounding = {"none": "f64", "f32": "f32", "bf16": "bf16", "f64": "f64"}[storage]
    e = rmat_case(16, 7, weights, rounding)
    kind = 0 if weights == "unit" else 1
    tol = 1e-9
    ref = oracle_pull(oracle, e, kind, tol, 200)
    for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
        dg = ga.DeviceGraph.from_edges(e, weight_storage=storage if kind else "f64", ctx=cx)
        cfg = ga.PageRankConfig(0.85, 200, tol)
        r = (ga.pagerank_weighted_run if kind else ga.pagerank_run)(dg, cfg, ctx=cx)
        check_run(r, ref, tol, prec, f"rmat16-{weights}")
        assert abs(r.scores.sum() - 1.0) < 1e-9 and r.scores.min() >= 0.0


def test_c2_scale20_fp32_weights(oracle, ctx):
    """C2: R-MAT scale 20, ef 16, seed 7, weights U[0.1, 1) stored f32,
    d 0.85, tol 1e-9, max 200 -- against the f64 oracle on the same weights."""
    e = generators.rmat(20, 16, seed=7, weights="uniform", p0=0.1, p1=1.0, rounding="f32")
    ref = oracle_pull(oracle, e, 1, 1e-9, 200)
    dg = ga.DeviceGraph.from_edges(e, weight_storage="f32", ctx=ctx)
    r = ga.pagerank_weighted_run(dg, ga.PageRankConfig(0.85, 200, 1e-9), ctx=ctx)
    check_run(r, ref, 1e-9, "mixed", "c2")
    assert np.abs(r.scores - ref.scores).sum() <= 1e-6
    top_ref = np.argsort(-ref.scores, kind="stable")[:100]
    top = np.argsort(-r.scores, kind="stable")[:100]
    assert len(set(top_ref.tolist()) & set(top.tolist())) >= 99


def test_ppr_against_oracle(oracle, ctx, ctx64):
    e = generators.rmat(14, 16, seed=13)
    n = e.n
    rng = np.random.default_rng(3)
    for trial in range(3):
        p = np.zeros(n)
        seeds = rng.choice(n, size=int(rng.integers(1, 65)), replace=False)
        p[seeds] = 1.0 if trial < 2 else rng.random(seeds.size)
        tol = 1e-8
        ref = oracle_pull(oracle, e, 2, tol, 300, p=p)
        for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
            dg = ga.DeviceGraph.from_edges(e, ctx=cx)
            r = ga.personalized_pagerank_run(dg, ga.PageRankConfig(0.85, 300, tol), p, ctx=cx)
            check_run(r, ref, tol, prec, f"ppr{trial}")


def test_petgraph_adapter_semantics(oracle, ctx):
    """DiGraph (petgraph semantics): parallel edges count twice; weighted run
    reads the find_edge weight for every parallel entry."""
    g = ga.DiGraph(4)
    for u, v, w in [(0, 1, 1.0), (0, 1, 3.0), (0, 2, 2.0), (1, 2, 1.0), (2, 0, 1.0)]:
        g.add_edge(u, v, w)
    off, nb, ew = g._lists(True)
    cfg = ga.PageRankConfig(0.85, 500, 1e-12)
    r = ga.pagerank_run(g, cfg, ctx=ctx)
    ref = oracle.pagerank_run(off, nb, 0.85, 500, 1e-12)
    check_run(r, ref, 1e-12, "mixed", "petgraph")
    r = ga.pagerank_weighted_run(g, cfg, ctx=ctx)
    ref = oracle.pagerank_weighted_run(off, nb, ew, 0.85, 500, 1e-12)
    check_run(r, ref, 1e-12, "mixed", "petgraph-w")
    # reference integration tests: tests/invariants.rs:20-50
    line = ga.DiGraph(3)
    line.add_edge(0, 1, 1.0)
    line.add_edge(1, 2, 1.0)
    for s in (ga.pagerank(line, ctx=ctx),
              ga.personalized_pagerank(line, ga.PageRankConfig(), [1.0, 0.0, 0.0], ctx=ctx)):
        assert s.size == 3 and np.all(np.isfinite(s)) and np.all(s >= 0)
        assert abs(s.sum() - 1.0) <= 1e-6


def test_edge_cases(ctx):
    empty = ga.DeviceGraph.from_lists(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32), ctx=ctx)
    r = ga.pagerank_run(empty, ctx=ctx)
    assert (r.scores.size, r.iterations, r.diff_l1, r.converged) == (0, 0, 0.0, True)
    g = ga.AdjacencyMatrix([[0, 1, 1, 0], [0, 0, 1, 0], [1, 0, 0, 0], [0, 0, 1, 0]])
    r = ga.pagerank_run(g, ga.PageRankConfig(0.85, 0, 1e-6), ctx=ctx)
    assert r.iterations == 0 and r.diff_l1 == float("inf") and not r.converged
    np.testing.assert_array_equal(r.scores, np.full(4, 0.25))
    r = ga.pagerank_run(g, ga.PageRankConfig(0.0, 100, 1e-6), ctx=ctx)
    assert r.iterations == 1 and r.converged and r.diff_l1 == 0.0
    np.testing.assert_allclose(r.scores, 0.25, atol=1e-15)
    r = ga.pagerank_run(g, ga.PageRankConfig(0.85, 7, 1e-15), ctx=ctx)
    assert r.iterations == 7 and not r.converged
    # all-zero personalization falls back to uniform (ppr.rs:105-109)
    a = ga.personalized_pagerank(g, ga.PageRankConfig(0.85, 100, 1e-10), np.zeros(4), ctx=ctx)
    b = ga.pagerank(g, ga.PageRankConfig(0.85, 100, 1e-10), ctx=ctx)
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_checked_errors(ctx):
    g = ga.AdjacencyMatrix([[0.0, 1.0], [0.0, 0.0]])
    with pytest.raises(ga.InvalidParameter, match=r"invalid parameter: damping must be in \[0,1\]"):
        ga.pagerank_checked(g, ga.PageRankConfig(1.5, 100, 1e-6), ctx=ctx)
    with pytest.raises(ga.InvalidParameter, match="max_iterations must be > 0"):
        ga.pagerank_checked_run(g, ga.PageRankConfig(0.85, 0, 1e-6), ctx=ctx)
    with pytest.raises(ga.InvalidParameter, match="tolerance must be finite and > 0"):
        ga.pagerank_weighted_checked(g, ga.PageRankConfig(0.85, 10, -1.0), ctx=ctx)
    # ppr.rs:163-181
    with pytest.raises(ga.InvalidParameter, match="personalization length"):
        ga.personalized_pagerank_checked(g, ga.PageRankConfig(), [1.0], ctx=ctx)
    with pytest.raises(ga.InvalidParameter, match="sum"):
        ga.personalized_pagerank_checked(g, ga.PageRankConfig(), [0.0, 0.0], ctx=ctx)
    with pytest.raises(ga.InvalidParameter, match="non-negative"):
        ga.personalized_pagerank_checked(g, ga.PageRankConfig(), [1.0, -1.0], ctx=ctx)
    # pagerank.rs:397-424: a -1 weight through a custom adapter is rejected
    bad = ga.DiGraph(3)
    bad.add_edge(0, 1, -1.0)
    with pytest.raises(ga.InvalidParameter, match="edge weights must be non-negative"):
        ga.pagerank_weighted_checked(bad, ga.PageRankConfig(), ctx=ctx)
    nan = ga.DiGraph(3)
    nan.add_edge(0, 1, float("nan"))
    nan.add_edge(1, 2, -1.0)
    with pytest.raises(ga.InvalidParameter, match="edge weights must be finite"):
        ga.pagerank_weighted_checked(nan, ga.PageRankConfig(), ctx=ctx)
    # unchecked weighted run clamps instead (pagerank.rs:257)
    s = ga.pagerank_weighted(bad, ga.PageRankConfig(), ctx=ctx)
    np.testing.assert_allclose(s, 1.0 / 3.0, atol=1e-12)
    with pytest.raises(ga.DimensionMismatch):
        ga.personalized_pagerank(g, ga.PageRankConfig(), [1.0], ctx=ctx)
    with pytest.raises(ga.IndexOutOfBounds):
        ga.DeviceGraph.from_lists(2, np.array([0, 1, 1], np.uint64), np.array([5], np.uint32),
                                  ctx=ctx)
    unweighted = ga.DeviceGraph.from_lists(2, np.array([0, 1, 1], np.uint64),
                                           np.array([1], np.uint32), ctx=ctx)
    with pytest.raises(ga.InvalidParameter, match="without edge weights"):
        ga.pagerank_weighted(unweighted, ctx=ctx)


def test_ppr_batch_matches_single(ctx):
    e = generators.rmat(13, 16, seed=13)
    off, seeds = generators.seed_sets(e, 12, 1, 16, seed=9)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 300, 1e-8) for q in range(12)]
    scores, runs = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx)
    for q in range(12):
        p = np.zeros(e.n)
        p[seeds[off[q]:off[q + 1]]] = 1.0
        r = ga.personalized_pagerank_run(dg, cfgs[q], p, ctx=ctx)
        assert runs[q].iterations == r.iterations and runs[q].converged == r.converged
        assert np.abs(scores[q] - r.scores).sum() <= 1e-2 * 1e-8 + 1e-12


def test_unaligned_runs_and_hub_chunks(oracle, ctx, ctx64):
    """Rows whose edges start off a 64-edge boundary, where the pull's waves
    begin: two hub rows split into chunks (the second starting 48 edges past
    a boundary), a long row with its own descriptor (1 past), a row of 4090
    in-edges 9 past a boundary (its offset and edges overflow one pass of
    the stage's loads, so its run takes a second pass), a row of 4000, and
    80 000 short rows at every offset; unweighted and f32-weighted, mixed and
    f64, against the oracle."""
    n = 80000
    rng = np.random.default_rng(21)
    src = [np.arange(1, 70001), np.arange(2, 66003), np.arange(3, 5003), np.arange(4, 4094),
           np.arange(5, 4005), rng.integers(0, n, size=2 * n)]
    dst = [np.full(70000, 0), np.full(66001, 1), np.full(5000, 2), np.full(4090, 3),
           np.full(4000, 4), rng.integers(5, n, size=2 * n)]
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    w = rng.uniform(0.1, 1.0, size=dst.size).astype(np.float32).astype(np.float64)
    e = SimpleNamespace(n=n, m=dst.size, offsets=off, targets=dst.astype(np.uint32), weights=w)
    in_deg = np.bincount(dst, minlength=n)
    assert in_deg[0] == 70000 and in_deg[1] == 66001 and in_deg[3] == 4090
    tol = 1e-10
    for kind, storage in ((0, "f64"), (1, "f32")):
        ref = oracle_pull(oracle, e, kind, tol, 300)
        for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
            dg = ga.DeviceGraph.from_edges(e, weight_storage=storage, ctx=cx)
            cfg = ga.PageRankConfig(0.85, 300, tol)
            r = (ga.pagerank_weighted_run if kind else ga.pagerank_run)(dg, cfg, ctx=cx)
            check_run(r, ref, tol, prec, f"unaligned-{kind}")
            dg.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_normalize_option(oracle, kind):
    """GOPS_OPT_NORMALIZE: the f64 sum-to-one rescale of the result (the
    reference never renormalises, lib.rs:11-12, so it is opt-in).  It must
    leave a sum of 1 to f64 rounding and move the scores by far less than the
    tolerance: within 1e-2 * tol of the same run without it, and within the
    usual bound of the oracle."""
    e = rmat_case(15, 9, "lognormal" if kind == 1 else "unit", "bf16" if kind == 1 else "f64")
    tol = 1e-9
    p = None
    if kind == 2:
        p = np.zeros(e.n)
        p[[3, 40, 777]] = [1.0, 2.0, 0.5]
    ref = oracle_pull(oracle, e, kind, tol, 300, p=p)
    runs = {}
    for norm in (False, True):
        cx = ga.Context(0, precision="mixed", normalize=norm)
        dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16" if kind == 1 else "f64", ctx=cx)
        cfg = ga.PageRankConfig(0.85, 300, tol)
        if kind == 0:
            runs[norm] = ga.pagerank_run(dg, cfg, ctx=cx)
        elif kind == 1:
            runs[norm] = ga.pagerank_weighted_run(dg, cfg, ctx=cx)
        else:
            runs[norm] = ga.personalized_pagerank_run(dg, cfg, p, ctx=cx)
        dg.close()
        cx.close()
    plain, normed = runs[False], runs[True]
    assert abs(normed.scores.sum() - 1.0) <= 64 * np.finfo(float).eps
    assert normed.iterations == plain.iterations and normed.diff_l1 == plain.diff_l1
    assert np.abs(normed.scores - plain.scores).sum() <= 1e-2 * tol
    # the device divides by its own fixed-order f64 sum of the scores
    np.testing.assert_allclose(normed.scores, plain.scores / plain.scores.sum(), rtol=1e-14,
                               atol=0)
    check_run(normed, ref, tol, "mixed", f"normalize-{kind}")


@pytest.mark.parametrize("delta", ["pull", "blocked"])
def test_graph_replay_is_the_queued_run(oracle, delta):
    """GOPS_LAUNCH_GRAPH (the sweep chain captured once as a HIP graph of
    a few sweeps and replayed until the device stops it) gives
    the same bits as GOPS_LAUNCH_QUEUED (kernel by kernel), for plain,
    weighted and personalized runs, with the iteration cap falling inside a
    replay; and replays the captured chain on later runs."""
    from graphops_amd import _lib as L

    e = rmat_case(14, 5, "lognormal", "bf16")
    p = np.zeros(e.n)
    p[[1, 50, 900]] = [1.0, 0.5, 2.0]
    out = {}
    for launch in (L.LAUNCH_QUEUED, L.LAUNCH_GRAPH):
        cx = ga.Context(0)
        cx.set_option(L.OPT_LAUNCH, launch)
        cx.set_option(L.OPT_DELTA_SWEEP, {"pull": L.DELTA_SWEEP_PULL,
                                          "blocked": L.DELTA_SWEEP_BLOCKED}[delta])
        dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16", ctx=cx)
        runs = []
        for rep in range(2):
            for cfg in (ga.PageRankConfig(0.85, 300, 1e-10), ga.PageRankConfig(0.85, 7, 1e-14)):
                runs.append(ga.pagerank_weighted_run(dg, cfg, ctx=cx))
                assert cx.last_stats()["graph_replay"] == (launch == L.LAUNCH_GRAPH)
                runs.append(ga.pagerank_run(dg, cfg, ctx=cx))
                runs.append(ga.personalized_pagerank_run(dg, cfg, p, ctx=cx))
        out[launch] = runs
        dg.close()
        cx.close()
    for a, b in zip(out[L.LAUNCH_QUEUED], out[L.LAUNCH_GRAPH]):
        assert a.iterations == b.iterations and a.converged == b.converged
        assert a.diff_l1 == b.diff_l1 and np.array_equal(a.scores, b.scores)
    capped = out[L.LAUNCH_GRAPH][3]  # the weighted run capped at 7 sweeps
    assert capped.iterations == 7 and not capped.converged
    ref = oracle_pull(oracle, e, 1, 1e-10, 300)
    check_run(out[L.LAUNCH_GRAPH][0], ref, 1e-10, "mixed", f"graph-{delta}")


@pytest.mark.parametrize("plan", [(4096, 65536, 8), (512, 16384, 4), (1024, 32768, 16),
                                  (64, 64, 1)])
def test_row_plans(oracle, ctx, ctx64, plan):
    """Every row plan the autotuner can choose (and the extreme one: a
    workgroup per row of more than 64 in-edges, hubs in chunks of 64, one row
    per lane group) gives the oracle's result, mixed and f64, on the graph
    with hub rows, long rows and unaligned runs."""
    n = 30000
    rng = np.random.default_rng(5)
    src = [np.arange(1, 20001), np.arange(2, 9003), np.arange(3, 1503), rng.integers(0, n, 3 * n)]
    dst = [np.full(20000, 0), np.full(9001, 1), np.full(1500, 2), rng.integers(3, n, 3 * n)]
    src, dst = np.concatenate(src), np.concatenate(dst)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    w = rng.uniform(0.1, 1.0, size=dst.size).astype(np.float32).astype(np.float64)
    e = SimpleNamespace(n=n, m=dst.size, offsets=off, targets=dst.astype(np.uint32), weights=w)
    tol = 1e-10
    ref = oracle_pull(oracle, e, 1, tol, 300)
    for cx, prec in ((ctx, "mixed"), (ctx64, "fp64")):
        dg = ga.DeviceGraph.from_edges(e, weight_storage="f32", ctx=cx)
        ga.graph_set_rows(dg, *plan)
        r = ga.pagerank_weighted_run(dg, ga.PageRankConfig(0.85, 300, tol), ctx=cx)
        check_run(r, ref, tol, prec, f"plan{plan}")
        dg.close()


def test_row_plan_autotune(oracle, ctx):
    """gops_graph_tune_rows on R-MAT scale 16: the chosen plan is one of the
    candidates, no slower than the default as measured, and the runs after
    it still match the oracle; the context's last statistics are the ones
    of the run before the tuning.  Bad cut-offs are refused."""
    e = rmat_case(16, 7, "uniform", "f32")
    tol = 1e-9
    ref = oracle_pull(oracle, e, 1, tol, 200)
    dg = ga.DeviceGraph.from_edges(e, weight_storage="f32", ctx=ctx)
    cfg = ga.PageRankConfig(0.85, 200, tol)
    r0 = ga.pagerank_weighted_run(dg, cfg, ctx=ctx)
    before = ctx.last_stats()
    t = ga.graph_tune_rows(dg)
    assert ctx.last_stats() == before
    assert t["candidates"] == 8
    assert t["row_min"] in (4096, 2048, 1024, 512) and t["hub_chunk"] in (65536, 32768, 16384)
    assert t["row_iters"] in (8, 4, 16)
    assert 0.0 < t["tuned_us"] <= t["default_us"]
    r1 = ga.pagerank_weighted_run(dg, cfg, ctx=ctx)
    check_run(r1, ref, tol, "mixed", "tuned")
    assert r1.iterations == r0.iterations
    assert np.abs(r1.scores - r0.scores).sum() <= 1e-12
    for bad in ((32, 65536, 8), (8192, 65536, 8), (4096, 1024, 8), (4096, 65536, 0)):
        with pytest.raises(ga.InvalidParameter):
            ga.graph_set_rows(dg, *bad)
    dg.close()
