"""GPU parity at the full benchmark sizes, through properties that do not
need the oracle to run at that size (it would take minutes):

  * R-MAT scale 24, bf16 log-normal weights, pagerank_weighted_run at
    tolerance 1e-12: the mixed-precision run against the all-f64 run of the
    same library (every sweep an f64 anchor, which follows the reference's
    f64 recurrence): L1 <= 1e-9 (the accuracy bound is 1e-6), the same top
    100, total mass 1 within 1e-9, no negative score, both converged;
  * R-MAT scale 22, 1024 personalized queries (|S| ~ U{1..64}, d 0.85, tol
    1e-8) in one batch: all converged; eight of them with scores: mass 1
    within 1e-9, and two against the oracle within 1e-2 * tol + 1e-12.
"""
import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


def test_c3_scale24_mixed_against_all_f64():
    e = generators.rmat(24, 16, 0.57, 0.19, 0.19, 11, "lognormal", 0.0, 1.0, "bf16")
    cfg = ga.PageRankConfig(0.85, 200, 1e-12)
    out = {}
    for prec in ("mixed", "fp64"):
        c = ga.Context(0, precision=prec)
        dg = ga.DeviceGraph.from_edges(e, weight_storage="bf16", ctx=c)
        out[prec] = ga.pagerank_weighted_run(dg, cfg, ctx=c)
        dg.close()
    m, f = out["mixed"], out["fp64"]
    assert m.converged and f.converged
    l1 = np.abs(m.scores - f.scores).sum()
    assert l1 <= 1e-9, l1
    assert abs(m.scores.sum() - 1.0) <= 1e-9 and m.scores.min() >= 0.0
    top_m = set(np.argsort(-m.scores, kind="stable")[:100].tolist())
    top_f = set(np.argsort(-f.scores, kind="stable")[:100].tolist())
    assert top_m == top_f


def test_c4_scale22_batch(oracle):
    e = generators.rmat(22, 16, 0.57, 0.19, 0.19, 13)
    c = ga.Context(0)
    dg = ga.DeviceGraph.from_edges(e, ctx=c)
    off, seeds = generators.seed_sets(e, 1024, 1, 64, log_uniform=False, seed=5)
    cfg = ga.PageRankConfig(0.85, 200, 1e-8)
    _, runs = ga.ppr_batch_run(dg, cfg, off, seeds, with_scores=False, ctx=c)
    assert all(r.converged for r in runs)
    sc, sruns = ga.ppr_batch_run(dg, cfg, off[:9], seeds[:off[8]], ctx=c)
    for q in range(8):
        assert sruns[q].iterations == runs[q].iterations
        assert abs(sc[q].sum() - 1.0) <= 1e-9 and sc[q].min() >= 0.0
    in_off, in_src, _ = oracle.transpose(e.offsets, e.targets)
    deg = np.diff(e.offsets.astype(np.int64)).astype(np.float64)
    for q in (0, 1):
        p = np.zeros(e.n)
        p[seeds[off[q]:off[q + 1]]] += 1.0
        ref = oracle.pull_run(2, in_off, in_src, deg, p=p, damping=0.85, max_iterations=200,
                              tolerance=1e-8, threads=16)
        assert ref.converged and abs(ref.iterations - sruns[q].iterations) <= 1
        assert np.abs(sc[q] - ref.scores).sum() <= 1e-2 * 1e-8 + 1e-12
    dg.close()
