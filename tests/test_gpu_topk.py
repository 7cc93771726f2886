"""GPU parity of top_k (topk.rs:17-42): gops_top_k and gops_result_top_k
against the oracle's heap restatement, bit for bit (ids and scores): the
same entries at ties on the k-th score, equal scores by ascending id."""
import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


def tie_heavy(rng, n):
    return rng.choice([0.25, 0.5, 0.5, 1.0, 2.0, 3.0, 0.0, -1.0, np.nan, np.inf, -np.inf, 5e-324],
                      size=n)


@pytest.mark.parametrize("n", [1, 7, 100, 5000, 1 << 20])
def test_random_and_ties(oracle, ctx, n):
    rng = np.random.default_rng(n)
    for kind in ("distinct", "ties"):
        s = rng.random(n) * 2 - 0.5 if kind == "distinct" else tie_heavy(rng, n)
        for k in sorted({0, 1, 7, 100, 4096, 4097, 10000, n, n + 5}):
            got = ga.top_k(s, k, ctx=ctx)
            assert got == oracle.top_k(s, k), (n, kind, k)


def test_all_equal_and_all_invalid(oracle, ctx):
    s = np.full(300000, 0.125)
    for k in (1, 5, 4096, 5000):
        assert ga.top_k(s, k, ctx=ctx) == oracle.top_k(s, k)
    bad = np.array([np.nan, 0.0, -1.0, np.inf] * 1000)
    assert ga.top_k(bad, 10, ctx=ctx) == []
    assert ga.top_k(np.zeros(0), 10, ctx=ctx) == []
    # doc example, topk.rs:9-16
    assert ga.top_k([0.1, 0.5, 0.3, 0.9, 0.2], 2, ctx=ctx) == [(3, 0.9), (1, 0.5)]


def test_result_top_k_after_runs(oracle, ctx):
    """The result of a run stays on the device; its top_k equals the
    oracle's top_k of the downloaded scores."""
    e = generators.rmat(16, 16, seed=7)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    with pytest.raises(ga.InvalidParameter, match="no PageRank result"):
        ga.result_top_k(dg, 10)
    cfg = ga.PageRankConfig(0.85, 100, 1e-9)
    r = ga.pagerank_run(dg, cfg, ctx=ctx)
    for k in (1, 100, 4096, 10000):
        assert ga.result_top_k(dg, k) == oracle.top_k(r.scores, k)
    p = np.zeros(e.n)
    p[[5, 9, 1000]] = 1.0
    r = ga.personalized_pagerank_run(dg, cfg, p, ctx=ctx)
    assert ga.result_top_k(dg, 50) == oracle.top_k(r.scores, 50)


def test_stop_at_a_bucket_bound_with_ties(oracle, ctx):
    """The radix passes stop once the bucket of the k-th largest holds
    exactly the keys still to take; T is then the bucket's lower bound, and
    keys equal to it are ties that all enter.  1.0 is such a bound (its low
    bits are zero): 60 scores of 3.0, then 10 of exactly 1.0 and 30 in
    (1.0, 1.0625) -- one 16-bit bucket of 40 keys, all needed for k = 100 --
    shuffled among 5000 smaller scores; also with the bucket one key short
    and one key over, so that the passes do not stop early."""
    rng = np.random.default_rng(17)
    for extra in (30, 29, 31):
        s = np.concatenate([np.full(60, 3.0), np.full(10, 1.0),
                            1.0 + rng.uniform(1e-9, 0.0625, size=extra),
                            rng.uniform(0.01, 0.5, size=5000)])
        s = s[rng.permutation(s.size)]
        for k in (100, 70, 65, 101):
            assert ga.top_k(s, k, ctx=ctx) == oracle.top_k(s, k), (extra, k)
