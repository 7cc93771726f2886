"""gops_result_top_k_noalloc: top_k of a run's resident result with the
scratch in the library's file-scope device array (no allocation), against
gops_result_top_k, the same selection with allocated scratch: equal ids and
scores bit for bit (topk.rs:17-42 semantics, ties included), on two calls in
a row, the second on a smaller input; and against the oracle's top_k."""
import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


def test_noalloc_top_k_matches_allocating_call(oracle):
    ctx = ga.Context(0)
    cfg = ga.PageRankConfig(0.85, 200, 1e-10)
    for scale, k in ((15, 1000), (12, 37)):  # the second input is smaller
        e = generators.rmat(scale, 16, seed=scale, weights="uniform", p0=0.1, p1=1.0,
                            rounding="f32")
        dg = ga.DeviceGraph.from_edges(e, weight_storage="f32", ctx=ctx)
        scores = ga.pagerank_weighted_run(dg, cfg, ctx=ctx).scores
        got = ga.result_top_k_noalloc(dg, k)
        ref = ga.result_top_k(dg, k)
        assert len(got) == k and got == ref
        assert got == oracle.top_k(scores, k)
        # k beyond the scratch, and k = 0
        with pytest.raises(ga.InvalidParameter, match="at most 4096"):
            ga.result_top_k_noalloc(dg, 4097)
        assert ga.result_top_k_noalloc(dg, 0) == []
        dg.close()
    # ties: every score equal, so the kept entries are fixed by the reference's heap
    n = 5000
    off = np.arange(n + 1, dtype=np.uint64)
    nb = ((np.arange(n) + 1) % n).astype(np.uint32)  # one cycle: every score 1/n
    dg = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx)
    scores = ga.pagerank_run(dg, cfg, ctx=ctx).scores
    assert ga.result_top_k_noalloc(dg, 100) == ga.result_top_k(dg, 100) == oracle.top_k(scores, 100)
    dg.close()
    ctx.close()
