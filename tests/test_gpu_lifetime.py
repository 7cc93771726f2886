"""Handles released in any order (gops_context_destroy before the graphs and
servers made on it, as Python's cycle collector may do): the context stays
until the last of them is destroyed, and every object still works until its
own handle is released."""
import gc

import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


def test_context_released_before_graph_and_server():
    e = generators.rmat(10, 8, seed=1)
    ctx = ga.Context(0)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    srv = ga.Server(dg, tile=8, max_slots=8)
    ref = ga.pagerank_run(dg, ga.PageRankConfig(0.85, 100, 1e-9), ctx=ctx).scores
    ctx.close()  # the caller's context handle first
    dg_h = dg.handle
    assert dg_h is not None
    srv.submit(ga.PageRankConfig(0.85, 200, 1e-8), [1, 2])
    got = None
    while got is None:
        srv.step(4)
        got = srv.poll()
    assert got[1].converged
    dg.close()   # then the graph: the server still holds it
    srv.close()  # the last reference: graph and context are freed
    assert np.isfinite(ref).all()


def test_cycle_collected_together():
    """A context and its graph in one reference cycle: the collector runs
    their finalizers in an order of its own."""
    for _ in range(3):
        ctx = ga.Context(0)
        dg = ga.DeviceGraph.from_lists(3, np.array([0, 1, 2, 3], np.uint64),
                                       np.array([1, 2, 0], np.uint32), ctx=ctx)
        holder = {"ctx": ctx, "dg": dg}
        holder["self"] = holder
        ctx.cycle = holder
        del ctx, dg, holder
        gc.collect()
