"""GPU parity of batched personalized PageRank and the serving loop.

Every query is checked against the f64 oracle's personalized_pagerank_run
(ppr.rs:90-156) on the same graph and personalization:
  * iterations and converged equal the oracle's (a residual landing within
    1e-4 of the tolerance, or within the reference's own f64 noise, may move
    the stop by one sweep);
  * L1(scores - oracle) <= 1e-2 * tolerance + 1e-12 for converged queries.
"""
import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


@pytest.fixture(scope="module")
def graph13():
    e = generators.rmat(13, 16, seed=13)
    return e


def oracle_ppr(oracle, e, p, d, tol, max_it):
    in_off, in_src, _ = oracle.transpose(e.offsets, e.targets)
    scale = np.diff(e.offsets).astype(np.float64)
    return oracle.pull_run(2, in_off, in_src, scale, p=p, damping=d, max_iterations=max_it,
                           tolerance=tol, threads=0)


def dense_p(n, seeds, w=None):
    p = np.zeros(n)
    np.add.at(p, seeds, 1.0 if w is None else w)
    return p


def check(run_it, run_conv, run_diff, scores, ref, tol):
    noise = 1e-15 + 4e-18 * ref.scores.size
    assert run_conv == ref.converged
    if run_it != ref.iterations:
        assert abs(run_it - ref.iterations) == 1
        assert abs(ref.diff_l1 - tol) <= 1e-4 * tol + noise or abs(run_diff - tol) <= 1e-4 * tol + noise
    else:
        assert abs(run_diff - ref.diff_l1) <= 1e-4 * ref.diff_l1 + noise
    if scores is not None:
        l1 = np.abs(scores - ref.scores).sum()
        assert l1 <= (1e-2 * tol + 1e-12 if ref.converged else 1e-7), l1


@pytest.mark.parametrize("tile", [8, 16, 32, 64])
def test_batch_matches_oracle(oracle, ctx, graph13, tile):
    e = graph13
    nq = 20
    off, seeds = generators.seed_sets(e, nq, 1, 64, seed=tile)
    rng = np.random.default_rng(tile)
    w = rng.uniform(0.5, 2.0, size=seeds.size) if tile == 16 else None
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 400, [1e-8, 1e-6, 1e-10][q % 3])
            for q in range(nq)]
    scores, runs = ga.ppr_batch_run(dg, cfgs, off, seeds, seed_weights=w, ctx=ctx, tile=tile)
    for q in range(nq):
        sl = slice(off[q], off[q + 1])
        p = dense_p(e.n, seeds[sl], None if w is None else w[sl])
        ref = oracle_ppr(oracle, e, p, cfgs[q].damping, cfgs[q].tolerance, 400)
        check(runs[q].iterations, runs[q].converged, runs[q].diff_l1, scores[q], ref,
              cfgs[q].tolerance)


def test_batch_f64_phase_schedule(oracle, ctx, graph13):
    """GOPS_OPT_TILE_F32_BELOW = 1e4: each query is admitted with an f64
    anchor sweep and stays in f64 until its residual is within 1e4 of its
    tolerance, then continues with the carried f32 delta sweeps.  Same bar as
    the default schedule."""
    from graphops_amd import _lib as L

    e = graph13
    nq = 24
    off, seeds = generators.seed_sets(e, nq, 1, 64, seed=21)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 400, [1e-8, 1e-6, 1e-10][(q // 3) % 3])
            for q in range(nq)]
    ctx.set_option(L.OPT_TILE_F32_BELOW, 1e4)
    try:
        scores, runs = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx, tile=16)
    finally:
        ctx.set_option(L.OPT_TILE_F32_BELOW, 0.0)
    for q in range(nq):
        p = dense_p(e.n, seeds[off[q]:off[q + 1]])
        ref = oracle_ppr(oracle, e, p, cfgs[q].damping, cfgs[q].tolerance, 400)
        check(runs[q].iterations, runs[q].converged, runs[q].diff_l1, scores[q], ref,
              cfgs[q].tolerance)


def test_carried_rounding_reaches_f64_accuracy(oracle, ctx, graph13):
    """The default tile schedule runs every sweep on f32 deltas.  Each row
    carries its rounding into its next delta, so the scores stay within a few
    1e-15 of the f64 oracle at a tolerance of 1e-12, where dropping the
    roundings would leave an error of order 1e-8 (f32 epsilon times the
    first delta's mass)."""
    e = graph13
    nq = 8
    off, seeds = generators.seed_sets(e, nq, 1, 64, seed=33)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 600, 1e-12) for q in range(nq)]
    scores, runs = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx, tile=8)
    for q in range(nq):
        p = dense_p(e.n, seeds[off[q]:off[q + 1]])
        ref = oracle_ppr(oracle, e, p, cfgs[q].damping, 1e-12, 600)
        assert runs[q].converged and ref.converged
        assert abs(runs[q].iterations - ref.iterations) <= 1
        l1 = np.abs(scores[q] - ref.scores).sum()
        assert l1 <= 1e-13, (q, l1)


def test_batch_refill_same_as_alone(ctx, graph13):
    """More queries than slots: finished slots are refilled while the others
    run on; every query's result equals the same query run alone, bit for
    bit (a query's arithmetic does not depend on its tile mates)."""
    e = graph13
    nq = 24
    off, seeds = generators.seed_sets(e, nq, 1, 32, seed=11)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.95, 0.85][q % 3], 300, [1e-6, 1e-9][q % 2])
            for q in range(nq)]
    scores, runs = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx, tile=8)
    for q in [0, 1, 2, 7, 13, 23]:
        s1, r1 = ga.ppr_batch_run(dg, [cfgs[q]], off[q:q + 2] - off[q], seeds[off[q]:off[q + 1]],
                                  ctx=ctx, tile=8)
        assert r1[0].iterations == runs[q].iterations
        assert r1[0].diff_l1 == runs[q].diff_l1
        assert np.array_equal(s1[0], scores[q])


def test_batch_zero_in_degree_seeds(oracle, ctx, graph13):
    """Rows without in-edges are swept only while they seed a query of the
    tile.  Queries seeded on such nodes (with out-edges, and isolated ones),
    alone and mixed with ordinary seeds, through refilled slots (a slot's
    previous query may have seeded other such rows): every result matches
    the oracle."""
    e = graph13
    indeg = np.bincount(np.asarray(e.targets, dtype=np.int64), minlength=e.n)
    outdeg = np.diff(e.offsets.astype(np.int64))
    src_only = np.nonzero((indeg == 0) & (outdeg > 0))[0]
    isolated = np.nonzero((indeg == 0) & (outdeg == 0))[0]
    assert src_only.size >= 40 and isolated.size >= 2
    rng = np.random.default_rng(29)
    sets = []
    for q in range(30):
        k = q % 5
        if k == 0:
            s = rng.choice(src_only, 3, replace=False)
        elif k == 1:
            s = np.array([isolated[q % isolated.size]])
        elif k == 2:
            s = np.concatenate([rng.choice(src_only, 2, replace=False),
                                rng.integers(0, e.n, 3)])
        elif k == 3:
            s = np.concatenate([isolated[:2], rng.choice(src_only, 1)])
        else:
            s = rng.integers(0, e.n, 4)
        sets.append(np.asarray(s, dtype=np.uint32))
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum([s.size for s in sets])
    seeds = np.concatenate(sets)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 400, [1e-8, 1e-10][q % 2])
            for q in range(len(sets))]
    scores, runs = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx, tile=8)
    for q in range(len(sets)):
        p = dense_p(e.n, sets[q])
        ref = oracle_ppr(oracle, e, p, cfgs[q].damping, cfgs[q].tolerance, 400)
        check(runs[q].iterations, runs[q].converged, runs[q].diff_l1, scores[q], ref,
              cfgs[q].tolerance)


def test_batch_graph_replay_same_bits(ctx, graph13):
    """GOPS_LAUNCH_GRAPH: the tile sweeps replayed from captured chains of
    four (captured again after each admission) give the bits of the queued
    kernels, with refills; the server's steps too."""
    from graphops_amd import _lib as L

    e = graph13
    nq = 40
    off, seeds = generators.seed_sets(e, nq, 1, 64, seed=19)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 300, [1e-6, 1e-9][q % 2])
            for q in range(nq)]
    s1, r1 = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx, tile=16)
    ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_GRAPH)
    try:
        s2, r2 = ga.ppr_batch_run(dg, cfgs, off, seeds, ctx=ctx, tile=16)
        srv = ga.Server(dg, tile=8, max_slots=16)
        tickets = {srv.submit(cfgs[q], seeds[off[q]:off[q + 1]]): q for q in range(12)}
        served = {}
        steps = 0
        while len(served) < 12:
            srv.step(3)
            steps += 1
            while (r := srv.poll()) is not None:
                served[tickets[r[0]]] = r[1]
            assert steps < 2000
        srv.close()
    finally:
        ctx.set_option(L.OPT_LAUNCH, L.LAUNCH_AUTO)
    for q in range(nq):
        assert r1[q].iterations == r2[q].iterations and r1[q].diff_l1 == r2[q].diff_l1
        assert np.array_equal(s1[q], s2[q])
    s8, r8 = ga.ppr_batch_run(dg, cfgs[:12], off[:13], seeds[:off[12]], ctx=ctx, tile=8)
    for q in range(12):
        assert served[q].iterations == r8[q].iterations
        assert np.array_equal(served[q].scores, s8[q])


def test_plain_path_matches_tiles(ctx, graph13):
    """tile=-1: one query per device run (the single-vector chain)."""
    e = graph13
    off, seeds = generators.seed_sets(e, 6, 1, 16, seed=3)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfg = ga.PageRankConfig(0.85, 300, 1e-8)
    s1, r1 = ga.ppr_batch_run(dg, cfg, off, seeds, ctx=ctx, tile=-1)
    s2, r2 = ga.ppr_batch_run(dg, cfg, off, seeds, ctx=ctx, tile=16)
    for q in range(6):
        assert r1[q].iterations == r2[q].iterations
        assert np.abs(s1[q] - s2[q]).sum() <= 2e-10


def test_batch_edge_cases(oracle, ctx, graph13):
    e = graph13
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    # duplicate seeds add up; zero-sum personalization falls back to uniform
    off = np.array([0, 3, 4, 4], np.uint64)
    seeds = np.array([5, 5, 9, 7], np.uint32)
    w = np.array([1.0, 2.0, 1.0, 0.0])
    cfg = ga.PageRankConfig(0.85, 200, 1e-9)
    scores, runs = ga.ppr_batch_run(dg, cfg, off, seeds, seed_weights=w, ctx=ctx)
    for q, p in enumerate([dense_p(e.n, [5, 5, 9], [1.0, 2.0, 1.0]), dense_p(e.n, [7], [0.0]),
                           np.zeros(e.n)]):
        ref = oracle_ppr(oracle, e, p, 0.85, 1e-9, 200)
        check(runs[q].iterations, runs[q].converged, runs[q].diff_l1, scores[q], ref, 1e-9)
    with pytest.raises(ga.IndexOutOfBounds):
        ga.ppr_batch_run(dg, cfg, np.array([0, 1], np.uint64), np.array([e.n], np.uint32), ctx=ctx)


def test_server_continuous_batching(oracle, ctx, graph13):
    """Queries arrive while others run; each result equals the oracle and,
    bit for bit, the same query served alone."""
    e = graph13
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    srv = ga.Server(dg, tile=8, max_slots=16)
    assert srv.stats()["slots"] == 16
    rng = np.random.default_rng(11)
    off, seeds = generators.seed_sets(e, 40, 1, 256, log_uniform=True, seed=17)
    cfgs = [ga.PageRankConfig(float(rng.choice([0.5, 0.85, 0.95])), 500, 1e-6) for _ in range(40)]
    tickets = {}
    results = {}
    nxt = 0
    steps = 0
    while len(results) < 40:
        # a few arrivals between device steps
        for _ in range(int(rng.integers(0, 4))):
            if nxt < 40:
                tickets[srv.submit(cfgs[nxt], seeds[off[nxt]:off[nxt + 1]])] = nxt
                nxt += 1
        if nxt < 40 and srv.stats()["queued"] == 0 and srv.stats()["active"] == 0:
            tickets[srv.submit(cfgs[nxt], seeds[off[nxt]:off[nxt + 1]])] = nxt
            nxt += 1
        srv.step(4)
        steps += 1
        while True:
            r = srv.poll()
            if r is None:
                break
            ticket, run, t_sub, t_fin, _ = r
            results[tickets[ticket]] = run
            assert t_fin >= t_sub
        assert steps < 5000
    for q in range(40):
        p = dense_p(e.n, seeds[off[q]:off[q + 1]])
        ref = oracle_ppr(oracle, e, p, cfgs[q].damping, 1e-6, 500)
        run = results[q]
        check(run.iterations, run.converged, run.diff_l1, run.scores, ref, 1e-6)
    # alone == in the mix
    alone = ga.Server(dg, tile=8, max_slots=8)
    for q in (0, 7, 23):
        alone.submit(cfgs[q], seeds[off[q]:off[q + 1]])
        got = None
        while got is None:
            alone.step(4)
            got = alone.poll()
        run = got[1]
        # every query keeps its own sweep mode and lane arithmetic, so the
        # result does not depend on the other queries or on the slot
        assert run.iterations == results[q].iterations
        assert run.diff_l1 == results[q].diff_l1
        assert np.array_equal(run.scores, results[q].scores)
    srv.close()
    alone.close()


def test_server_admission_refused_without_memory(ctx, graph13):
    dg = ga.DeviceGraph.from_edges(graph13, ctx=ctx)
    with pytest.raises(ga.OutOfMemory):
        ga.Server(dg, tile=16, hbm_budget_bytes=1024.0)


def test_server_late_tiles_start_mid_flight(oracle, ctx, graph13):
    """A server whose later tiles take their first queries while earlier ones
    run (a few submissions per step), created after another server's run left
    the allocator's memory dirty: every result matches the oracle.  (Idle
    tiles' device state used to be read back uninitialised and uploaded again
    at their first admission.)"""
    e = graph13
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    nq = 48
    off, seeds = generators.seed_sets(e, nq, 1, 64, seed=5)
    cfgs = [ga.PageRankConfig([0.5, 0.85, 0.95][q % 3], 500, 1e-7) for q in range(nq)]

    def serve(per_step):
        srv = ga.Server(dg, tile=8, max_slots=32)  # four tiles of 8 queries
        tickets, results, nxt, steps = {}, {}, 0, 0
        while len(results) < nq:
            for _ in range(per_step):
                if nxt < nq:
                    tickets[srv.submit(cfgs[nxt], seeds[off[nxt]:off[nxt + 1]])] = nxt
                    nxt += 1
            srv.step(2)
            steps += 1
            while True:
                r = srv.poll()
                if r is None:
                    break
                results[tickets[r[0]]] = r[1]
            assert steps < 5000
        srv.close()
        return results

    serve(nq)  # runs to completion, then frees its tiles
    results = serve(6)
    for q in range(nq):
        p = dense_p(e.n, seeds[off[q]:off[q + 1]])
        ref = oracle_ppr(oracle, e, p, cfgs[q].damping, 1e-7, 500)
        run = results[q]
        check(run.iterations, run.converged, run.diff_l1, run.scores, ref, 1e-7)


def test_server_topk(oracle, ctx, graph13):
    """A top-k server selects each result's top_k on the device and returns
    only those pairs; they equal the oracle's top_k (topk.rs:17-42) of the
    same query's full scores from a server without top-k (the same tile
    width and slot, so bitwise the same scores)."""
    e = graph13
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    cfg = ga.PageRankConfig(0.85, 300, 1e-8)
    res = []
    for topk in (0, 10):
        srv = ga.Server(dg, tile=8, max_slots=8, topk=topk)
        srv.submit(cfg, [3, 17])
        got = None
        while got is None:
            srv.step(8)
            got = srv.poll()
        res.append(got)
        if topk:
            with pytest.raises(ga.InvalidParameter, match="top-k server"):
                srv.poll(with_scores=True)
        srv.close()
    full, top = res[0][1].scores, res[1][4]
    assert res[1][1].scores is None and res[1][1].iterations == res[0][1].iterations
    assert top == oracle.top_k(full, 10)
