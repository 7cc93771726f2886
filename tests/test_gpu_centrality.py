"""GPU parity of Katz, eigenvector and HITS centrality (through the C ABI)
against the f64 oracle.

Tolerances (f64 throughout on both sides; only the order of the additions
differs, pull sums by row instead of the reference's push loops):
  * after the same number of iterations, every score agrees to 1e-12 of the
    largest score (1e-11 for HITS, whose normalisations run twice per step),
    plus, on graphs with very long rows, the bound of a reordered sum of L
    terms (L * eps of the sum per iteration);
  * iteration counts are equal unless the oracle's last residual lies
    within 1e-6 (relative) of the tolerance.
"""
import numpy as np
import pytest

import graphops_amd as ga
from graphops_amd import generators

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


SMALL = {
    "path4": [[1], [0, 2], [1, 3], [2]],
    "triangle": [[1, 2], [0, 2], [0, 1]],
    "star": [[1, 2, 3, 4], [0], [0], [0], [0]],
    "l2case": [[1, 2, 3], [0, 2], [0, 1, 3], [0, 2]],
    "directed": [[1, 2], [2], [0, 3], [], [0, 0, 4, 1]],
    "star_out": [[1, 2, 3, 4], [], [], [], []],
    "isolated": [[], [], []],
}


def lists_from(nbrs):
    off = np.zeros(len(nbrs) + 1, dtype=np.uint64)
    for u, lst in enumerate(nbrs):
        off[u + 1] = off[u] + len(lst)
    nb = np.array([v for lst in nbrs for v in lst], dtype=np.uint32)
    return off, nb


def close(got, want, rel):
    scale = max(np.abs(want).max() if want.size else 0.0, 1e-300)
    return np.abs(got - want).max(initial=0.0) <= rel * scale


def residual_close(got, want, scores, rel):
    """Residuals cancel (|x_k - x_k+1| of nearly equal iterates): each term
    carries the iterates' own rounding, n * eps * max|x| in all."""
    noise = 4 * scores.size * 2.3e-16 * np.abs(scores).max(initial=0.0)
    return abs(got - want) <= rel * want + noise


def same_stop(got_it, ref, tol):
    if got_it == ref.iterations:
        return True
    return abs(got_it - ref.iterations) == 1 and abs(ref.diff_l1 - tol) <= 1e-6 * tol


@pytest.mark.parametrize("name", sorted(SMALL))
def test_small_graphs_against_oracle(oracle, ctx, name):
    off, nb = lists_from(SMALL[name])
    n = len(SMALL[name])
    dg = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx)
    for alpha, beta, tol in ((0.1, 1.0, 1e-6), (0.3, 0.5, 1e-12)):
        ref = oracle.katz_run(off, nb, alpha, beta, 1000, tol)
        r = ga.katz_centrality_run(dg, ga.KatzConfig(alpha, beta, 1000, tol), ctx=ctx)
        assert same_stop(r.iterations, ref, tol) and r.converged == ref.converged, name
        assert close(r.scores, ref.scores, 1e-12), (name, r.scores, ref.scores)
    for tol in (1e-6, 1e-12):
        ref = oracle.eigenvector_run(off, nb, 1000, tol)
        r = ga.eigenvector_centrality_run(dg, 1000, tol, ctx=ctx)
        assert same_stop(r.iterations, ref, tol) and r.converged == ref.converged, name
        assert close(r.scores, ref.scores, 1e-12), (name, r.scores, ref.scores)
    ref = oracle.hits(off, nb, 100, 1e-10)
    hub, auth, info = ga.hits_run(dg, 100, 1e-10, ctx=ctx)
    assert info["iterations"] == ref.iterations, name
    assert close(hub, ref.hub, 1e-11) and close(auth, ref.auth, 1e-11), name


def rmat(scale, seed):
    e = generators.rmat(scale, 16, seed=seed)
    return e, np.asarray(e.offsets), np.asarray(e.targets)


@pytest.mark.parametrize("sweeps", [1, 7])
def test_rmat_fixed_iterations(oracle, ctx, sweeps):
    """R-MAT scale 14: the same number of iterations (tolerance 0) on both
    sides; Katz with alpha below 1/max in-degree."""
    e, off, nb = rmat(14, 3)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    alpha = 0.5 / np.bincount(nb, minlength=e.n).max()
    ref = oracle.katz_run(off, nb, alpha, 1.0, sweeps, 0.0)
    r = ga.katz_centrality_run(dg, ga.KatzConfig(alpha, 1.0, sweeps, 0.0), ctx=ctx)
    assert r.iterations == sweeps and not r.converged
    assert close(r.scores, ref.scores, 1e-12)
    assert residual_close(r.diff_l1, ref.diff_l1, ref.scores, 1e-9)
    ref = oracle.eigenvector_run(off, nb, sweeps, 0.0)
    r = ga.eigenvector_centrality_run(dg, sweeps, 0.0, ctx=ctx)
    assert r.iterations == sweeps and close(r.scores, ref.scores, 1e-12)
    assert residual_close(r.diff_l1, ref.diff_l1, ref.scores, 1e-9)
    ref = oracle.hits(off, nb, sweeps, 0.0)
    hub, auth, info = ga.hits_run(dg, sweeps, 0.0, ctx=ctx)
    assert info["iterations"] == sweeps
    assert close(hub, ref.hub, 1e-11) and close(auth, ref.auth, 1e-11)
    assert residual_close(info["hub_diff"], ref.hub_diff, ref.hub, 1e-8)


def test_rmat_converged(oracle, ctx):
    e, off, nb = rmat(12, 5)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx)
    alpha = 0.9 / np.bincount(nb, minlength=e.n).max()
    ref = oracle.katz_run(off, nb, alpha, 1.0, 500, 1e-10)
    r = ga.katz_centrality_run(dg, ga.KatzConfig(alpha, 1.0, 500, 1e-10), ctx=ctx)
    assert r.converged and same_stop(r.iterations, ref, 1e-10)
    assert close(r.scores, ref.scores, 1e-12)
    ref = oracle.eigenvector_run(off, nb, 2000, 1e-9)
    r = ga.eigenvector_centrality_run(dg, 2000, 1e-9, ctx=ctx)
    assert r.converged == ref.converged and same_stop(r.iterations, ref, 1e-9)
    assert close(r.scores, ref.scores, 1e-11)


def test_hub_rows_both_orientations(oracle, ctx):
    """A node with 70000 in-edges (a hub row of the in-edge CSR) and one with
    70000 out-edges (a hub row of the out-edge CSR), rows between 4096 and
    65536 edges, and short rows: every descriptor kind on both sides."""
    n = 80000
    rng = np.random.default_rng(2)
    lists = [[] for _ in range(n)]
    lists[1] = list(range(2, 70002))           # out-hub
    for u in range(2, n):
        if u < 72000:
            lists[u].append(0)                 # in-hub
        if u % 11 == 0:
            lists[u].append(3)                 # a mid-size in-row
        lists[u].extend(rng.integers(0, n, size=int(rng.integers(0, 3))).tolist())
    lists[5] = rng.integers(0, n, size=9000).tolist()  # a mid-size out-row
    off, nb = lists_from(lists)
    dg = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx)
    indeg = np.bincount(nb, minlength=n)
    alpha = 0.5 / indeg.max()
    # a sum of L terms in another order differs by up to L * eps of the sum
    # (the reference's sequential sum carries that error itself): 72000-term
    # rows allow about 1e-11 per iteration
    maxdeg = max(int(indeg.max()), int(np.diff(off).max()))
    for k in (1, 4):
        rel = 1e-12 + 2 * k * maxdeg * 1.2e-16
        ref = oracle.katz_run(off, nb, alpha, 1.0, k, 0.0)
        r = ga.katz_centrality_run(dg, ga.KatzConfig(alpha, 1.0, k, 0.0), ctx=ctx)
        assert close(r.scores, ref.scores, rel)
        ref = oracle.eigenvector_run(off, nb, k, 0.0)
        r = ga.eigenvector_centrality_run(dg, k, 0.0, ctx=ctx)
        assert close(r.scores, ref.scores, rel)
        ref = oracle.hits(off, nb, k, 0.0)
        hub, auth, _ = ga.hits_run(dg, k, 0.0, ctx=ctx)
        assert close(hub, ref.hub, 2 * rel) and close(auth, ref.auth, 2 * rel)


def test_edge_cases_and_checked(ctx):
    empty = ga.DeviceGraph.from_lists(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32), ctx=ctx)
    r = ga.katz_centrality_run(empty, ctx=ctx)
    assert (r.scores.size, r.iterations, r.diff_l1, r.converged) == (0, 0, 0.0, True)
    r = ga.eigenvector_centrality_run(empty, ctx=ctx)
    assert (r.scores.size, r.iterations, r.converged) == (0, 0, True)
    hub, auth = ga.hits(empty, 100, 1e-10, ctx=ctx)
    assert hub.size == 0 and auth.size == 0
    off, nb = lists_from(SMALL["directed"])
    g = ga.DeviceGraph.from_lists(5, off, nb, ctx=ctx)
    r = ga.katz_centrality_run(g, ga.KatzConfig(0.1, 2.0, 0, 1e-6), ctx=ctx)
    assert r.iterations == 0 and r.diff_l1 == float("inf") and not r.converged
    np.testing.assert_array_equal(r.scores, np.full(5, 2.0))
    r = ga.eigenvector_centrality_run(g, 0, 1e-6, ctx=ctx)
    assert r.iterations == 0 and r.diff_l1 == 0.0 and not r.converged
    np.testing.assert_array_equal(r.scores, np.full(5, 1.0 / np.sqrt(5.0)))
    hub, auth = ga.hits(g, 0, 1e-6, ctx=ctx)
    np.testing.assert_array_equal(hub, np.full(5, 1.0 / np.sqrt(5.0)))
    # katz.rs:57-84, 279-299
    for cfg, msg in ((ga.KatzConfig(alpha=1.0), "alpha must be < 1"),
                     (ga.KatzConfig(alpha=2.5), "alpha must be < 1"),
                     (ga.KatzConfig(alpha=0.0), "alpha must be finite and > 0"),
                     (ga.KatzConfig(beta=float("nan")), "beta must be finite"),
                     (ga.KatzConfig(max_iterations=0), "max_iterations must be > 0"),
                     (ga.KatzConfig(tolerance=-1.0), "tolerance must be finite and > 0")):
        with pytest.raises(ga.InvalidParameter, match="invalid parameter: " + msg):
            ga.katz_centrality_checked(g, cfg, ctx=ctx)
        with pytest.raises(ga.InvalidParameter, match=msg):
            cfg.validate()
    s = ga.katz_centrality_checked(g, ga.KatzConfig(), ctx=ctx)
    assert s.size == 5 and np.all(s > 0)
    # the reference's unit tests on the device path (katz.rs:233-275,
    # eigenvector.rs:145-197, centrality.rs:244-265)
    path = ga.DeviceGraph.from_lists(4, *lists_from(SMALL["path4"]), ctx=ctx)
    s = ga.katz_centrality(path, ctx=ctx)
    assert s[1] > s[0] and s[2] > s[3]
    tri = ga.DeviceGraph.from_lists(3, *lists_from(SMALL["triangle"]), ctx=ctx)
    r = ga.eigenvector_centrality_run(tri, 100, 1e-8, ctx=ctx)
    assert r.converged and np.all(np.abs(r.scores - 1.0 / np.sqrt(3.0)) < 1e-6)
    star = ga.DeviceGraph.from_lists(5, *lists_from(SMALL["star_out"]), ctx=ctx)
    hub, auth = ga.hits(star, 100, 1e-10, ctx=ctx)
    assert np.all(hub[0] > hub[1:]) and np.all(auth[1:] > auth[0])


@pytest.fixture(scope="module")
def ctx_kb():
    from graphops_amd import _lib as L

    c = ga.Context(0)
    c.set_option(L.OPT_CENTRALITY_SWEEP, L.CENTRALITY_SWEEP_BLOCKED)
    return c


def katz_graphs():
    rng = np.random.default_rng(9)
    out = {k: lists_from(v) + (len(v),) for k, v in SMALL.items()}
    e, off, nb = rmat(14, 3)
    out["rmat14"] = (off.copy(), nb.copy(), e.n)  # e owns the generated arrays
    # rows of every kind in the blocked layout: a row with 70 000 in-edges
    # from every chunk, duplicate entries, rows without in-edges that have
    # out-edges, and n not a multiple of a chunk (16 384) or a bin
    n = 70001
    lists = [rng.integers(0, n, size=int(rng.integers(0, 4))).tolist() for _ in range(n)]
    for u in range(1, 70001):
        lists[u].append(5)
    lists[7].extend([9, 9, 9])
    off, nb = lists_from(lists)
    out["hub70k"] = (off, nb, n)
    return out


KATZ_GRAPHS = katz_graphs()


@pytest.mark.parametrize("name", sorted(KATZ_GRAPHS))
def test_katz_blocked_against_oracle(oracle, ctx_kb, ctx, name):
    """Katz over the blocked layout (GOPS_CENTRALITY_SWEEP_BLOCKED: f64 run sums,
    128-bit fixed-point row sums in units set by a bound on the row sums)
    against the oracle, after a fixed number of iterations and to
    convergence, with betas of either sign and of large magnitude; and
    against the pull on the same graph."""
    off, nb, n = KATZ_GRAPHS[name]
    indeg = np.bincount(nb.astype(np.int64), minlength=n).max(initial=0)
    dgb = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx_kb)
    dgp = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx)
    maxdeg = max(int(indeg), int(np.diff(off.astype(np.int64)).max(initial=0)), 1)
    alpha = 0.5 / max(int(indeg), 1)
    for beta, k in ((1.0, 1), (1.0, 6), (-0.75, 3), (1e200, 2)):
        rel = 1e-12 + 2 * k * maxdeg * 1.2e-16
        ref = oracle.katz_run(off, nb, alpha, beta, k, 0.0)
        r = ga.katz_centrality_run(dgb, ga.KatzConfig(alpha, beta, k, 0.0), ctx=ctx_kb)
        assert ctx_kb.last_stats()["sweeps_blocked"] == k
        assert r.iterations == k and not r.converged
        assert close(r.scores, ref.scores, rel), (name, beta, k)
        assert residual_close(r.diff_l1, ref.diff_l1, ref.scores, 1e-9)
        p = ga.katz_centrality_run(dgp, ga.KatzConfig(alpha, beta, k, 0.0), ctx=ctx)
        assert close(r.scores, p.scores, 2 * rel)
    tol = 1e-10
    ref = oracle.katz_run(off, nb, alpha, 1.0, 500, tol)
    r = ga.katz_centrality_run(dgb, ga.KatzConfig(alpha, 1.0, 500, tol), ctx=ctx_kb)
    assert r.converged == ref.converged and same_stop(r.iterations, ref, tol)
    assert close(r.scores, ref.scores, 1e-12 + 2 * r.iterations * maxdeg * 1.2e-16)
    # a growing run (alpha above 1 / spectral radius on graphs with cycles):
    # scores far beyond the first sweep's scale
    ref = oracle.katz_run(off, nb, 0.9, 1.0, 20, 0.0)
    r = ga.katz_centrality_run(dgb, ga.KatzConfig(0.9, 1.0, 20, 0.0), ctx=ctx_kb)
    assert np.all(np.isfinite(ref.scores))
    assert close(r.scores, ref.scores, 1e-11 + 40 * maxdeg * 1.2e-16), name
    dgb.close()
    dgp.close()


def test_katz_blocked_bitwise_repeatable(ctx_kb):
    e, off, nb = rmat(13, 4)
    dg = ga.DeviceGraph.from_edges(e, ctx=ctx_kb)
    alpha = 0.5 / np.bincount(nb, minlength=e.n).max()
    a = ga.katz_centrality_run(dg, ga.KatzConfig(alpha, 1.0, 100, 1e-12), ctx=ctx_kb)
    b = ga.katz_centrality_run(dg, ga.KatzConfig(alpha, 1.0, 100, 1e-12), ctx=ctx_kb)
    assert a.iterations == b.iterations and a.diff_l1 == b.diff_l1
    assert np.array_equal(a.scores, b.scores)


@pytest.mark.parametrize("name", sorted(KATZ_GRAPHS))
def test_eigen_hits_blocked_against_oracle(oracle, ctx_kb, ctx, name):
    """Eigenvector and HITS over blocked layouts (the out-edge CSR's, and for
    the HITS authority the in-edge one; unit-norm gathered vectors, 128-bit
    row sums in units fixed by the longest row) against the oracle and the
    pulls, after fixed iteration counts and to convergence."""
    off, nb, n = KATZ_GRAPHS[name]
    dgb = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx_kb)
    dgp = ga.DeviceGraph.from_lists(n, off, nb, ctx=ctx)
    maxdeg = max(int(np.bincount(nb.astype(np.int64), minlength=n).max(initial=0)),
                 int(np.diff(off.astype(np.int64)).max(initial=0)), 1)
    for k in (1, 5):
        rel = 1e-12 + 2 * k * maxdeg * 1.2e-16
        ref = oracle.eigenvector_run(off, nb, k, 0.0)
        r = ga.eigenvector_centrality_run(dgb, k, 0.0, ctx=ctx_kb)
        assert ctx_kb.last_stats()["sweeps_blocked"] == k
        assert r.iterations == k and close(r.scores, ref.scores, rel), (name, k)
        assert residual_close(r.diff_l1, ref.diff_l1, ref.scores, 1e-9)
        p = ga.eigenvector_centrality_run(dgp, k, 0.0, ctx=ctx)
        assert close(r.scores, p.scores, 2 * rel)
        ref = oracle.hits(off, nb, k, 0.0)
        hub, auth, info = ga.hits_run(dgb, k, 0.0, ctx=ctx_kb)
        assert ctx_kb.last_stats()["sweeps_blocked"] == 2 * k
        assert info["iterations"] == k
        assert close(hub, ref.hub, 2 * rel) and close(auth, ref.auth, 2 * rel), (name, k)
    ref = oracle.eigenvector_run(off, nb, 1000, 1e-9)
    r = ga.eigenvector_centrality_run(dgb, 1000, 1e-9, ctx=ctx_kb)
    assert r.converged == ref.converged and same_stop(r.iterations, ref, 1e-9)
    assert close(r.scores, ref.scores, 1e-11 + 2 * r.iterations * maxdeg * 1.2e-16)
    ref = oracle.hits(off, nb, 100, 1e-10)
    hub, auth, info = ga.hits_run(dgb, 100, 1e-10, ctx=ctx_kb)
    assert info["iterations"] == ref.iterations, name
    tolh = 1e-11 + 4 * info["iterations"] * maxdeg * 1.2e-16
    assert close(hub, ref.hub, tolh) and close(auth, ref.auth, tolh), name
    dgb.close()
    dgp.close()
