"""Pins of the oracle's Katz, eigenvector and HITS restatements (CPU only).

What pins them, independently of the oracle's own code:
  * exact arithmetic: Katz trajectories in rationals (fractions); eigenvector
    and HITS trajectories in 60-digit decimals (their square roots are
    irrational);
  * the mathematics of the fixed points: Katz = beta (I - alpha A^T)^-1 1
    (linear solve), eigenvector = the dominant eigenvector of A + I, HITS =
    the top singular vectors of A (numpy);
  * the reference's own unit tests (katz.rs:231-329, eigenvector.rs:145-197,
    centrality.rs:244-273), its validation messages and special cases;
  * planted mistakes (a transposed operand, on the CPU) that fail the pins.
"""
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

GRAPHS = {
    # katz.rs:125-130 / 234-257 (undirected path as symmetric lists)
    "path4": [[1], [0, 2], [1, 3], [2]],
    # katz.rs:261-275, eigenvector.rs:146-156
    "triangle": [[1, 2], [0, 2], [0, 1]],
    # eigenvector.rs:159-167
    "star": [[1, 2, 3, 4], [0], [0], [0], [0]],
    # eigenvector.rs:190-197
    "l2case": [[1, 2, 3], [0, 2], [0, 1, 3], [0, 2]],
    # directed: a parallel edge, a self-loop, a dangling node; A != A^T
    "directed": [[1, 2], [2], [0, 3], [], [0, 0, 4, 1]],
    # centrality.rs:244-265 (directed star, out from the centre)
    "star_out": [[1, 2, 3, 4], [], [], [], []],
}


def lists_from(nbrs):
    off = np.zeros(len(nbrs) + 1, dtype=np.uint64)
    for u, lst in enumerate(nbrs):
        off[u + 1] = off[u] + len(lst)
    nb = np.array([v for lst in nbrs for v in lst], dtype=np.uint32)
    return off, nb


def adjacency(nbrs):
    """A[u][v] = number of entries v in neighbors(u)."""
    n = len(nbrs)
    a = np.zeros((n, n))
    for u, lst in enumerate(nbrs):
        for v in lst:
            a[u, v] += 1.0
    return a


# ---------------------------------------------------------------- Katz
def katz_exact(nbrs, alpha, beta, sweeps):
    """x <- alpha A^T x + beta 1 in rationals (alpha, beta: the exact values
    of the f64 inputs); returns the iterates and residuals."""
    n = len(nbrs)
    a, b = Fraction(alpha), Fraction(beta)
    x = [b] * n
    xs, diffs = [], []
    for _ in range(sweeps):
        y = [b] * n
        for u, lst in enumerate(nbrs):
            for v in lst:
                y[v] += a * x[u]
        diffs.append(sum(abs(p - q) for p, q in zip(x, y)))
        x = y
        xs.append(x)
    return xs, diffs


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("alpha,beta", [(0.1, 1.0), (0.3, 0.5)])
def test_katz_exact_trajectories(oracle, name, alpha, beta):
    off, nb = lists_from(GRAPHS[name])
    xs, diffs = katz_exact(GRAPHS[name], alpha, beta, 6)
    for k in (1, 3, 6):
        r = oracle.katz_run(off, nb, alpha, beta, k, 0.0)
        assert r.iterations == k and not r.converged
        want = np.array([float(v) for v in xs[k - 1]])
        np.testing.assert_allclose(r.scores, want, rtol=4e-16 * 8 * k, atol=0)
        # the residual cancels: each |x_k - x_k+1| carries the iterates' rounding
        slack = 8 * want.size * 2.3e-16 * np.abs(want).max()
        assert abs(r.diff_l1 - float(diffs[k - 1])) <= 1e-14 * float(diffs[k - 1]) + slack


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_katz_fixed_point_is_linear_solve(oracle, name):
    a = adjacency(GRAPHS[name])
    n = a.shape[0]
    alpha, beta = 0.2, 1.5
    rho = max(abs(np.linalg.eigvals(a)))
    assert alpha * rho < 1.0
    want = np.linalg.solve(np.eye(n) - alpha * a.T, np.full(n, beta))
    off, nb = lists_from(GRAPHS[name])
    r = oracle.katz_run(off, nb, alpha, beta, 10000, 1e-14)
    assert r.converged
    np.testing.assert_allclose(r.scores, want, rtol=1e-11, atol=1e-12)


def test_katz_reference_unit_tests(oracle):
    # katz.rs:233-257: inner nodes of the path out-score the endpoints
    off, nb = lists_from(GRAPHS["path4"])
    s = oracle.katz_run(off, nb).scores
    assert s.size == 4 and s[1] > s[0] and s[2] > s[3]
    # katz.rs:261-275: triangle converges to equal scores
    off, nb = lists_from(GRAPHS["triangle"])
    r = oracle.katz_run(off, nb)
    assert r.converged and np.all(np.abs(r.scores - r.scores[0]) < 1e-6)
    # katz.rs:301-328: random undirected graphs give finite, non-negative scores
    rng = np.random.default_rng(5)
    for _ in range(50):
        n = int(rng.integers(1, 10))
        adj = [set() for _ in range(n)]
        for u, v in rng.integers(0, 10, (int(rng.integers(0, 30)), 2)):
            if u < n and v < n and u != v:
                adj[u].add(int(v))
                adj[v].add(int(u))
        off, nb = lists_from([sorted(x) for x in adj])
        s = oracle.katz_run(off, nb).scores
        assert s.size == n and np.all(np.isfinite(s)) and np.all(s >= 0.0)


def test_katz_validation_and_special_cases(oracle):
    ok = dict(alpha=0.1, beta=1.0, max_iterations=100, tolerance=1e-6)
    oracle.validate_katz(**ok)
    cases = [
        (dict(alpha=float("nan")), "alpha must be finite and > 0"),
        (dict(alpha=0.0), "alpha must be finite and > 0"),
        (dict(alpha=-1.0), "alpha must be finite and > 0"),
        (dict(alpha=1.0), "alpha must be < 1 (required for convergence"),  # katz.rs:279-299
        (dict(alpha=2.5), "alpha must be < 1"),
        (dict(beta=float("inf")), "beta must be finite"),
        (dict(max_iterations=0), "max_iterations must be > 0"),
        (dict(tolerance=0.0), "tolerance must be finite and > 0"),
        (dict(tolerance=float("nan")), "tolerance must be finite and > 0"),
    ]
    for change, msg in cases:
        with pytest.raises(oracle.OracleError, match=msg.replace("(", r"\(")):
            oracle.validate_katz(**{**ok, **change})
    r = oracle.katz_run(np.zeros(1, np.uint64), np.zeros(0, np.uint32))  # katz.rs:138-145
    assert (r.scores.size, r.iterations, r.diff_l1, r.converged) == (0, 0, 0.0, True)
    off, nb = lists_from(GRAPHS["triangle"])
    r = oracle.katz_run(off, nb, 0.1, 2.0, 0, 1e-6)  # the loop never runs
    assert r.iterations == 0 and r.diff_l1 == float("inf") and not r.converged
    np.testing.assert_array_equal(r.scores, [2.0, 2.0, 2.0])
    # no edges: every node keeps beta; converged at the first sweep with diff 0
    r = oracle.katz_run(np.zeros(4, np.uint64), np.zeros(0, np.uint32), 0.1, 0.7)
    assert r.iterations == 1 and r.converged and r.diff_l1 == 0.0
    np.testing.assert_array_equal(r.scores, [0.7] * 3)


# ---------------------------------------------------------------- eigenvector
getcontext().prec = 60


def eig_decimal(nbrs, sweeps):
    n = len(nbrs)
    x = [Decimal(1) / Decimal(n).sqrt()] * n
    xs, diffs = [], []
    for _ in range(sweeps):
        y = [x[v] + sum((x[u] for u in nbrs[v]), Decimal(0)) for v in range(n)]
        norm = sum(t * t for t in y).sqrt()
        if norm > 0:
            sign = 1 if sum(y) >= 0 else -1
            y = [t * sign / norm for t in y]
        diffs.append(sum(abs(p - q) for p, q in zip(x, y)))
        x = y
        xs.append(x)
    return xs, diffs


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_eigenvector_decimal_trajectories(oracle, name):
    off, nb = lists_from(GRAPHS[name])
    xs, diffs = eig_decimal(GRAPHS[name], 8)
    for k in (1, 4, 8):
        r = oracle.eigenvector_run(off, nb, k, 0.0)
        assert r.iterations == k
        want = np.array([float(v) for v in xs[k - 1]])
        np.testing.assert_allclose(r.scores, want, rtol=1e-14, atol=1e-16)
        slack = 8 * want.size * 2.3e-16 * np.abs(want).max()
        assert abs(r.diff_l1 - float(diffs[k - 1])) <= 1e-13 * float(diffs[k - 1]) + slack


@pytest.mark.parametrize("name", ["path4", "triangle", "star", "l2case", "directed"])
def test_eigenvector_fixed_point_is_perron_vector(oracle, name):
    a = adjacency(GRAPHS[name])
    w, v = np.linalg.eig(a + np.eye(a.shape[0]))
    i = int(np.argmax(w.real))
    vec = np.abs(v[:, i].real)
    vec /= np.linalg.norm(vec)
    off, nb = lists_from(GRAPHS[name])
    r = oracle.eigenvector_run(off, nb, 100000, 1e-14)
    assert r.converged
    np.testing.assert_allclose(r.scores, vec, atol=1e-9)


def test_eigenvector_reference_unit_tests(oracle):
    off, nb = lists_from(GRAPHS["triangle"])  # eigenvector.rs:145-156
    r = oracle.eigenvector_run(off, nb, 100, 1e-8)
    assert r.converged and np.all(np.abs(r.scores - 1.0 / np.sqrt(3.0)) < 1e-6)
    off, nb = lists_from(GRAPHS["star"])  # :158-167
    s = oracle.eigenvector_run(off, nb).scores
    assert s[0] > s[1] and abs(s[1] - s[2]) < 1e-6
    r = oracle.eigenvector_run(np.zeros(1, np.uint64), np.zeros(0, np.uint32))  # :169-175
    assert r.scores.size == 0 and r.converged and r.iterations == 0
    s = oracle.eigenvector_run(np.zeros(4, np.uint64), np.zeros(0, np.uint32)).scores  # :177-187
    assert np.all(np.abs(s) < 1e-10) or np.all(np.abs(s - 1.0 / np.sqrt(3.0)) < 1e-10)
    off, nb = lists_from(GRAPHS["l2case"])  # :189-197
    assert abs(np.linalg.norm(oracle.eigenvector_run(off, nb).scores) - 1.0) < 1e-6
    # max_iterations = 0: the start vector, diff 0.0 (eigenvector.rs:64-69), not converged
    r = oracle.eigenvector_run(off, nb, 0, 1e-6)
    assert r.iterations == 0 and r.diff_l1 == 0.0 and not r.converged
    np.testing.assert_array_equal(r.scores, np.full(4, 0.5))


# ---------------------------------------------------------------- HITS
def hits_decimal(nbrs, sweeps):
    n = len(nbrs)
    x = Decimal(1) / Decimal(n).sqrt()
    hub, auth = [x] * n, [x] * n

    def normalize(v):
        norm = sum(t * t for t in v).sqrt()
        return [t / norm for t in v] if norm > 0 else v

    out = []
    for _ in range(sweeps):
        na = [Decimal(0)] * n
        for u, lst in enumerate(nbrs):
            for v in lst:
                na[v] += hub[u]
        na = normalize(na)
        nh = normalize([sum((na[v] for v in nbrs[u]), Decimal(0)) for u in range(n)])
        ad = sum((p - q) ** 2 for p, q in zip(auth, na)).sqrt()
        hd = sum((p - q) ** 2 for p, q in zip(hub, nh)).sqrt()
        hub, auth = nh, na
        out.append((hub, auth, ad, hd))
    return out


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_hits_decimal_trajectories(oracle, name):
    off, nb = lists_from(GRAPHS[name])
    traj = hits_decimal(GRAPHS[name], 6)
    for k in (1, 3, 6):
        r = oracle.hits(off, nb, k, 0.0)
        hub, auth, ad, hd = traj[k - 1]
        assert r.iterations == k
        np.testing.assert_allclose(r.hub, [float(t) for t in hub], rtol=1e-14, atol=1e-16)
        np.testing.assert_allclose(r.auth, [float(t) for t in auth], rtol=1e-14, atol=1e-16)
        # an L2 change of nearly equal unit vectors: absolute rounding ~ sqrt(n) eps
        assert abs(r.auth_diff - float(ad)) <= 1e-12 * float(ad) + 1e-14
        assert abs(r.hub_diff - float(hd)) <= 1e-12 * float(hd) + 1e-14


@pytest.mark.parametrize("name", ["directed", "l2case", "triangle", "star_out"])
def test_hits_fixed_point_is_top_singular_pair(oracle, name):
    a = adjacency(GRAPHS[name])
    u, s, vt = np.linalg.svd(a)
    assert s[0] > 1.01 * s[1]  # a simple top singular value
    off, nb = lists_from(GRAPHS[name])
    r = oracle.hits(off, nb, 100000, 1e-14)
    np.testing.assert_allclose(r.hub, np.abs(u[:, 0]), atol=1e-9)
    np.testing.assert_allclose(r.auth, np.abs(vt[0]), atol=1e-9)


def test_hits_reference_unit_tests(oracle):
    off, nb = lists_from(GRAPHS["star_out"])  # centrality.rs:244-265
    r = oracle.hits(off, nb, 100, 1e-10)
    assert r.hub.size == 5 and np.all(r.hub[0] > r.hub[1:]) and np.all(r.auth[1:] > r.auth[0])
    r = oracle.hits(np.zeros(1, np.uint64), np.zeros(0, np.uint32), 100, 1e-10)  # :267-273
    assert r.hub.size == 0 and r.auth.size == 0
    # max_iter = 0: the start vectors
    off, nb = lists_from(GRAPHS["directed"])
    r = oracle.hits(off, nb, 0, 1e-10)
    np.testing.assert_array_equal(r.hub, np.full(5, 1.0 / np.sqrt(5.0)))
    np.testing.assert_array_equal(r.auth, r.hub)


def test_planted_transposes_are_caught(oracle):
    """The oracle run on A^T (a transposed operand planted in its input) must
    fail the fixed-point pins of all three on a directed graph."""
    g = GRAPHS["directed"]
    a = adjacency(g)
    n = a.shape[0]
    at_lists = [[v for v in range(n) for _ in range(int(a[v, u]))] for u in range(n)]
    off_t, nb_t = lists_from(at_lists)
    katz = np.linalg.solve(np.eye(n) - 0.2 * a.T, np.full(n, 1.5))
    assert np.abs(oracle.katz_run(off_t, nb_t, 0.2, 1.5, 10000, 1e-14).scores - katz).max() > 1e-3
    w, v = np.linalg.eig(a + np.eye(n))
    vec = np.abs(v[:, int(np.argmax(w.real))].real)
    vec /= np.linalg.norm(vec)
    assert np.abs(oracle.eigenvector_run(off_t, nb_t, 100000, 1e-14).scores - vec).max() > 1e-3
    u, s, vt = np.linalg.svd(a)
    r = oracle.hits(off_t, nb_t, 100000, 1e-14)
    assert np.abs(r.hub - np.abs(u[:, 0])).max() > 1e-3
