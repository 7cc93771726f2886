"""Pins of the oracle's top_k (topk.rs:17-42), CPU only.

  * the reference's doc example (topk.rs:9-16);
  * brute force on inputs without ties: the k largest valid scores;
  * at ties, a closed description of the entries the reference's heap keeps,
    derived from its rules rather than by running a heap: every score above
    T (the k-th largest), and among the entries equal to T those that
    entered the heap (fewer than k entries >= T before them) minus the ones
    evicted later (smallest index first): the last k - #above of them;
  * invalid scores (not finite, <= 0) never selected; k = 0 and n = 0.
"""
import numpy as np


def described(scores, k):
    s = np.asarray(scores, dtype=np.float64)
    valid = np.nonzero(np.isfinite(s) & (s > 0.0))[0]
    if k == 0 or valid.size == 0:
        return []
    if valid.size <= k:
        keep = valid
    else:
        t = np.sort(s[valid])[::-1][k - 1]
        above = valid[s[valid] > t]
        ge_before = np.cumsum(np.isfinite(s) & (s >= t)) - (np.isfinite(s) & (s >= t))
        ties = valid[s[valid] == t]
        entered = ties[ge_before[ties] < k]
        keep = np.concatenate([above, entered[len(entered) - (k - above.size):]])
    order = sorted(keep.tolist(), key=lambda i: (-s[i], i))
    return [(int(i), float(s[i])) for i in order]


def test_doc_example(oracle):
    assert oracle.top_k([0.1, 0.5, 0.3, 0.9, 0.2], 2) == [(3, 0.9), (1, 0.5)]


def test_tie_free_is_brute_force(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        n = int(rng.integers(1, 300))
        s = rng.random(n) - 0.2
        k = int(rng.integers(0, n + 5))
        valid = [i for i in range(n) if s[i] > 0]
        want = sorted(valid, key=lambda i: -s[i])[:k]
        assert oracle.top_k(s, k) == [(i, float(s[i])) for i in want]


def test_ties_follow_the_heap_rule(oracle):
    # [1, 1, 1, 2], k = 2: entries 0 and 1 enter, 3 evicts 0 -> (3, 2), (1, 1)
    assert oracle.top_k([1.0, 1.0, 1.0, 2.0], 2) == [(3, 2.0), (1, 1.0)]
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = int(rng.integers(1, 200))
        s = rng.choice([0.25, 0.5, 0.5, 1.0, 2.0, 0.0, -1.0, np.nan, np.inf], size=n)
        k = int(rng.integers(0, n + 3))
        assert oracle.top_k(s, k) == described(s, k), (s.tolist(), k)


def test_invalid_and_empty(oracle):
    s = [np.nan, np.inf, -np.inf, 0.0, -0.0, -3.0, 5e-324, 1.0]
    assert oracle.top_k(s, 10) == [(7, 1.0), (6, 5e-324)]
    assert oracle.top_k(s, 0) == []
    assert oracle.top_k([], 3) == []


def test_planted_smallest_index_rule_is_caught():
    """Keeping the smallest indices at a tie (a plausible simplification)
    disagrees with the reference's rule on [1, 1, 1, 2], k = 2."""
    assert described([1.0, 1.0, 1.0, 2.0], 2) != [(3, 2.0), (0, 1.0)]
