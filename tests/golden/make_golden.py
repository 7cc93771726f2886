"""Writes the golden fixtures under tests/golden/ (run in the build container).

Everything here is computed from the mathematics of the reference's Markov
chain, independently of the oracle and of the product kernels:

* closed forms: exact rational fixed points (Fraction Gaussian elimination) of
  every graph in the reference's own tests, docs and example
  (pagerank.rs:100-113, 344-377; ppr.rs:50-65; tests/invariants.rs:20-50;
  examples/pagerank.rs:14-19);
* exact trajectories: the reference's iteration (pagerank.rs:122-187,
  210-289; ppr.rs:90-156) in exact rational arithmetic on tiny graphs, which
  fixes the iteration count / residual / converged semantics without any
  rounding;
* dense solves: numpy solve of (I - d M^T - d q 1_dangling^T) x = (1-d) q on
  random graphs with duplicates, self-loops, dangling nodes, weights;
* networkx.pagerank fixed points (an independent implementation) on simple
  graphs, weighted and personalized;
* the C1 workload (directed G(2000, 0.005), seed 42): a hash of its
  neighbour lists and its exact fixed point by dense solve.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


# ---------------------------------------------------------------- exact math
def solve_fraction(A, b):
    n = len(b)
    M = [list(A[i]) + [b[i]] for i in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        piv = M[c][c]
        M[c] = [x / piv for x in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] for i in range(n)]


def chain(n, nbrs, weights=None):
    """Transition shares T[u][v] (sum over neighbour entries of u) and the
    dangling mask, for the unweighted (weights None) or weighted chain."""
    T = [[Fraction(0)] * n for _ in range(n)]
    dangling = [False] * n
    for u in range(n):
        if weights is None:
            deg = len(nbrs[u])
            if deg == 0:
                dangling[u] = True
                continue
            for v in nbrs[u]:
                T[u][v] += Fraction(1, deg)
        else:
            ws = sum((max(Fraction(w), Fraction(0)) for w in weights[u]), Fraction(0))
            if ws == 0:
                dangling[u] = True
                continue
            for v, w in zip(nbrs[u], weights[u]):
                w = max(Fraction(w), Fraction(0))
                if w > 0:
                    T[u][v] += w / ws
    return T, dangling


def fixed_point_exact(n, nbrs, d, weights=None, p=None):
    d = Fraction(d)
    T, dangling = chain(n, nbrs, weights)
    if p is None:
        q = [Fraction(1, n)] * n
    else:
        s = sum(Fraction(x) for x in p)
        q = [Fraction(x) / s for x in p] if s > 0 else [Fraction(1, n)] * n
    A = [[Fraction(0)] * n for _ in range(n)]
    for v in range(n):
        for u in range(n):
            a = (Fraction(1) if u == v else Fraction(0)) - d * T[u][v]
            if dangling[u]:
                a -= d * q[v]
            A[v][u] = a
    b = [(1 - d) * q[v] for v in range(n)]
    return solve_fraction(A, b)


def trajectory_exact(n, nbrs, d, max_iterations, tol, weights=None, p=None):
    """The reference loop in exact arithmetic: returns iterations, the exact
    last residual, converged, and the last iterate."""
    d = Fraction(d)
    tol = Fraction(tol)
    T, dangling = chain(n, nbrs, weights)
    if p is None:
        q = [Fraction(1, n)] * n
        x = list(q)
    else:
        s = sum(Fraction(v) for v in p)
        q = [Fraction(v) / s for v in p] if s > 0 else [Fraction(1, n)] * n
        x = list(q)
    last = None
    it = 0
    conv = False
    for _ in range(max_iterations):
        it += 1
        dsum = sum((x[i] for i in range(n) if dangling[i]), Fraction(0))
        new = [(1 - d) * q[v] + d * dsum * q[v] for v in range(n)]
        for u in range(n):
            for v in range(n):
                if T[u][v]:
                    new[v] += d * x[u] * T[u][v]
        diff = sum((abs(a - b) for a, b in zip(x, new)), Fraction(0))
        last = diff
        x = new
        if diff < tol:
            conv = True
            break
    return it, last, conv, x


def ffmt(fr):
    return f"{fr.numerator}/{fr.denominator}"


# ---------------------------------------------------------------- graphs
def dense_to_lists(mat):
    """AdjacencyMatrix adapter, graph.rs:128-141."""
    nb, ws = [], []
    for row in mat:
        nb.append([j for j, w in enumerate(row) if w > 0.0])
        ws.append([w for w in row if w > 0.0])
    return nb, ws


CYCLE3 = [[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]
EXAMPLE4 = [[0.0, 1.0, 1.0, 0.0], [0.0, 0.0, 1.0, 0.0], [1.0, 0.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0]]
WEIGHTED3 = [[0.0, 2.0, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]]
HEAVIER3 = [[0.0, 2.0, 1.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
LINE3 = [[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]]
EMPTY4 = [[0.0] * 4 for _ in range(4)]


def closed_forms():
    cases = []

    def add(name, source, mat, weighted=False, p=None):
        nb, ws = dense_to_lists(mat)
        x = fixed_point_exact(len(mat), nb, Fraction(85, 100), ws if weighted else None, p)
        cases.append({"name": name, "source": source, "matrix": mat, "weighted": weighted,
                      "personalization": p, "damping": 0.85,
                      "fixed_point": [ffmt(v) for v in x]})

    add("cycle3", "pagerank.rs:104-108", CYCLE3)
    add("example4", "examples/pagerank.rs:14-19", EXAMPLE4)
    add("weighted3_dangling", "pagerank.rs:350-354", WEIGHTED3, weighted=True)
    add("heavier_edge", "pagerank.rs:364-368", HEAVIER3, weighted=True)
    add("line3", "tests/invariants.rs:22-28", LINE3)
    add("line3_ppr_e0", "tests/invariants.rs:45-47", LINE3, p=[1.0, 0.0, 0.0])
    add("cycle3_ppr_e0", "ppr.rs:59-64", CYCLE3, p=[1.0, 0.0, 0.0])
    add("empty4", "no edges", EMPTY4)
    return cases


def trajectories():
    out = []
    rng = np.random.default_rng(1234)
    specs = [
        ("example4", EXAMPLE4, False, None, 0.85, 100, 1e-8),
        ("line3", LINE3, False, None, 0.85, 100, 1e-6),
        ("line3_ppr", LINE3, False, [1.0, 0.0, 0.0], 0.85, 100, 1e-6),
        ("weighted3", WEIGHTED3, True, None, 0.85, 100, 1e-9),
        ("cycle3", CYCLE3, False, None, 0.85, 100, 1e-6),
        ("empty4", EMPTY4, False, None, 0.85, 100, 1e-6),
        ("example4_cap", EXAMPLE4, False, None, 0.85, 5, 1e-12),
        ("example4_d0", EXAMPLE4, False, None, 0.0, 100, 1e-6),
        ("example4_d1", EXAMPLE4, False, None, 1.0, 50, 1e-6),
    ]
    for k in range(4):
        n = int(rng.integers(3, 7))
        mat = [[0.0] * n for _ in range(n)]
        for _ in range(int(rng.integers(n, 3 * n))):
            u, v = int(rng.integers(0, n)), int(rng.integers(0, n))
            mat[u][v] = float(rng.choice([0.5, 1.0, 2.0, 3.0]))
        specs.append((f"random{k}", mat, bool(k % 2), None, 0.85, 200, 1e-9))
    specs.append(("random_ppr", specs[-1][1], False,
                  [0.0] * (len(specs[-1][1]) - 1) + [2.0], 0.85, 200, 1e-9))
    for name, mat, weighted, p, d, mi, tol in specs:
        nb, ws = dense_to_lists(mat)
        it, diff, conv, x = trajectory_exact(len(mat), nb, Fraction(d), mi, Fraction(tol),
                                             ws if weighted else None, p)
        out.append({"name": name, "matrix": mat, "weighted": weighted, "personalization": p,
                    "damping": d, "max_iterations": mi, "tolerance": tol, "iterations": it,
                    "diff_l1": float(diff), "converged": conv, "scores": [float(v) for v in x]})
    return out


def dense_solve(n, nbrs, d, ew=None, p=None):
    """numpy solve of the fixed point for neighbour lists with duplicates."""
    T = np.zeros((n, n))
    dangling = np.zeros(n, dtype=bool)
    for u in range(n):
        if ew is None:
            deg = len(nbrs[u])
            if deg == 0:
                dangling[u] = True
            for v in nbrs[u]:
                T[u, v] += 1.0 / deg
        else:
            w = np.maximum(np.asarray(ew[u], dtype=float), 0.0) if len(ew[u]) else np.zeros(0)
            ws = float(w.sum())
            if ws == 0.0:
                dangling[u] = True
                continue
            for v, wi in zip(nbrs[u], w):
                T[u, v] += wi / ws
    if p is None:
        q = np.full(n, 1.0 / n)
    else:
        p = np.asarray(p, dtype=float)
        q = p / p.sum() if p.sum() > 0 else np.full(n, 1.0 / n)
    A = np.eye(n) - d * T.T - d * np.outer(q, dangling.astype(float))
    return np.linalg.solve(A, (1.0 - d) * q)


def random_lists(rng, n, m, dup=True, loops=True):
    nbrs = [[] for _ in range(n)]
    for _ in range(m):
        u, v = int(rng.integers(0, n)), int(rng.integers(0, n))
        if not loops and u == v:
            continue
        if not dup and v in nbrs[u]:
            continue
        nbrs[u].append(v)
    return nbrs


def dense_solve_cases():
    rng = np.random.default_rng(20240607)
    out = []
    for k in range(6):
        n = int(rng.integers(5, 60))
        nbrs = random_lists(rng, n, int(rng.integers(n, 5 * n)))
        for u in rng.choice(n, size=max(1, n // 6), replace=False):
            nbrs[int(u)] = []  # force dangling nodes
        kind = ["plain", "weighted", "ppr"][k % 3]
        ew = None
        p = None
        if kind == "weighted":
            # per-entry weights, with zero / negative ones that the chain clamps
            ew = [[float(rng.choice([0.0, -1.0, 0.25, 1.0, 3.5])) for _ in lst] for lst in nbrs]
        if kind == "ppr":
            p = [float(v) if rng.random() < 0.2 else 0.0 for v in rng.random(n)]
            if sum(p) == 0:
                p[0] = 1.0
        x = dense_solve(n, nbrs, 0.85, ew, p)
        out.append({"name": f"dense_{kind}_{k}", "kind": kind, "n": n, "neighbors": nbrs,
                    "edge_weights": ew, "personalization": p, "damping": 0.85,
                    "fixed_point": x.tolist()})
    return out


def networkx_cases():
    import networkx as nx

    rng = np.random.default_rng(99)
    out = []
    for k in range(3):
        n = int(rng.integers(20, 80))
        nbrs = random_lists(rng, n, 4 * n, dup=False, loops=False)
        G = nx.DiGraph()
        G.add_nodes_from(range(n))
        ew = [[float(rng.uniform(0.1, 2.0)) for _ in lst] for lst in nbrs]
        for u in range(n):
            for v, w in zip(nbrs[u], ew[u]):
                G.add_edge(u, v, weight=w)
        kind = ["plain", "weighted", "ppr"][k]
        kw = dict(alpha=0.85, tol=1e-15, max_iter=10000)
        p = None
        if kind == "plain":
            x = nx.pagerank(G, weight=None, **kw)
        elif kind == "weighted":
            x = nx.pagerank(G, weight="weight", **kw)
        else:
            p = [0.0] * n
            for s in rng.choice(n, size=5, replace=False):
                p[int(s)] = 1.0
            x = nx.pagerank(G, weight=None, personalization={i: p[i] for i in range(n)}, **kw)
        out.append({"name": f"networkx_{kind}", "kind": kind, "n": n, "neighbors": nbrs,
                    "edge_weights": ew if kind == "weighted" else None, "personalization": p,
                    "damping": 0.85, "fixed_point": [x[i] for i in range(n)]})
    return out


def c1_case():
    """Directed G(2000, 0.005), seed 42, via the library's generator (host
    code only), and its fixed point by dense solve."""
    sys.path.insert(0, ROOT)
    from graphops_amd import generators

    e = generators.erdos_renyi(2000, 0.005, 42)
    off, nb = e.offsets, e.targets
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, dtype=np.uint64).tobytes())
    h.update(np.ascontiguousarray(nb, dtype=np.uint32).tobytes())
    nbrs = [nb[off[u]:off[u + 1]].tolist() for u in range(2000)]
    x = dense_solve(2000, nbrs, 0.85)
    np.save(os.path.join(HERE, "c1_er2000_fixed_point.npy"), x)
    return {"n": 2000, "m": int(off[-1]), "sha256": h.hexdigest(),
            "dangling": int(sum(1 for lst in nbrs if not lst))}


def main():
    data = {
        "closed_forms": closed_forms(),
        "trajectories": trajectories(),
        "dense_solves": dense_solve_cases(),
        "networkx": networkx_cases(),
        "c1": c1_case(),
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
