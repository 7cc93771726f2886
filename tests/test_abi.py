"""C-ABI library checks that need no GPU: it loads, exports every declared
symbol, and its host-side logic (generators) behaves."""
import ctypes
import os
import re

import numpy as np

from graphops_amd import _lib as L
from graphops_amd import generators

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "graphops_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gops_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    # the ctypes table covers the header exactly
    assert sorted(name for name, _, _ in L.SIGNATURES) == syms


def test_struct_layouts():
    assert ctypes.sizeof(L.PageRankConfigC) == 24
    assert ctypes.sizeof(L.RunInfoC) == 24
    assert ctypes.sizeof(L.RunStatsC) == 88
    assert ctypes.sizeof(L.RowPlanC) == 32


def test_rmat_properties_and_determinism():
    a = generators.rmat(12, 8, seed=3, weights="uniform", p0=0.1, p1=1.0, rounding="f32")
    b = generators.rmat(12, 8, seed=3, weights="uniform", p0=0.1, p1=1.0, rounding="f32")
    assert a.n == 4096 and 0 < a.m <= 8 * 4096
    assert np.array_equal(a.offsets, b.offsets) and np.array_equal(a.targets, b.targets)
    assert np.array_equal(a.weights, b.weights)
    for u in range(a.n):
        row = a.targets[a.offsets[u]:a.offsets[u + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)  # sorted, unique
        assert u not in row  # no self-loops
    assert a.weights.min() >= 0.1 and a.weights.max() < 1.0
    assert np.array_equal(a.weights, a.weights.astype(np.float32).astype(np.float64))
    c = generators.rmat(12, 8, seed=4)
    assert not np.array_equal(a.targets[:100], c.targets[:100])
    assert c.weights is None
    # skew: low ids (quadrant a) dominate
    indeg = np.bincount(a.targets, minlength=a.n)
    assert indeg[0] > 20 * max(1.0, indeg.mean())


def test_bf16_lognormal_weights_are_bf16_exact():
    e = generators.rmat(10, 16, seed=11, weights="lognormal", p0=0.0, p1=1.0, rounding="bf16")
    w = e.weights
    bits = w.astype(np.float32).view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)
    assert np.all(w.astype(np.float32).astype(np.float64) == w)
    lw = np.log(w)
    assert abs(lw.mean()) < 0.05 and abs(lw.std() - 1.0) < 0.05


def test_erdos_renyi_density():
    e = generators.erdos_renyi(2000, 0.005, 42)
    assert abs(e.m - 0.005 * 2000 * 1999) < 6 * np.sqrt(0.005 * 2000 * 1999)
    for u in range(0, 2000, 97):
        row = e.targets[e.offsets[u]:e.offsets[u + 1]]
        assert u not in row and np.all(np.diff(row.astype(np.int64)) > 0)


def test_seed_sets_distinct_and_in_range():
    e = generators.rmat(12, 16, seed=13)
    off, seeds = generators.seed_sets(e, 64, 1, 64, seed=5)
    sizes = np.diff(off)
    assert sizes.min() >= 1 and sizes.max() <= 64
    deg = np.diff(e.offsets)
    for q in range(64):
        s = seeds[off[q]:off[q + 1]]
        assert len(set(s.tolist())) == len(s)
        assert np.all(deg[s] > 0)  # drawn as sources of edges
    off2, seeds2 = generators.seed_sets(e, 64, 1, 256, log_uniform=True, seed=5)
    assert np.diff(off2).max() <= 256
