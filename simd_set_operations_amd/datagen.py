"""Seeded synthetic workloads (wrappers over libsetops_datagen.so, SplitMix64 streams).

The workloads are the benchmark configurations of this path:

* ``plain_pair``   -- one pair of 2^16 + 2^16 sorted unique u32 over {0..2^20-1} with exactly
  19,661 shared elements (the reference's CPU-runnable 2-set case).
* ``uneven_batch`` -- 65,536 pairs; larger size log-uniform in [2^6, 2^14], smaller = larger / r
  with r log-uniform in [1, 1024]; element space = 10 x larger; shared = round(sel x smaller),
  sel ~ U[0,1].

* ``rmat_graph``   -- R-MAT (a, b, c, d) = (0.57, 0.19, 0.19, 0.05), 2^22 vertices, 16 x 2^22
  sampled edges, symmetrized and deduplicated, as a CSR of sorted u32 neighbour lists.
* ``skewed``       -- one resident large set of 2^24 over {0..2^28-1} against 4,096 small sets of
  2^8, each with round(0.1 x 256) = 26 elements drawn from the large set and the rest from its
  complement.

All follow the reference generator's planted-shared-then-disjoint semantics
(benchmark/src/generators.rs:30-110).
"""
import ctypes

import numpy as np

from . import _native

_c = ctypes

PLAIN_PAIR = dict(universe=1 << 20, size=1 << 16, shared=19661, seed=1001)
SKEWED = dict(universe=1 << 28, large=1 << 24, nsmall=4096, small=256, shared=26, seed=1003)
RMAT = dict(scale=22, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1004)
UNEVEN = dict(npairs=65536, lmin=6.0, lmax=14.0, log2_skew=10.0, density=0.1, seed=1002)


def _p(a):
    return a.ctypes.data_as(_c.c_void_p)


def splitmix64(seed, n):
    out = np.zeros(n, np.uint64)
    _native.datagen().setops_gen_splitmix64(seed, _p(out), n)
    return out


def planted_pair(seed, universe, shared, only_a, only_b):
    a = np.zeros(shared + only_a, np.uint32)
    b = np.zeros(shared + only_b, np.uint32)
    rc = _native.datagen().setops_gen_planted_pair(seed, universe, shared, only_a, only_b,
                                                   _p(a) if a.size else None, _p(b) if b.size else None)
    if rc != 0:
        raise ValueError("planted pair does not fit in its element space")
    return a, b


def plain_pair(seed=PLAIN_PAIR["seed"]):
    """The 1-pair workload: |A| = |B| = 2^16 over {0..2^20-1}, 19,661 shared."""
    n, s = PLAIN_PAIR["size"], PLAIN_PAIR["shared"]
    return planted_pair(seed, PLAIN_PAIR["universe"], s, n - s, n - s)


def uneven_batch(npairs=UNEVEN["npairs"], seed=UNEVEN["seed"], lmin=UNEVEN["lmin"], lmax=UNEVEN["lmax"],
                 log2_skew=UNEVEN["log2_skew"], density=UNEVEN["density"], with_shared=False):
    """Uneven CSR batch: returns (A, a_off, B, b_off[, shared]) with A = the smaller sets."""
    lib = _native.datagen()
    a_len = np.zeros(npairs, np.uint64)
    b_len = np.zeros(npairs, np.uint64)
    lib.setops_gen_uneven_sizes(seed, npairs, lmin, lmax, log2_skew, density, _p(a_len), _p(b_len))
    a_off = np.zeros(npairs + 1, np.uint64)
    b_off = np.zeros(npairs + 1, np.uint64)
    np.cumsum(a_len, out=a_off[1:])
    np.cumsum(b_len, out=b_off[1:])
    A = np.zeros(int(a_off[-1]), np.uint32)
    B = np.zeros(int(b_off[-1]), np.uint32)
    shared = np.zeros(npairs, np.uint32)
    rc = lib.setops_gen_uneven_fill(seed, npairs, lmin, lmax, log2_skew, density, _p(a_off), _p(b_off), _p(A), _p(B),
                                    _p(shared))
    if rc != 0:
        raise RuntimeError("uneven batch generation failed")
    if with_shared:
        return A, a_off, B, b_off, shared
    return A, a_off, B, b_off


def fixed_batch(npairs, larger, smaller, selectivity, density=0.1, seed=1006):
    """Fixed-shape CSR batch: ``npairs`` pairs of ``smaller`` vs ``larger`` sorted unique u32 over
    an element space of ``larger / density``, each with round(selectivity x smaller) shared
    elements planted first.  Returns (A, a_off, B, b_off, shared)."""
    universe = max(int(np.ceil(larger / density)), larger + smaller)
    shared = int(round(selectivity * smaller))
    A = np.zeros(npairs * smaller, np.uint32)
    B = np.zeros(npairs * larger, np.uint32)
    rc = _native.datagen().setops_gen_fixed_batch(seed, npairs, larger, smaller, universe, shared, _p(A), _p(B))
    if rc != 0:
        raise ValueError("fixed batch does not fit its element space")
    a_off = np.arange(npairs + 1, dtype=np.uint64) * np.uint64(smaller)
    b_off = np.arange(npairs + 1, dtype=np.uint64) * np.uint64(larger)
    return A, a_off, B, b_off, shared


def skewed(seed=SKEWED["seed"], universe=SKEWED["universe"], large=SKEWED["large"], nsmall=SKEWED["nsmall"],
           small=SKEWED["small"], shared=SKEWED["shared"]):
    """Skewed batch: (large[large], smalls[nsmall, small]) -- every small set has exactly `shared`
    elements of the large set."""
    L = np.zeros(large, np.uint32)
    S = np.zeros((nsmall, small), np.uint32)
    rc = _native.datagen().setops_gen_skewed(seed, universe, large, nsmall, small, shared, _p(L), _p(S))
    if rc != 0:
        raise ValueError("skewed workload does not fit its element space")
    return L, S


def rmat_graph(scale=RMAT["scale"], edge_factor=RMAT["edge_factor"], a=RMAT["a"], b=RMAT["b"], c=RMAT["c"],
               seed=RMAT["seed"]):
    """R-MAT graph as CSR: (rowptr[n+1] u64, col[nnz] u32), lists sorted and duplicate-free,
    no self loops, symmetric."""
    n = 1 << scale
    m = edge_factor << scale
    rowptr = np.zeros(n + 1, np.uint64)
    col = np.zeros(2 * m, np.uint32)
    nnz = _native.datagen().setops_gen_rmat(seed, scale, edge_factor, a, b, c, _p(rowptr), _p(col))
    if nnz < 0:
        raise ValueError("bad R-MAT parameters")
    return rowptr, col[:nnz].copy()


def graph_edges(rowptr, col):
    """Edges (u, v) with u < v of a symmetric CSR, in CSR order."""
    n = rowptr.size - 1
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rowptr).astype(np.int64))
    keep = src < col
    return src[keep], col[keep]


def upper_starts(rowptr, col):
    """For every vertex u, the index in col of its first neighbour greater than u (the start of
    the oriented list N+(u))."""
    n = rowptr.size - 1
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(rowptr).astype(np.int64))
    keys = (src << np.uint64(32)) | col.astype(np.uint64)
    probe = (np.arange(n, dtype=np.uint64) << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    return np.searchsorted(keys, probe, side="right").astype(np.uint64)
