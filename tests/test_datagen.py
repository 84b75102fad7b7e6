"""Seeded generators (CPU): determinism, sizes, sortedness and the exact planted shared counts."""
import numpy as np

import oracle
from simd_set_operations_amd import datagen


def test_splitmix64_known_values():
    # SplitMix64 from seed 0 (published reference values of the algorithm)
    v = datagen.splitmix64(0, 3)
    assert v.tolist() == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_plain_pair_shape():
    a, b = datagen.plain_pair()
    assert a.size == b.size == 1 << 16
    for s in (a, b):
        assert np.all(np.diff(s.astype(np.int64)) > 0)
        assert s.max() < 1 << 20
    # exactly 19,661 shared elements (planted first, the rest disjoint)
    assert np.intersect1d(a, b).size == 19661
    assert oracle.count2("branchless_merge", a, b) == 19661
    a2, b2 = datagen.plain_pair()
    assert np.array_equal(a, a2) and np.array_equal(b, b2)


def test_uneven_batch_small():
    A, a_off, B, b_off, shared = datagen.uneven_batch(npairs=512, seed=77, with_shared=True)
    sa, sb = np.diff(a_off), np.diff(b_off)
    assert np.all(sb >= 64) and np.all(sb <= 1 << 14)
    assert np.all(sa >= 1) and np.all(sa <= sb)
    # S = max(1, floor(L / r)) with r <= 1024, so S >= max(1, floor(L / 1024))
    assert np.all(sa >= np.maximum(1, sb // 1024))
    counts, _, _ = oracle.batch_intersect("branchless_merge", A, a_off, B, b_off, nthreads=4, with_output=False)
    assert np.array_equal(counts, shared)
    for p in range(0, 512, 37):
        a = A[a_off[p]:a_off[p + 1]]
        b = B[b_off[p]:b_off[p + 1]]
        assert np.all(np.diff(a.astype(np.int64)) > 0) and np.all(np.diff(b.astype(np.int64)) > 0)
        assert b.max() < 10 * b.size + a.size + 1
    A2, a_off2, B2, b_off2 = datagen.uneven_batch(npairs=512, seed=77)
    assert np.array_equal(A, A2) and np.array_equal(B, B2)


def test_rmat_graph_is_a_simple_symmetric_csr():
    from simd_set_operations_amd import datagen

    rp, col = datagen.rmat_graph(scale=12)
    n = rp.size - 1
    d = np.diff(rp).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), d)
    assert np.all(src != col)  # no self loops
    key = src * n + col
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates
    assert np.array_equal(key, np.sort(col.astype(np.int64) * n + src))  # symmetric
    # R-MAT with a = 0.57 is skewed: the top vertex has far more than the mean degree
    assert d.max() > 10 * d.mean()
    up = datagen.upper_starts(rp, col)
    for u in range(0, n, 97):
        lst = col[rp[u]:rp[u + 1]]
        assert up[u] - rp[u] == np.searchsorted(lst, u, side="right")


def test_csr_to_bsr_matches_oracle_encoding():
    import oracle
    from simd_set_operations_amd import datagen
    from simd_set_operations_amd.graph import csr_to_bsr

    rp, col = datagen.rmat_graph(scale=11)
    brp, bases, states = csr_to_bsr(rp, col)
    for u in range(rp.size - 1):
        eb, es = oracle.bsr_from_sorted(col[rp[u]:rp[u + 1]])
        assert np.array_equal(bases[brp[u]:brp[u + 1]], eb), u
        assert np.array_equal(states[brp[u]:brp[u + 1]], es), u


def test_graph_triangle_identity_on_the_oracle():
    """Sum over edges u<v of |N(u) & N(v)| is three times the oriented count (CPU oracle)."""
    import oracle
    from simd_set_operations_amd import datagen

    rp, col = datagen.rmat_graph(scale=10)
    u, v = datagen.graph_edges(rp, col)
    up = datagen.upper_starts(rp, col)
    full, _, _ = oracle.batch_intersect_gather("branchless_merge", col, rp[u], rp[u + 1], col, rp[v], rp[v + 1],
                                               nthreads=4, with_output=False)
    ori, _, _ = oracle.batch_intersect_gather("branchless_merge", col, up[u], rp[u + 1], col, up[v], rp[v + 1],
                                              nthreads=4, with_output=False)
    s, t = int(full.sum()), int(ori.sum())
    assert s == 3 * t and t > 0
    # brute force on the adjacency matrix: trace(A^3) / 6
    n = rp.size - 1
    A = np.zeros((n, n), np.int64)
    A[np.repeat(np.arange(n), np.diff(rp).astype(np.int64)), col] = 1
    assert np.trace(A @ A @ A) // 6 == t


def test_served_workload_generators():
    from types import SimpleNamespace

    from simd_set_operations_amd import serving

    lens = serving.zipf_lengths(1 << 16, 16, 1 << 18, 1.1, seed=5)
    assert lens.min() >= 16 and lens.max() <= 1 << 18
    # bounded Zipf(1.1) on [16, 2^18]: the mean follows from the pmf
    sup = np.arange(16, (1 << 18) + 1, dtype=np.float64)
    w = sup ** -1.1
    mean = (sup * w).sum() / w.sum()
    assert abs(lens.mean() / mean - 1) < 0.1
    assert np.mean(lens == 16) > np.mean(lens == 17) > np.mean(lens == 64)
    rk = serving.zipf_lengths(1000, 16, 1 << 18, 1.1, seed=5, mode="rank")
    assert rk.max() == 1 << 18 and rk.min() == max(16, int(np.floor((1 << 18) * 1000.0 ** -1.1)))
    c = SimpleNamespace(nlists=1000, lens=lens[:1000])
    ids, ks = serving.make_queries(c, 500, seed=9)
    assert ks.min() >= 2 and ks.max() <= 6 and set(np.unique(ks)) == {2, 3, 4, 5, 6}
    for q in range(500):
        row = ids[q, :ks[q]]
        assert np.all(ids[q, ks[q]:] == -1) and np.unique(row).size == row.size
        assert np.all(np.diff(c.lens[row].astype(np.int64)) >= 0)  # smallest first
    t = serving.poisson_arrivals(20000, 1000.0, seed=3)
    assert np.all(np.diff(t) > 0) and abs(t[-1] / 20.0 - 1) < 0.05
