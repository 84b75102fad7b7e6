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
