"""Pins the oracle (CPU, no GPU): every oracle function is checked against something other than
its own formula -- the reference's known answers, the reference's compiled qfilter.cpp, the
committed golden vectors, numpy.intersect1d, brute force on tiny universes and the reference's
invariant trio (setops/tests/testlib/properties.rs:2-49)."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pairs.npz")
TWOSET = ["naive_merge", "branchless_merge", "galloping", "binary_search", "shuffling_sse", "shuffling_avx2",
          "shuffling_avx512", "qfilter", "broadcast_sse", "galloping_sse", "galloping_avx512"]
BSR_KINDS = ["branchless_merge_bsr", "galloping_bsr", "shuffling_bsr", "qfilter_bsr", "broadcast_bsr",
             "simd_galloping_bsr"]
WIDTHS = {"shuffling_bsr": (4, 8, 16), "broadcast_bsr": (4, 8, 16), "simd_galloping_bsr": (4, 8, 16)}


def golden():
    return np.load(GOLDEN)


def golden_pairs(g):
    i = 0
    while f"pair{i}_a" in g:
        yield i, g[f"pair{i}_a"], g[f"pair{i}_b"]
        i += 1


def sorted_set(rng, n, universe):
    return np.unique(rng.integers(0, universe, size=n)).astype(np.uint32)


# -- properties.rs:2-49 --------------------------------------------------------------------------
def prop_intersection_correct(result, sets):
    result = np.asarray(result, np.int64)
    increasing = bool(np.all(np.diff(result) > 0))
    all_common = all(np.isin(result, np.asarray(s, np.int64)).all() for s in sets)
    common = np.asarray(sets[0], np.int64)
    for s in sets[1:]:
        common = common[np.isin(common, np.asarray(s, np.int64))]
    complete = bool(np.isin(common, result).all())
    return increasing and all_common and complete


@pytest.mark.parametrize("algo", TWOSET)
def test_known_answers(algo):
    """setops/tests/unit_tests.rs:6-59 through every oracle 2-set function."""
    g = golden()
    for i in range(6):
        got = oracle.intersect2(algo, g[f"kat{i}_a"], g[f"kat{i}_b"])
        assert np.array_equal(got, g[f"kat{i}_out"]), (algo, i)


def test_const_intersect_answer():
    # unit_tests.rs:28-33: const_intersect gives [1,3,9,10,15,20]
    a = [1, 3, 5, 8, 9, 10, 14, 15, 18, 20]
    b = [1, 2, 3, 4, 9, 10, 11, 15, 20, 21]
    assert oracle.branchless_merge(a, b).tolist() == [1, 3, 9, 10, 15, 20]


@pytest.mark.parametrize("algo", TWOSET)
def test_golden_vectors(algo):
    """Outputs of the reference's compiled intersect_qfilter_uint_b4_v2 (tests/golden/)."""
    g = golden()
    for i, a, b in golden_pairs(g):
        # galloping-family functions take (small, large); orientation changes cost, not output
        got = oracle.intersect2(algo, a, b)
        assert np.array_equal(got, g[f"pair{i}_out"]), (algo, i)


@pytest.mark.parametrize("algo", TWOSET)
def test_against_numpy(algo):
    rng = np.random.Generator(np.random.PCG64(7))
    for trial in range(60):
        la, lb = rng.integers(0, 3000, size=2)
        u = int(rng.integers(1, 20000))
        a, b = sorted_set(rng, la, u), sorted_set(rng, lb, u)
        exp = np.intersect1d(a, b).astype(np.uint32)
        assert np.array_equal(oracle.intersect2(algo, a, b), exp), (algo, trial)
        assert oracle.count2(algo, a, b) == exp.size


@pytest.mark.parametrize("algo", TWOSET)
def test_brute_force_tiny(algo):
    """Every pair of subsets of a 6-element universe: membership by brute force."""
    u = 6
    for ma in range(1 << u):
        a = np.array([x for x in range(u) if ma >> x & 1], np.uint32)
        for mb in range(0, 1 << u, 3):
            b = np.array([x for x in range(u) if mb >> x & 1], np.uint32)
            exp = [x for x in range(u) if (ma & mb) >> x & 1]
            assert oracle.intersect2(algo, a, b).tolist() == exp


@pytest.mark.parametrize("algo", TWOSET)
def test_invariant_trio_skewed(algo):
    # testlib/mod.rs:123-153 SkewedSetPair shape: small 1..128, large 128..8319
    rng = np.random.Generator(np.random.PCG64(11))
    for _ in range(40):
        small = sorted_set(rng, int(rng.integers(1, 128)), 20000)
        large = np.union1d(sorted_set(rng, int(rng.integers(128, 8319)), 20000), small[: small.size // 2])
        large = large.astype(np.uint32)
        r = oracle.intersect2(algo, small, large)
        assert prop_intersection_correct(r, [small, large])


@pytest.mark.parametrize("algo", ["galloping_sse", "galloping_avx512"])
def test_simd_galloping_block_walk(algo):
    """simd_galloping.rs:64-107: the probe side passes every full block of the larger side while it
    is itself at least a block long (the two sides swap), probes that land in each quarter of a
    block, and a larger side shorter than a block (the merge tail only)."""
    cases = [
        (np.arange(5000, 5700), np.concatenate([np.arange(0, 1000), np.arange(5100, 5110)])),
        (np.arange(0, 40000, 7), np.arange(0, 40000, 3)),
        (np.array([3, 130, 260, 390, 511]), np.arange(0, 512)),
        (np.array([5, 9, 100]), np.arange(0, 100, 2)),
        (np.array([], np.uint32), np.arange(0, 1000)),
    ]
    for a, b in cases:
        a, b = np.asarray(a, np.uint32), np.asarray(b, np.uint32)
        exp = np.intersect1d(a, b).astype(np.uint32)
        assert np.array_equal(oracle.intersect2(algo, a, b), exp)
        assert np.array_equal(oracle.intersect2(algo, b, a), exp)
    # per-quarter probes against a brute-force membership count over one block of 32 W elements
    big = np.arange(0, 4096, 2, dtype=np.uint32)
    probes = np.arange(0, 4096, 5, dtype=np.uint32)
    assert oracle.count2(algo, probes, big) == sum(1 for x in probes.tolist() if x % 2 == 0)


def test_binary_search_semantics():
    # galloping.rs:110-131: index of the target, else the insertion point lo
    s = np.array([2, 4, 6, 8], np.uint32)
    L = oracle.lib()
    p = s.ctypes.data
    assert L.oracle_binary_search(p, 6, 0, 3) == 2
    assert L.oracle_binary_search(p, 5, 0, 3) == 2
    assert L.oracle_binary_search(p, 1, 0, 3) == 0
    assert L.oracle_binary_search(p, 9, 0, 3) == 4


def test_qfilter_dictionary():
    # qfilter.rs:254-357 classification: none / single offsets / multi
    cls = oracle.lib().oracle_qfilter_classify
    assert cls(0) == -2
    assert cls(0xFFFF) == -1
    assert cls(0b0011) == -1  # lane 0 matches b0 and b1
    # lane 0 -> b2, lane 1 none (points at itself), lane 2 -> b0, lane 3 none
    m = (1 << 2) | (1 << 8)
    assert cls(m) == (2 | (1 << 2) | (0 << 4) | (3 << 6))


@pytest.mark.skipif(not oracle.ref_available(), reason="reference qfilter build absent")
def test_against_reference_qfilter():
    """The oracle's qfilter (and every other 2-set function) equals the reference's compiled
    intersect_qfilter_uint_b4 and _v2 on fresh seeded inputs."""
    rng = np.random.Generator(np.random.PCG64(99))
    for _ in range(80):
        la, lb = rng.integers(0, 5000, size=2)
        u = int(rng.integers(1, 1 << 16))
        a, b = sorted_set(rng, la, u), sorted_set(rng, lb, u)
        r2 = oracle.ref_qfilter(a, b, v2=True)
        r1 = oracle.ref_qfilter(a, b, v2=False)
        assert np.array_equal(r1, r2)
        for algo in TWOSET:
            assert np.array_equal(oracle.intersect2(algo, a, b), r2)


def test_bsr_round_trip():
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(50):
        v = sorted_set(rng, int(rng.integers(0, 4000)), int(rng.integers(1, 100000)))
        b, s = oracle.bsr_from_sorted(v)
        assert np.all(np.diff(b.astype(np.int64)) > 0)
        assert np.all(s != 0)
        assert int(oracle.lib().oracle_bsr_count(s.ctypes.data, s.size)) == v.size
        assert np.array_equal(oracle.bsr_to_sorted(b, s), v)


def test_bsr_known_encoding():
    # bsr.rs:140-164 by hand: 1, 3 -> base 0 state 0b1010; 32 -> base 1 state 1; 95 -> base 2 bit 31
    b, s = oracle.bsr_from_sorted([1, 3, 32, 95])
    assert b.tolist() == [0, 1, 2]
    assert s.tolist() == [0b1010, 1, 1 << 31]


@pytest.mark.parametrize("kind", BSR_KINDS)
def test_bsr_intersections(kind):
    """decode(bsr_a ∩ bsr_b) == a ∩ b; word list equals from_sorted(a ∩ b)."""
    rng = np.random.Generator(np.random.PCG64(5))
    for _ in range(40):
        u = int(rng.integers(64, 40000))
        a, b = sorted_set(rng, int(rng.integers(0, 3000)), u), sorted_set(rng, int(rng.integers(0, 3000)), u)
        ba, sa = oracle.bsr_from_sorted(a)
        bb, sb = oracle.bsr_from_sorted(b)
        for w in WIDTHS.get(kind, (4,)):
            ob, os_ = oracle.bsr_intersect(kind, ba, sa, bb, sb, width=w)
            exp = np.intersect1d(a, b).astype(np.uint32)
            eb, es = oracle.bsr_from_sorted(exp)
            assert np.array_equal(ob, eb) and np.array_equal(os_, es), (kind, w)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference qfilter build absent")
def test_bsr_against_reference():
    g = golden()
    for i, a, b in golden_pairs(g):
        ba, sa = oracle.bsr_from_sorted(a)
        bb, sb = oracle.bsr_from_sorted(b)
        for kind in BSR_KINDS:
            ob, os_ = oracle.bsr_intersect(kind, ba, sa, bb, sb)
            assert np.array_equal(ob, g[f"pair{i}_bsr_bases"]), (kind, i)
            assert np.array_equal(os_, g[f"pair{i}_bsr_states"]), (kind, i)


@pytest.mark.parametrize("algo", ["branchless_merge", "galloping", "shuffling_avx2", "qfilter"])
def test_svs_fold(algo):
    """svs.rs:50-73: the fold equals the k-way intersection and satisfies the invariant trio
    (property_tests.rs svs_correct with SetCollection, testlib/mod.rs:168-212)."""
    rng = np.random.Generator(np.random.PCG64(13))
    for _ in range(30):
        k = int(rng.integers(2, 7))
        common = sorted_set(rng, int(rng.integers(0, 50)), 5000)
        sets = [np.union1d(sorted_set(rng, int(rng.integers(0, 2000)), 5000), common).astype(np.uint32)
                for _ in range(k)]
        sets.sort(key=len)
        exp = sets[0]
        for s in sets[1:]:
            exp = np.intersect1d(exp, s)
        got = oracle.svs(algo, sets)
        assert np.array_equal(got, exp.astype(np.uint32))
        assert prop_intersection_correct(got, sets)


def test_small_adaptive():
    """adaptive.rs:107-147: equals the k-way intersection (numpy) in any set order, satisfies the
    invariant trio, handles empty and identical sets; brute force over a tiny universe."""
    rng = np.random.Generator(np.random.PCG64(17))
    for trial in range(60):
        k = int(rng.integers(2, 12))
        common = sorted_set(rng, int(rng.integers(0, 40)), 3000)
        sets = [np.union1d(sorted_set(rng, int(rng.integers(0, 1500)), 3000), common).astype(np.uint32)
                for _ in range(k)]
        if trial % 2:
            sets.sort(key=len)
        if trial % 7 == 0:
            sets[int(rng.integers(0, k))] = np.zeros(0, np.uint32)
        if trial % 5 == 0:
            sets[-1] = sets[0].copy()
        exp = sets[0]
        for s in sets[1:]:
            exp = np.intersect1d(exp, s)
        got = oracle.small_adaptive(sets)
        assert np.array_equal(got, exp.astype(np.uint32)), trial
        assert prop_intersection_correct(got, sets)
    # brute force: every 3-subset pattern of a 12-value universe
    for mask in range(1 << 12):
        a = np.array([v for v in range(12) if mask >> v & 1], np.uint32)
        b = np.array([v for v in range(12) if (mask * 7 + 3) >> v & 1], np.uint32)
        c = np.array([v for v in range(12) if v % 3 != mask % 3], np.uint32)
        want = np.array([v for v in range(12) if v in a and v in b and v in c], np.uint32)
        assert np.array_equal(oracle.small_adaptive([a, b, c]), want), mask


def test_batch_runner_matches_pairwise():
    rng = np.random.Generator(np.random.PCG64(17))
    P = 200
    la = rng.integers(0, 300, size=P)
    lb = rng.integers(0, 300, size=P)
    As = [sorted_set(rng, int(n), 1000) for n in la]
    Bs = [sorted_set(rng, int(n), 1000) for n in lb]
    a_off = np.concatenate([[0], np.cumsum([x.size for x in As])]).astype(np.uint64)
    b_off = np.concatenate([[0], np.cumsum([x.size for x in Bs])]).astype(np.uint64)
    A = np.concatenate(As).astype(np.uint32)
    B = np.concatenate(Bs).astype(np.uint32)
    counts, cap_off, strided = oracle.batch_intersect("branchless_merge", A, a_off, B, b_off, nthreads=4)
    off, data = oracle.compact(counts, cap_off, strided)
    for p in range(P):
        exp = np.intersect1d(As[p], Bs[p])
        assert counts[p] == exp.size
        assert np.array_equal(data[int(off[p]):int(off[p + 1])], exp)


# -- the croaring comparator (oracle/roaring.c) ----------------------------------------------------
def _roaring_cases(rng):
    """Sets whose roaring forms use array containers (<= 4096 values under a high key), bitmap
    containers (more), several high keys, keys present in only one set, and empty sets."""
    cases = []
    for _ in range(30):
        u = int(rng.choice([1 << 12, 1 << 16, 1 << 18, 1 << 22]))
        na, nb = int(rng.integers(0, 30000)), int(rng.integers(0, 30000))
        cases.append((sorted_set(rng, na, u), sorted_set(rng, nb, u)))
    dense = np.arange(0, 1 << 17, 3, dtype=np.uint32)  # two bitmap containers
    cases.append((dense, np.arange(0, 1 << 17, 2, dtype=np.uint32)))
    cases.append((dense, np.array([3, 4, 65538, 65539, 131070], np.uint32)))  # array & bitmap
    cases.append((np.arange(0, 4096, dtype=np.uint32) * 16, dense))  # 4096-value array container
    cases.append((np.zeros(0, np.uint32), dense))
    cases.append((np.array([0xFFFFFFFF, 5], np.uint32)[::-1].copy(), np.array([5, 0xFFFFFFFF], np.uint32)))
    return cases


def test_roaring_against_numpy():
    rng = np.random.Generator(np.random.PCG64(23))
    for a, b in _roaring_cases(rng):
        exp = np.intersect1d(a, b).astype(np.uint32)
        assert np.array_equal(oracle.roaring_intersect(a, b), exp)
        assert np.array_equal(oracle.roaring_intersect(b, a), exp)


def test_roaring_skewed_arrays_gallop():
    # array & array with one side at least 64x smaller takes the galloping path
    big = np.arange(0, 4000, dtype=np.uint32)
    small = np.array([0, 17, 63, 1999, 3999, 4000], np.uint32)
    assert oracle.roaring_intersect(small, big).tolist() == [0, 17, 63, 1999, 3999]


def test_roaring_brute_force_tiny():
    u = 6
    for ma in range(1 << u):
        a = np.array([x for x in range(u) if ma >> x & 1], np.uint32)
        for mb in range(0, 1 << u, 5):
            b = np.array([x for x in range(u) if mb >> x & 1], np.uint32)
            assert oracle.roaring_intersect(a, b).tolist() == [x for x in range(u) if (ma & mb) >> x & 1]


def test_roaring_svs_and_batch():
    rng = np.random.Generator(np.random.PCG64(29))
    sets = [sorted_set(rng, int(rng.integers(5000, 40000)), 1 << 17) for _ in range(4)]
    exp = sets[0]
    for s in sets[1:]:
        exp = np.intersect1d(exp, s)
    assert np.array_equal(oracle.roaring_svs(sets), exp.astype(np.uint32))
    # batch: counts equal the pairwise oracle's, count-only and materialised, 1 and 3 threads
    a_sets = [sorted_set(rng, int(rng.integers(0, 9000)), 1 << 16) for _ in range(150)]
    b_sets = [sorted_set(rng, int(rng.integers(0, 9000)), 1 << 16) for _ in range(150)]
    A = np.concatenate(a_sets)
    B = np.concatenate(b_sets)
    a_off = np.concatenate([[0], np.cumsum([x.size for x in a_sets])]).astype(np.uint64)
    b_off = np.concatenate([[0], np.cumsum([x.size for x in b_sets])]).astype(np.uint64)
    want = np.array([np.intersect1d(x, y).size for x, y in zip(a_sets, b_sets)], np.uint32)
    rb = oracle.RoaringBatch(A, a_off, B, b_off)
    for nthreads, mat in ((1, False), (3, True)):
        assert rb.run(nthreads, materialize=mat) >= 0
        assert np.array_equal(rb.counts[:150], want)
    rb.close()


# -- x86 SIMD restatements (simd_x86.c), the CPU baseline's block intersections ----------------------
SIMD_X86 = [("shuffling_avx512_simd", 2, "shuffling_avx512"), ("broadcast_avx512_simd", 2, "shuffling_avx512"),
            ("shuffling_avx2_simd", 1, "shuffling_avx2")]


@pytest.mark.parametrize("algo,level,scalar", SIMD_X86)
def test_x86_simd_restatements(algo, level, scalar):
    """shuffling_avx512 / broadcast_avx512 / shuffling_avx2 with intrinsics: the reference's known
    answers, the golden vectors, numpy on block-aligned and ragged sizes, brute force on a
    48-element universe (so whole blocks are compared, not only the merge tail), and the scalar
    restatement of the same block walk."""
    if oracle.simd_level() < level:
        pytest.skip("this host lacks the instructions")
    g = golden()
    for i in range(6):
        assert np.array_equal(oracle.intersect2(algo, g[f"kat{i}_a"], g[f"kat{i}_b"]), g[f"kat{i}_out"]), i
    for i, a, b in golden_pairs(g):
        assert np.array_equal(oracle.intersect2(algo, a, b), g[f"pair{i}_out"]), i
    rng = np.random.Generator(np.random.PCG64(97))
    for trial in range(300):
        if trial % 3 == 0:  # whole blocks only
            la, lb = 16 * int(rng.integers(0, 60)), 16 * int(rng.integers(0, 60))
            u = max(la, lb, 1) * int(rng.integers(1, 4))
            a = np.sort(rng.choice(u, size=min(la, u), replace=False)).astype(np.uint32)
            b = np.sort(rng.choice(u, size=min(lb, u), replace=False)).astype(np.uint32)
        else:
            la, lb = rng.integers(0, 1200, size=2)
            u = int(rng.integers(1, 5000))
            a, b = sorted_set(rng, la, u), sorted_set(rng, lb, u)
        exp = np.intersect1d(a, b).astype(np.uint32)
        assert np.array_equal(oracle.intersect2(algo, a, b), exp), trial
        assert oracle.count2(algo, a, b) == exp.size
        assert np.array_equal(oracle.intersect2(scalar, a, b), exp)
    for trial in range(2000):
        ma, mb = rng.random(48) < 0.5, rng.random(48) < rng.random()
        a = np.flatnonzero(ma).astype(np.uint32)
        b = np.flatnonzero(mb).astype(np.uint32)
        assert oracle.intersect2(algo, a, b).tolist() == np.flatnonzero(ma & mb).tolist(), trial


def test_stream_read_bandwidth_is_measured():
    gbs = oracle.stream_read_gbs(2, nbytes=64 << 20, reps=2)
    assert 0.1 < gbs < 10000
