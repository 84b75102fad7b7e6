"""GPU parity: every result of the HIP library, called through the C ABI, is compared bit for
bit with the oracle (and with the golden vectors written from the reference's own code)."""
import os
import sys

import numpy as np
import pytest

import oracle
from simd_set_operations_amd import _native, datagen, intersect
from simd_set_operations_amd.bsr import BsrVec
from simd_set_operations_amd.visitor import Counter, VecWriter

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pairs.npz")
PLAIN = ["hip_branchless_merge", "hip_galloping", "hip_binary_search", "hip_shuffling", "hip_qfilter",
         "hip_small_adaptive", "hip_simd_galloping", "hip_broadcast"]
BSR = ["hip_branchless_merge_bsr", "hip_galloping_bsr", "hip_shuffling_bsr", "hip_qfilter_bsr", "hip_simd_galloping_bsr",
       "hip_broadcast_bsr"]


def announce(msg):
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


def sorted_set(rng, n, universe):
    return np.unique(rng.integers(0, universe, size=n, dtype=np.int64)).astype(np.uint32)


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    yield


# ---- 2-set, host pointers -----------------------------------------------------------------------
@pytest.mark.parametrize("name", PLAIN)
def test_known_answers(name):
    g = np.load(GOLDEN)
    f = intersect.lookup(name)
    for i in range(6):
        announce(f"{name} kat{i}")
        got = intersect.run_2set(g[f"kat{i}_a"], g[f"kat{i}_b"], f)
        assert np.array_equal(got, g[f"kat{i}_out"]), i


@pytest.mark.parametrize("name", PLAIN)
def test_golden_vectors(name):
    """Golden outputs come from the reference's compiled intersect_qfilter_uint_b4_v2."""
    g = np.load(GOLDEN)
    i = 0
    while f"pair{i}_a" in g:
        a, b = g[f"pair{i}_a"], g[f"pair{i}_b"]
        announce(f"{name} golden pair{i} |A|={a.size} |B|={b.size}")
        assert np.array_equal(intersect.run_2set(a, b, name), g[f"pair{i}_out"]), i
        assert np.array_equal(intersect.run_2set(b, a, name), g[f"pair{i}_out"]), i
        assert np.array_equal(intersect.run_2set_c(a, b, name), g[f"pair{i}_out"]), i
        i += 1


EDGE = {
    "both_empty": ([], []),
    "a_empty": ([], [1, 2, 3]),
    "b_empty": ([5], []),
    "single_hit": ([7], [7]),
    "single_miss": ([7], [8]),
    "identical": (list(range(0, 10000, 3)), list(range(0, 10000, 3))),
    "disjoint_interleaved": (list(range(0, 20000, 2)), list(range(1, 20000, 2))),
    "disjoint_ranges": (list(range(0, 5000)), list(range(5000, 9000))),
    "u32_extremes": ([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF], [0, 0x80000000, 0xFFFFFFFF]),
    "all_of_a": (list(range(0, 40000, 2)), list(range(0, 40000))),
    "tile_edges": (list(range(4095)), list(range(4097))),
    "one_vs_many": ([123456], list(range(0, 1000000, 3))),
}


@pytest.mark.parametrize("name", PLAIN)
@pytest.mark.parametrize("case", sorted(EDGE))
def test_edge_cases(name, case):
    a, b = (np.array(x, np.uint32) for x in EDGE[case])
    announce(f"{name} edge {case}")
    exp = oracle.branchless_merge(a, b)
    assert np.array_equal(intersect.run_2set(a, b, name), exp)
    assert np.array_equal(intersect.run_2set(b, a, name), exp)
    c = Counter()
    intersect.lookup(name)(a, b, c)
    assert c.count() == exp.size


@pytest.mark.parametrize("name", PLAIN)
def test_random_pairs(name):
    rng = np.random.Generator(np.random.PCG64(2024))
    for trial in range(40):
        la, lb = (int(x) for x in rng.integers(0, 30000, size=2))
        u = int(rng.integers(1, 200000))
        a, b = sorted_set(rng, la, u), sorted_set(rng, lb, u)
        announce(f"{name} random trial={trial} |A|={a.size} |B|={b.size} universe={u}")
        assert np.array_equal(intersect.run_2set(a, b, name), oracle.branchless_merge(a, b)), trial


@pytest.mark.parametrize("name", PLAIN)
def test_large_pair_split_across_tiles(name):
    """One pair of 2^20 + 2^20 with dense matches: merge-path splits land between equal values."""
    rng = np.random.Generator(np.random.PCG64(31))
    a = sorted_set(rng, 1 << 20, 1 << 21)
    b = np.union1d(sorted_set(rng, 1 << 20, 1 << 21), a[::2]).astype(np.uint32)
    announce(f"{name} large pair |A|={a.size} |B|={b.size}")
    assert np.array_equal(intersect.run_2set(a, b, name), oracle.branchless_merge(a, b))


@pytest.mark.parametrize("name", PLAIN)
def test_plain_pair_workload(name):
    a, b = datagen.plain_pair()
    announce(f"{name} plain pair 2^16+2^16")
    got = intersect.run_2set(a, b, name)
    assert got.size == 19661
    assert np.array_equal(got, oracle.branchless_merge(a, b))


def test_capacity_error(hip):
    a = np.arange(100, dtype=np.uint32)
    out = np.zeros(10, np.uint32)
    announce("capacity error")
    with pytest.raises(_native.SetopsError) as e:
        intersect.intersect2_c("hip_branchless_merge", a, a, out)
    assert e.value.code == _native.SETOPS_ERR_CAPACITY


# ---- BSR ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", BSR)
def test_bsr_golden(name):
    g = np.load(GOLDEN)
    i = 0
    while f"pair{i}_a" in g:
        announce(f"{name} golden pair{i}")
        ba, bb = BsrVec.from_sorted(g[f"pair{i}_a"]), BsrVec.from_sorted(g[f"pair{i}_b"])
        r = intersect.run_2set_bsr(ba, bb, name)
        assert np.array_equal(r.bases, g[f"pair{i}_bsr_bases"]), i
        assert np.array_equal(r.states, g[f"pair{i}_bsr_states"]), i
        c = Counter()
        intersect.lookup(name)(ba, bb, c)
        assert c.count() == np.intersect1d(g[f"pair{i}_a"], g[f"pair{i}_b"]).size
        i += 1


@pytest.mark.parametrize("name", BSR)
def test_bsr_random_dense(name):
    rng = np.random.Generator(np.random.PCG64(41))
    for trial in range(20):
        u = int(rng.integers(1000, 400000))
        a = sorted_set(rng, int(rng.integers(0, u // 2)), u)
        b = sorted_set(rng, int(rng.integers(0, u // 2)), u)
        announce(f"{name} dense trial={trial} |A|={a.size} |B|={b.size}")
        ba, bb = BsrVec.from_sorted(a), BsrVec.from_sorted(b)
        eb, es = oracle.bsr_intersect("branchless_merge_bsr", ba.bases, ba.states, bb.bases, bb.states)
        r = intersect.run_2set_bsr(ba, bb, name)
        assert np.array_equal(r.bases, eb) and np.array_equal(r.states, es), trial
        assert np.array_equal(r.to_sorted_set(), np.intersect1d(a, b))


# ---- k-set --------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", PLAIN)
def test_svs(name):
    rng = np.random.Generator(np.random.PCG64(51))
    for trial in range(15):
        k = int(rng.integers(2, 7))
        common = sorted_set(rng, int(rng.integers(0, 300)), 100000)
        sets = [np.union1d(sorted_set(rng, int(rng.integers(0, 20000)), 100000), common).astype(np.uint32)
                for _ in range(k)]
        sets.sort(key=len)
        announce(f"{name} svs trial={trial} k={k}")
        exp = oracle.svs("branchless_merge", sets)
        assert np.array_equal(intersect.run_kset(sets, name), exp)
        assert np.array_equal(intersect.run_svs(sets, name), exp)


# ---- device batches -----------------------------------------------------------------------------
def _batch_expect(A, a_off, B, b_off):
    counts, cap_off, strided = oracle.batch_intersect("branchless_merge", A, a_off, B, b_off, nthreads=8)
    off, data = oracle.compact(counts, cap_off, strided)
    return counts, off, data


def _csr(sets):
    off = np.zeros(len(sets) + 1, np.uint64)
    np.cumsum([s.size for s in sets], out=off[1:])
    data = np.concatenate(sets).astype(np.uint32) if off[-1] else np.zeros(0, np.uint32)
    return data, off


@pytest.mark.parametrize("name", PLAIN)
def test_batch_mixed_sizes(name):
    from simd_set_operations_amd.batch import DeviceBatch

    rng = np.random.Generator(np.random.PCG64(61))
    P = 3000
    As, Bs = [], []
    for p in range(P):
        kind = p % 7
        if kind == 0:
            la, lb = 0, int(rng.integers(0, 50))  # empty A
        elif kind == 1:
            la, lb = int(rng.integers(0, 5)), int(rng.integers(0, 5))  # tiny
        elif kind == 2:
            la, lb = int(rng.integers(5000, 20000)), int(rng.integers(5000, 20000))  # spans tiles
        else:
            la, lb = int(rng.integers(1, 600)), int(rng.integers(1, 600))
        u = max(la + lb, 1) * int(rng.integers(1, 6))
        As.append(sorted_set(rng, la, u))
        Bs.append(sorted_set(rng, lb, u))
    A, a_off = _csr(As)
    B, b_off = _csr(Bs)
    announce(f"{name} batch mixed P={P} |A|={A.size} |B|={B.size}")
    counts, off, data = _batch_expect(A, a_off, B, b_off)
    db = DeviceBatch(name, A, a_off, B, b_off)
    for rep in range(3):  # repeated runs reuse the handle and its plan
        db.run()
        g_off, g_cnt, g_data = db.results()
        assert np.array_equal(g_cnt, counts), rep
        assert np.array_equal(g_off, off), rep
        assert np.array_equal(g_data, data), rep
    db.run(count_only=True)
    words, elems = db.totals()
    assert words == elems == int(counts.sum())
    g_off, g_cnt, _ = db.results()
    assert np.array_equal(g_cnt, counts)
    assert np.array_equal(g_off, off)  # count-only runs still write the per-pair offsets
    db.close()


@pytest.mark.parametrize("name", BSR)
def test_batch_bsr(name):
    from simd_set_operations_amd.batch import DeviceBatch

    rng = np.random.Generator(np.random.PCG64(71))
    P = 800
    As, Bs, Ass, Bss, exp_b, exp_s = [], [], [], [], [], []
    for p in range(P):
        u = int(rng.integers(64, 50000))
        a = sorted_set(rng, int(rng.integers(0, u // 3 + 1)), u)
        b = sorted_set(rng, int(rng.integers(0, u // 3 + 1)), u)
        ba, bb = BsrVec.from_sorted(a), BsrVec.from_sorted(b)
        As.append(ba.bases), Ass.append(ba.states), Bs.append(bb.bases), Bss.append(bb.states)
        eb, es = oracle.bsr_intersect("branchless_merge_bsr", ba.bases, ba.states, bb.bases, bb.states)
        exp_b.append(eb), exp_s.append(es)
    A, a_off = _csr(As)
    As_, _ = _csr(Ass)
    B, b_off = _csr(Bs)
    Bs_, _ = _csr(Bss)
    announce(f"{name} batch bsr P={P}")
    db = DeviceBatch(name, A, a_off, B, b_off, A_states=As_, B_states=Bs_)
    db.run()
    g_off, g_cnt, g_b, g_s = db.results()
    exp_cnt = np.array([x.size for x in exp_b], np.uint32)
    assert np.array_equal(g_cnt, exp_cnt)
    eb_all, _ = _csr(exp_b)
    es_all, _ = _csr(exp_s)
    assert np.array_equal(g_b, eb_all) and np.array_equal(g_s, es_all)
    words, elems = db.totals()
    assert words == int(exp_cnt.sum())
    assert elems == int(np.unpackbits(es_all.view(np.uint8)).sum())
    db.close()


@pytest.mark.parametrize("name", ["hip_branchless_merge", "hip_galloping", "hip_qfilter"])
def test_uneven_batch_full_size(name):
    """The 65,536-pair uneven workload: per-pair counts equal the planted shared counts (a
    size-independent property of the generator), every output run is strictly increasing, and
    sampled pairs match the oracle exactly."""
    from simd_set_operations_amd.batch import DeviceBatch

    A, a_off, B, b_off, shared = datagen.uneven_batch(with_shared=True)
    announce(f"{name} uneven batch P={a_off.size - 1} E={A.size + B.size}")
    db = DeviceBatch(name, A, a_off, B, b_off)
    db.run()
    g_off, g_cnt, g_data = db.results()
    assert np.array_equal(g_cnt, shared)
    assert np.array_equal(np.diff(g_off), shared.astype(np.uint64))
    d = np.diff(g_data.astype(np.int64))
    boundaries = g_off[1:-1].astype(np.int64) - 1
    ok = d > 0
    ok[boundaries[(boundaries >= 0) & (boundaries < d.size)]] = True
    assert ok.all()
    rng = np.random.Generator(np.random.PCG64(81))
    for p in rng.integers(0, a_off.size - 1, size=300):
        exp = oracle.branchless_merge(A[a_off[p]:a_off[p + 1]], B[b_off[p]:b_off[p + 1]])
        assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], exp), p
    db.close()


# ---- gather batches -----------------------------------------------------------------------------
def _gather_case(rng, nsets, P, max_len):
    """A pool of sorted sets and P pairs drawn from it (pairs share and repeat sets)."""
    sets = []
    for i in range(nsets):
        ln = int(rng.integers(0, max_len)) if i % 5 else int(rng.integers(0, 8))
        sets.append(sorted_set(rng, ln, max(4 * ln, 16)))
    pool, off = _csr(sets)
    ia = rng.integers(0, nsets, size=P)
    ib = rng.integers(0, nsets, size=P)
    return sets, pool, off, ia, ib


@pytest.mark.parametrize("name", PLAIN)
def test_gather_batch(name):
    from simd_set_operations_amd.batch import GatherBatch

    rng = np.random.Generator(np.random.PCG64(91))
    sets, pool, off, ia, ib = _gather_case(rng, 400, 5000, 3000)
    # a second pool so A and B come from different allocations; B sets are sub-ranges too
    announce(f"{name} gather P={ia.size} pool={pool.size}")
    a_beg, a_end = off[ia], off[ia + 1]
    b_beg = off[ib] + (np.diff(off)[ib] // 4)  # suffix of each set: still sorted
    b_end = off[ib + 1]
    gb = GatherBatch(name, pool, a_beg, a_end, pool, b_beg, b_end)
    for rep in range(2):
        gb.run()
        g_off, g_cnt, g_data = gb.results()
        for p in range(ia.size):
            exp = oracle.branchless_merge(pool[a_beg[p]:a_end[p]], pool[b_beg[p]:b_end[p]])
            assert g_cnt[p] == exp.size, (rep, p)
            assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], exp), (rep, p)
    gb.run(count_only=True)
    words, _ = gb.totals()
    assert words == int(g_cnt.sum())
    gb.close()


@pytest.mark.parametrize("name", ["hip_branchless_merge", "hip_galloping"])
def test_batch_reserve_keeps_results(name):
    """setops_hip_batch_reserve grows a handle's workspace and staging ahead (and shrinks
    nothing): runs after a large and after a small reservation give the oracle's results."""
    from simd_set_operations_amd import _native
    from simd_set_operations_amd.batch import GatherBatch

    rng = np.random.Generator(np.random.PCG64(93))
    sets, pool, off, ia, ib = _gather_case(rng, 200, 1500, 2000)
    announce(f"{name} reserve P={ia.size}")
    gb = GatherBatch(name, pool, off[ia], off[ia + 1], pool, off[ib], off[ib + 1])
    lib = _native.hip()
    for npairs, units in ((8 * ia.size, 1 << 24), (1, 1), (ia.size, 0)):
        _native.check(lib.setops_hip_batch_reserve(gb.handle, npairs, units))
        gb.run()
        g_off, g_cnt, g_data = gb.results()
        for p in range(0, ia.size, 7):
            exp = oracle.branchless_merge(sets[ia[p]], sets[ib[p]])
            assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], exp), (npairs, p)
    gb.close()


@pytest.mark.parametrize("P", [3, 2000])
def test_rebind_sizes_that_differ_from_the_device_are_reported(P):
    """Host sizes given to setops_hip_batch_rebind that make fewer units than the device offsets
    are reported by setops_hip_batch_check: in a batch of one search tile (whose direct form
    compares the units it scans with the plan's) and in a batch of many tiles (the plan's check);
    consistent sizes again give the oracle's results."""
    from simd_set_operations_amd import _native
    from simd_set_operations_amd.batch import GatherBatch

    rng = np.random.Generator(np.random.PCG64(95))
    sets = [sorted_set(rng, int(rng.integers(8, 60)), 4000) for _ in range(50)]
    pool, off = _csr(sets)
    ia, ib = rng.integers(0, 50, size=P), rng.integers(0, 50, size=P)
    announce(f"rebind size mismatch P={P}")
    gb = GatherBatch("hip_galloping", pool, off[ia], off[ia + 1], pool, off[ib], off[ib + 1])
    assert (gb.info()["tiles"] == 1) == (P == 3)
    lib = _native.hip()
    na = (off[ia + 1] - off[ia]).astype(np.uint64)
    nb = (off[ib + 1] - off[ib]).astype(np.uint64)
    na2, nb2 = na.copy(), nb.copy()
    na2[0] -= 1  # every set holds 8 or more: the pair's probes drop by one
    nb2[0] -= 1
    for hna, hnb, ok in ((na2, nb2, False), (na, nb, True)):
        _native.check(lib.setops_hip_batch_rebind(gb.handle, gb.dA_beg.data_ptr(), gb.dA_end.data_ptr(),
                                                  gb.dB_beg.data_ptr(), gb.dB_end.data_ptr(), P, hna.ctypes.data,
                                                  hnb.ctypes.data))
        gb.run()
        if ok:
            gb.check()
            g_off, g_cnt, g_data = gb.results()
            for p in range(P):
                exp = oracle.branchless_merge(sets[ia[p]], sets[ib[p]])
                assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], exp), p
        else:
            with pytest.raises(_native.SetopsError) as ei:
                gb.check()
            assert ei.value.code == _native.SETOPS_ERR_INVALID and "differ" in str(ei.value)
    gb.close()


@pytest.mark.parametrize("name", BSR)
def test_gather_batch_bsr(name):
    from simd_set_operations_amd.batch import GatherBatch

    rng = np.random.Generator(np.random.PCG64(93))
    sets = [sorted_set(rng, int(rng.integers(0, 5000)), int(rng.integers(64, 60000))) for _ in range(200)]
    bsrs = [BsrVec.from_sorted(x) for x in sets]
    bases, off = _csr([b.bases for b in bsrs])
    states, _ = _csr([b.states for b in bsrs])
    P = 3000
    ia, ib = rng.integers(0, 200, size=P), rng.integers(0, 200, size=P)
    announce(f"{name} gather bsr P={P}")
    gb = GatherBatch(name, bases, off[ia], off[ia + 1], bases, off[ib], off[ib + 1], A_states=states,
                     B_states=states)
    gb.run()
    g_off, g_cnt, g_b, g_s = gb.results()
    for p in range(P):
        eb, es = oracle.bsr_intersect("branchless_merge_bsr", bsrs[ia[p]].bases, bsrs[ia[p]].states,
                                      bsrs[ib[p]].bases, bsrs[ib[p]].states)
        assert np.array_equal(g_b[g_off[p]:g_off[p + 1]], eb), p
        assert np.array_equal(g_s[g_off[p]:g_off[p + 1]], es), p
    gb.close()


@pytest.mark.parametrize("indexed", [False, True])
@pytest.mark.parametrize("name", ["hip_galloping", "hip_binary_search", "hip_branchless_merge", "hip_simd_galloping"])
def test_skewed_shared_large_set(name, indexed):
    """The skewed workload's shape at reduced size: one large set shared by every pair (searched
    in global memory when its narrowed window does not fit LDS) against many small sets, each with
    a planted share drawn from the large set."""
    from simd_set_operations_amd.batch import GatherBatch

    rng = np.random.Generator(np.random.PCG64(95))
    large = sorted_set(rng, 1 << 20, 1 << 24)
    smalls, shared = [], []
    for i in range(512):
        m = int(rng.integers(0, 300))
        hit = rng.choice(large, size=m // 10, replace=False)
        miss = np.setdiff1d(sorted_set(rng, m, 1 << 24), large)
        s_ = np.union1d(hit, miss).astype(np.uint32)
        smalls.append(s_)
        shared.append(hit.size)
    pool, off = _csr(smalls)
    P = len(smalls)
    announce(f"{name} skewed P={P} |L|={large.size} indexed={indexed}")
    gb = GatherBatch(name, pool, off[:-1], off[1:], large, np.zeros(P, np.uint64),
                     np.full(P, large.size, np.uint64))
    if indexed:
        gb.index_b()
    gb.run()
    g_off, g_cnt, g_data = gb.results()
    assert np.array_equal(g_cnt, np.array(shared, np.uint32))
    for p in range(P):
        assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], np.intersect1d(smalls[p], large)), p
    gb.close()


@pytest.mark.parametrize("name", ["hip_galloping", "hip_binary_search", "hip_galloping_bsr", "hip_simd_galloping",
                                  "hip_simd_galloping_bsr"])
def test_indexed_b_pool_ragged_ranges(name):
    """B-pool index: B ranges that start and end anywhere inside long sets (so index lines straddle
    range ends and neighbouring sets), empty ranges, and a range that ends at the pool's last
    element; indexed results equal unindexed ones and the oracle's, on two runs."""
    from simd_set_operations_amd.batch import GatherBatch

    bsr = name.endswith("_bsr")
    rng = np.random.Generator(np.random.PCG64(97))
    sets = [sorted_set(rng, int(rng.integers(4000, 120000)), 1 << 26) for _ in range(24)]
    sets += [sorted_set(rng, int(rng.integers(0, 40)), 1 << 26) for _ in range(8)]
    if bsr:
        bs = [BsrVec.from_sorted(x) for x in sets]
        pool, off = _csr([b.bases for b in bs])
        spool, _ = _csr([b.states for b in bs])
    else:
        pool, off = _csr(sets)
    n = np.diff(off)
    P = 1500
    ib = rng.integers(0, len(sets), size=P)
    lo = (n[ib] * rng.random(P) * 0.3).astype(np.uint64)
    hi = n[ib] - (n[ib] * rng.random(P) * 0.3).astype(np.uint64)
    hi[::97] = lo[::97]  # empty ranges
    b_beg, b_end = off[ib] + lo, off[ib] + hi
    b_beg[5], b_end[5] = off[-2], off[-1]  # the pool's last set, whole
    # small A sets: half drawn from the B range (hits), half random
    a_sets = []
    for p in range(P):
        rngB = pool[b_beg[p]:b_end[p]]
        m = int(rng.integers(0, 200))
        hit = rng.choice(rngB, size=min(m // 2, rngB.size), replace=False) if rngB.size else np.zeros(0, np.uint32)
        a_sets.append(np.union1d(hit, rng.integers(0, 1 << (21 if bsr else 26), size=m - hit.size)).astype(np.uint32))
    if bsr:
        a_st = [rng.integers(1, 1 << 32, size=x.size, dtype=np.uint64).astype(np.uint32) for x in a_sets]
        apool, aoff = _csr(a_sets)
        aspool, _ = _csr(a_st)
    else:
        apool, aoff = _csr(a_sets)
    announce(f"{name} indexed B pool={pool.size} P={P}")
    res = []
    for indexed in (False, True):
        kw = {"A_states": aspool, "B_states": spool} if bsr else {}
        gb = GatherBatch(name, apool, aoff[:-1], aoff[1:], pool, b_beg, b_end, **kw)
        if indexed:
            gb.index_b()
        for rep in range(2):
            gb.run()
            res.append(gb.results())
        gb.close()
    for r in res[1:]:
        for x, y in zip(res[0], r):
            assert np.array_equal(x, y)
    g_off = res[0][0]
    for p in range(0, P, 7):
        if bsr:
            eb, es = oracle.bsr_intersect("branchless_merge_bsr", apool[aoff[p]:aoff[p + 1]],
                                          aspool[aoff[p]:aoff[p + 1]], pool[b_beg[p]:b_end[p]],
                                          spool[b_beg[p]:b_end[p]])
            assert np.array_equal(res[0][2][g_off[p]:g_off[p + 1]], eb), p
            assert np.array_equal(res[0][3][g_off[p]:g_off[p + 1]], es), p
        else:
            exp = oracle.branchless_merge(apool[aoff[p]:aoff[p + 1]], pool[b_beg[p]:b_end[p]])
            assert np.array_equal(res[0][2][g_off[p]:g_off[p + 1]], exp), p


@pytest.mark.parametrize("name", ["hip_galloping", "hip_binary_search@1", "hip_simd_galloping", "hip_galloping@2",
                                  "hip_galloping_bsr"])
def test_search_batches_without_padding(name):
    """Search-policy gather batches whose every pair has at least 2*KS probes (ceil(U/128)) carry
    no padding units: exactly ceil(sum of probes / U) tiles, up to 128 pairs in one tile (pairs of
    the minimum size), pair boundaries anywhere inside lanes.  One pair below the minimum brings
    the padding back.  Results equal the oracle's for every pair, on two runs."""
    from simd_set_operations_amd.batch import GatherBatch

    bsr = name.endswith("_bsr")
    ks = {"@1": 1, "@2": 4}.get(name[-2:], 2)
    U, pad = 256 * ks, 2 * ks
    rng = np.random.Generator(np.random.PCG64(151))
    large = sorted_set(rng, 1 << 18, 1 << 24)
    for small_pair in (False, True):
        # the smaller side of each pair: the minimum size for runs of pairs (128 per tile), then
        # random sizes; half of each small set drawn from the large set
        sizes = [pad] * 700 + [int(rng.integers(pad, 3 * U)) for _ in range(600)]
        if small_pair:
            sizes[333] = pad - 1
        smalls = []
        for m in sizes:
            hit = rng.choice(large, size=m // 2, replace=False)
            s_ = np.union1d(hit, rng.integers(0, 1 << 24, size=m - m // 2)).astype(np.uint32)
            while s_.size < m:  # collisions: top up to the exact size
                s_ = np.union1d(s_, rng.integers(0, 1 << 24, size=m - s_.size)).astype(np.uint32)
            smalls.append(s_)
        P = len(smalls)
        pool, off = _csr(smalls)
        probes = int(np.diff(off).sum())
        kw = {}
        if bsr:
            lb = BsrVec.from_sorted(large)
            sb = [BsrVec.from_sorted(x) for x in smalls]
            pool, off = _csr([b.bases for b in sb])
            spool, _ = _csr([b.states for b in sb])
            probes = int(np.minimum(np.diff(off), lb.bases.size).sum())
            kw = {"A_states": spool, "B_states": lb.states}
            L = lb.bases
        else:
            L = large
        announce(f"{name} unpadded={not small_pair} P={P} probes={probes}")
        gb = GatherBatch(name, pool, off[:-1], off[1:], L, np.zeros(P, np.uint64), np.full(P, L.size, np.uint64),
                         **kw)
        unpadded = not small_pair and int(np.diff(off).min()) >= pad
        expect_tiles = -(-(probes + (0 if unpadded else pad * P)) // U)
        assert gb.info()["tiles"] == expect_tiles
        if bsr:
            expect = [oracle.bsr_intersect("branchless_merge_bsr", pool[off[p]:off[p + 1]], spool[off[p]:off[p + 1]],
                                           lb.bases, lb.states) for p in range(P)]
        else:
            expect = [(oracle.branchless_merge(smalls[p], large),) for p in range(P)]
        for rep in range(2):
            gb.run()
            res = gb.results()
            g_off = res[0]
            for p in range(P):
                for j, e in enumerate(expect[p]):
                    assert np.array_equal(res[2 + j][g_off[p]:g_off[p + 1]], e), (rep, p, j)
        gb.close()


# ---- graph edges (count-only gather batches) ------------------------------------------------------
def test_graph_triangle_sums():
    """R-MAT edges: per-edge counts of every entry equal the oracle's; full-list sums are three times
    the oriented sum; the BSR form gives the same sum through popcounts; a skew split (search for
    hub-leaf edges, merge for the rest) adds up to the same total."""
    import torch

    from simd_set_operations_amd import datagen
    from simd_set_operations_amd.batch import GatherBatch
    from simd_set_operations_amd.graph import EdgeIntersections, csr_to_bsr, split_by_skew

    rp, col = datagen.rmat_graph(scale=13)
    u, v = datagen.graph_edges(rp, col)
    up = datagen.upper_starts(rp, col)
    brp, bases, states = csr_to_bsr(rp, col)
    exp, _, _ = oracle.batch_intersect_gather("branchless_merge", col, rp[u], rp[u + 1], col, rp[v], rp[v + 1],
                                              nthreads=8, with_output=False)
    s3 = int(exp.sum())
    announce(f"graph scale=13 edges={u.size} 3T={s3}")
    dcol = torch.from_numpy(col.view(np.int32)).cuda()
    dbases = torch.from_numpy(bases.view(np.int32)).cuda()
    dstates = torch.from_numpy(states.view(np.int32)).cuda()
    for name in PLAIN:
        gb = GatherBatch(name, dcol, rp[u], rp[u + 1], dcol, rp[v], rp[v + 1], outputs=False)
        gb.run(count_only=True)
        torch.cuda.synchronize()
        assert gb.totals()[0] == s3, name
        assert np.array_equal(gb.d_counts.cpu().numpy().view(np.uint32), exp), name
        gb.close()
    ori = EdgeIntersections("hip_branchless_merge", dcol, up, rp[1:], u, v)
    ori.run()
    assert 3 * ori.totals()[0] == s3
    ori.close()
    for name in BSR:
        e = EdgeIntersections(name, dbases, brp[:-1], brp[1:], u, v, states=dstates)
        e.run()
        assert e.totals()[1] == s3, name
        e.close()
    deg = np.diff(rp).astype(np.int64)
    sk = split_by_skew(deg[u], deg[v], 16.0)
    assert 0 < sk.sum() < sk.size
    a = EdgeIntersections("hip_galloping", dcol, rp[:-1], rp[1:], u, v, select=sk)
    b = EdgeIntersections("hip_branchless_merge", dcol, rp[:-1], rp[1:], u, v, select=~sk)
    a.run()
    b.run()
    assert a.totals()[0] + b.totals()[0] == s3
    a.close()
    b.close()


# ---- served k-set queries -------------------------------------------------------------------------
def test_served_queries_alone_and_in_the_mix():
    """Continuous batching: every query's result equals the oracle's SvS over its lists, the plain
    per-query path's result, its result when served alone, and the Python request loop's result
    (the native loop, setops_hip_serve, is the default)."""
    from simd_set_operations_amd import serving

    corpus = serving.Corpus(nlists=4096, lmin=16, lmax=1 << 12, universe=1 << 18, seed=77)
    for i in range(0, corpus.nlists, 511):
        lst = corpus.list(i)
        assert lst.size == corpus.lens[i]
        assert np.all(np.diff(lst.astype(np.int64)) > 0) and (lst.size == 0 or lst[-1] < corpus.universe)
    ids, ks = serving.make_queries(corpus, 1500, seed=78)
    arrivals = serving.poisson_arrivals(ids.shape[0], 2e5, seed=79)
    announce(f"serving nq={ids.shape[0]}")
    eng = serving.Engine(corpus, "hip_galloping", max_pairs=256)
    lat, res, stats = eng.serve(ids, ks, arrivals, keep_results=True)
    assert stats["served"] == ids.shape[0]
    assert stats["steps"] > 1 and stats["pairs"] == ids.shape[0]  # one step answers a query
    lists = {}

    def L(i):
        if i not in lists:
            lists[i] = corpus.list(i)
        return lists[i]

    for q in range(ids.shape[0]):
        sets = [L(int(i)) for i in ids[q, :ks[q]]]
        exp = oracle.svs("branchless_merge", sets)
        assert np.array_equal(res[q], exp), q
    plat, pres, _ = serving.plain_serve(corpus, ids[:300], ks[:300], arrivals[:300] * 0, keep_results=True)
    for q in range(300):
        assert np.array_equal(pres[q], res[q]), q
    for q in range(0, ids.shape[0], 100):  # alone: a stream of one query
        _, r1, _ = eng.serve(ids[q:q + 1], ks[q:q + 1], np.zeros(1), keep_results=True)
        assert np.array_equal(r1[0], res[q]), q
    announce("serving python loop")
    _, rp, sp = eng.serve(ids, ks, arrivals, keep_results=True, loop="python")
    assert sp["served"] == ids.shape[0]
    for q in range(ids.shape[0]):
        assert np.array_equal(rp[q], res[q]), q
    # all at once (every query pending from the start: steps of max_pairs queries), then a stream
    # stopped by its time limit before anything is served (a host-side condition)
    _, ra, sa = eng.serve(ids, ks, np.zeros(ids.shape[0]), keep_results=True)
    assert sa["served"] == ids.shape[0] and sa["steps"] >= ids.shape[0] // 256
    for q in range(ids.shape[0]):
        assert np.array_equal(ra[q], res[q]), q
    lat0, r0, s0 = eng.serve(ids, ks, arrivals + 10.0, keep_results=True, timeout_s=0.0)
    assert s0["served"] == 0 and s0["steps"] == 0 and np.all(np.isnan(lat0)) and all(r is None for r in r0)
    eng.close()
    announce("serving three slots, 64 pairs per step")
    eng3 = serving.Engine(corpus, "hip_galloping", max_pairs=64, nslots=3)
    _, r3, s3 = eng3.serve(ids, ks, arrivals, keep_results=True)
    assert s3["served"] == ids.shape[0] and s3["pairs"] == ids.shape[0]
    for q in range(ids.shape[0]):
        assert np.array_equal(r3[q], res[q]), q
    eng3.close()
    # an admission budget that holds a few queries at a time: every query still answered, in more
    # steps (the budget, not the pair slots, stops each admission)
    _, cost = serving.query_cost_words(corpus, ids)
    budget_words = 4 * int(cost.max())
    announce(f"serving budget {budget_words} words")
    engb = serving.Engine(corpus, "hip_galloping", max_pairs=256, budget_bytes=4 * budget_words)
    _, rb, sb = engb.serve(ids, ks, np.zeros(ids.shape[0]), keep_results=True)
    assert sb["served"] == ids.shape[0] and sb["steps"] > sa["steps"]
    for q in range(ids.shape[0]):
        assert np.array_equal(rb[q], res[q]), q
    engb.close()


def test_batch_graph_replays_match():
    """Repeated runs with the same arguments replay a captured graph: every replay, with other
    device work and synchronisations in between, gives the oracle's results (and the totals)."""
    import torch

    from simd_set_operations_amd.batch import DeviceBatch

    rng = np.random.Generator(np.random.PCG64(101))
    As = [sorted_set(rng, int(rng.integers(0, 3000)), 8000) for _ in range(700)]
    Bs = [sorted_set(rng, int(rng.integers(0, 3000)), 8000) for _ in range(700)]
    A, a_off = _csr(As)
    B, b_off = _csr(Bs)
    counts, off, data = _batch_expect(A, a_off, B, b_off)
    announce("graph replays P=700")
    db = DeviceBatch("hip_branchless_merge", A, a_off, B, b_off)
    scratch = torch.zeros(1 << 20, device="cuda")
    for rep in range(6):
        db.d_out.fill_(-1)
        db.d_counts.fill_(-1)
        db.run()
        scratch.add_(1.0)  # other work on the same stream between replays
        torch.cuda.synchronize()
        g_off, g_cnt, g_data = db.results()
        assert np.array_equal(g_cnt, counts), rep
        assert np.array_equal(g_off, off), rep
        assert np.array_equal(g_data, data), rep
        assert db.totals()[0] == int(counts.sum()), rep
    for rep in range(3):
        db.run(count_only=True)
        torch.cuda.synchronize()
        assert db.totals()[0] == int(counts.sum()), rep
    db.close()


def test_datafile_pairs_through_every_entry(tmp_path):
    """A datafile written in the reference's format is read back and intersected by every entry."""
    from simd_set_operations_amd import datafile

    rng = np.random.Generator(np.random.PCG64(103))
    a, b = sorted_set(rng, 5000, 40000), sorted_set(rng, 9000, 40000)
    p = tmp_path / "0"
    datafile.write(p, [a, b])
    sa, sb = datafile.read(p)
    announce("datafile pair")
    exp = oracle.branchless_merge(sa, sb)
    for name in PLAIN:
        assert np.array_equal(intersect.run_2set(sa, sb, name), exp), name


def test_auto_entry_resolves_through_the_tuning_table():
    """hip_auto batches take the entry the tuning table prices cheapest for their pair sizes;
    their results are exact whichever entry that is."""
    from simd_set_operations_amd import tuning
    from simd_set_operations_amd.batch import DeviceBatch, GatherBatch

    rng = np.random.Generator(np.random.PCG64(107))
    As, Bs = [], []
    for _ in range(600):
        n = int(rng.integers(50, 3000))
        As.append(sorted_set(rng, n, 8 * n))
        Bs.append(sorted_set(rng, int(rng.integers(50, 3000)), 8 * n))
    A, a_off = _csr(As)
    B, b_off = _csr(Bs)
    announce("hip_auto balanced CSR batch")
    db = DeviceBatch("hip_auto", A, a_off, B, b_off)
    assert db.algo == tuning.choose(np.diff(a_off), np.diff(b_off))
    counts, off, data = _batch_expect(A, a_off, B, b_off)
    db.run()
    g_off, g_cnt, g_data = db.results()
    assert np.array_equal(g_cnt, counts) and np.array_equal(g_off, off) and np.array_equal(g_data, data)
    db.close()
    # one resident large set against many small sets
    L = sorted_set(rng, 1 << 20, 1 << 26)
    smalls = [np.union1d(rng.choice(L, 4, replace=False), sorted_set(rng, 12, 1 << 26)).astype(np.uint32)
              for _ in range(256)]
    pool, s_off = _csr(smalls)
    a_beg, a_end = s_off[:-1], s_off[1:]
    b_beg = np.zeros(256, np.uint64)
    b_end = np.full(256, L.size, np.uint64)
    announce("hip_auto skewed gather batch")
    gb = GatherBatch("hip_auto", pool, a_beg, a_end, L, b_beg, b_end)
    assert gb.algo == tuning.choose(a_end - a_beg, b_end - b_beg)
    gb.run()
    _, cnt, _ = gb.results()
    exp = [oracle.branchless_merge(s, L).size for s in smalls]
    assert np.array_equal(cnt, np.array(exp, np.uint32))
    gb.close()


def test_experiment_harness_on_gpu_entries(tmp_path):
    """The reference's experiment flow (generate, then time every algorithm per x) over the GPU
    entries: reference names resolve to them, every datafile is timed, and each entry's result
    count on every datafile equals the oracle's."""
    from simd_set_operations_amd import datafile
    from simd_set_operations_amd import experiment as ex

    toml = tmp_path / "e.toml"
    toml.write_text("""
[algorithm_sets]
all = ["branchless_merge", "galloping", "binary_search", "shuffling_avx2", "qfilter", "branchless_merge_bsr",
       "croaring"]
[[experiment]]
name = "two"
title = "2-set skew"
dataset = "two"
algorithm_set = "all"
[[experiment]]
name = "four"
title = "4-set"
dataset = "four"
algorithms = ["branchless_merge", "galloping", "shuffling_sse"]
[[dataset]]
name = "two"
type = "synthetic"
set_count = 2
gen_count = 2
vary = "skew"
skewness_factor = 0
to = 4000
step = 2000
density = 10
selectivity = 300
max_len = 12
[[dataset]]
name = "four"
type = "synthetic"
set_count = 4
gen_count = 2
vary = "set_count"
set_count_unused = 0
skewness_factor = 1000
to = 5
step = 1
density = 50
selectivity = 500
max_len = 11
""".replace("set_count_unused = 0\n", ""))
    doc = ex.load(toml)
    d = str(tmp_path / "datasets")
    ex.generate(doc, d, log=lambda m: None)
    announce("experiment harness: run")
    res = ex.run(doc, d, resolver=ex.gpu_resolver(), warmup_s=0.05, log=lambda m: None)
    two = res["datasets"]["two"]["algos"]
    assert two["croaring"] == []  # no GPU entry: skipped, as the reference skips unknown names
    for name in ("branchless_merge", "galloping", "binary_search", "shuffling_avx2", "qfilter", "branchless_merge_bsr"):
        assert [r["x"] for r in two[name]] == [0, 2000, 4000]
        assert all(len(r["times"]) == 2 and min(r["times"]) > 0 for r in two[name])
    four = res["datasets"]["four"]["algos"]
    assert [r["x"] for r in four["galloping"]] == [4, 5]
    resolve = ex.gpu_resolver()
    resolve_co = ex.gpu_resolver(count_only=True)
    for ds in ("two", "four"):
        for x in ex.xvalues(res["datasets"][ds]["info"]):
            for i in range(2):
                sets = datafile.read(os.path.join(d, ds, str(x), str(i)))
                exp = oracle.svs("branchless_merge", sets).size
                for name in res["datasets"][ds]["algos"]:
                    t = resolve(name)
                    if t is None or (t.bsr and len(sets) > 2):
                        continue
                    assert t.run(t.prepare(sets)) == exp, (ds, x, i, name)
                    tc = resolve_co(name)
                    assert tc.run(tc.prepare(sets)) == exp, (ds, x, i, name, "count-only")
    # device-resident 2-set timing
    res = ex.run(doc, d, experiments=["two"], resolver=ex.gpu_resolver(resident=True), warmup_s=0.02,
                 log=lambda m: None)
    assert all(len(r["times"]) == 2 for r in res["datasets"]["two"]["algos"]["galloping"])


def test_small_adaptive_many_sets():
    """hip_small_adaptive on up to 21 sets (three launches of 1 + 8 further sets), in any order,
    with empty and repeated sets: equals the oracle's small_adaptive and the k-way intersection."""
    rng = np.random.Generator(np.random.PCG64(109))
    for trial in range(24):
        k = int(rng.integers(3, 22))
        common = sorted_set(rng, int(rng.integers(0, 400)), 1 << 22)
        sets = [np.union1d(sorted_set(rng, int(rng.integers(0, 60000)), 1 << 22), common).astype(np.uint32)
                for _ in range(k)]
        if trial % 2:
            sets.sort(key=len)
        if trial % 5 == 0:
            sets[int(rng.integers(0, k))] = np.zeros(0, np.uint32)
        if trial % 3 == 0:
            sets[-1] = sets[0].copy()
        announce(f"hip_small_adaptive k={k} trial={trial}")
        exp = oracle.small_adaptive(sets)
        assert np.array_equal(exp, oracle.svs("branchless_merge", sets))
        assert np.array_equal(intersect.run_kset(sets, "hip_small_adaptive"), exp), trial


# ---- ordering against the caller's streams -------------------------------------------------------------
@pytest.mark.parametrize("name", ["hip_galloping", "hip_branchless_merge"])
def test_gather_offsets_pending_on_the_callers_stream(name):
    """A gather batch created (explicit capacity, no host read of the sizes) from device offsets that
    work still pending on torch's stream writes: creation counts the units only after that work."""
    import torch

    from simd_set_operations_amd.batch import GatherBatch

    rng = np.random.Generator(np.random.PCG64(131))
    sets, pool, off, ia, ib = _gather_case(rng, 300, 4000, 3000)
    announce(f"{name} gather offsets pending P={ia.size}")
    dpool = torch.from_numpy(pool.view(np.int32)).cuda()
    d_off = torch.from_numpy(off.view(np.int64)).cuda()
    d_ia = torch.from_numpy(ia.astype(np.int64)).cuda()
    d_ib = torch.from_numpy(ib.astype(np.int64)).cuda()
    cap = int(np.minimum(np.diff(off)[ia], np.diff(off)[ib]).sum())
    for rep in range(3):
        torch.cuda.synchronize()
        x = torch.randn(4096, 4096, device="cuda")
        for _ in range(8):
            x = x @ x  # keeps the stream busy while the offsets below are still to be written
            x = x / x.abs().amax()
        a_beg = torch.empty_like(d_ia)
        a_end = torch.empty_like(d_ia)
        b_beg = torch.empty_like(d_ib)
        b_end = torch.empty_like(d_ib)
        a_beg.copy_(d_off[d_ia] + (x[0, 0] * 0).long())  # written after the products
        a_end.copy_(d_off[d_ia + 1])
        b_beg.copy_(d_off[d_ib])
        b_end.copy_(d_off[d_ib + 1])
        gb = GatherBatch(name, dpool, a_beg, a_end, dpool, b_beg, b_end, capacity=cap)
        gb.run()
        g_off, g_cnt, g_data = gb.results()
        gb.close()
        for p in range(0, ia.size, 7):
            exp = oracle.branchless_merge(sets[ia[p]], sets[ib[p]])
            assert g_cnt[p] == exp.size, (rep, p)
            assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], exp), (rep, p)


def test_batch_runs_alternating_streams():
    """Runs of one handle on two streams in turn share its workspace: each run waits for the last."""
    import torch

    from simd_set_operations_amd.batch import DeviceBatch

    rng = np.random.Generator(np.random.PCG64(137))
    As = [sorted_set(rng, int(rng.integers(0, 6000)), 20000) for _ in range(900)]
    Bs = [sorted_set(rng, int(rng.integers(0, 6000)), 20000) for _ in range(900)]
    A, a_off = _csr(As)
    B, b_off = _csr(Bs)
    counts, off, data = _batch_expect(A, a_off, B, b_off)
    announce("alternating streams P=900")
    db = DeviceBatch("hip_branchless_merge", A, a_off, B, b_off)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(6):
        db.run(s1 if rep % 2 == 0 else s2)
        db.run(s2 if rep % 2 == 0 else s1, count_only=True)
        db.run(s1 if rep % 2 == 0 else s2)
        torch.cuda.synchronize()
        g_off, g_cnt, g_data = db.results()
        assert np.array_equal(g_cnt, counts), rep
        assert np.array_equal(g_off, off), rep
        assert np.array_equal(g_data, data), rep
    db.close()


# ---- tile shapes (SETOPS_SHAPE): the autotuner's alternatives give the same results ------------------
@pytest.mark.parametrize("name", ["hip_branchless_merge@1", "hip_branchless_merge@2", "hip_branchless_merge@3",
                                  "hip_galloping@1", "hip_galloping@2", "hip_binary_search@1",
                                  "hip_binary_search@2", "hip_simd_galloping@1", "hip_small_adaptive@2"])
def test_tile_shapes_exact(name):
    from simd_set_operations_amd.batch import DeviceBatch, GatherBatch

    rng = np.random.Generator(np.random.PCG64(149))
    As, Bs = [], []
    for p in range(900):
        n_l = int(2 ** rng.uniform(0, 14))
        n_s = max(0, int(n_l / 2 ** rng.uniform(0, 10)))
        u = max(10 * n_l, 16)
        As.append(sorted_set(rng, n_s, u))
        Bs.append(sorted_set(rng, n_l, u))
    A, a_off = _csr(As)
    B, b_off = _csr(Bs)
    counts, off, data = _batch_expect(A, a_off, B, b_off)
    announce(f"shape {name} CSR P=900")
    db = DeviceBatch(name, A, a_off, B, b_off)
    for rep in range(2):
        db.run()
        g_off, g_cnt, g_data = db.results()
        assert np.array_equal(g_cnt, counts) and np.array_equal(g_off, off) and np.array_equal(g_data, data), rep
    db.close()
    sets, pool, soff, ia, ib = _gather_case(rng, 300, 3000, 4000)
    gb = GatherBatch(name, pool, soff[ia], soff[ia + 1], pool, soff[ib], soff[ib + 1])
    gb.run()
    g_off, g_cnt, g_data = gb.results()
    gb.close()
    for p in range(0, ia.size, 3):
        exp = oracle.branchless_merge(sets[ia[p]], sets[ib[p]])
        assert np.array_equal(g_data[g_off[p]:g_off[p + 1]], exp), p
    for case in range(6):  # single pairs through the host-pointer call, one to many tiles
        a = sorted_set(rng, int(rng.integers(0, 1 << (6 + 2 * case))), 1 << 20)
        b = sorted_set(rng, int(rng.integers(0, 1 << (6 + 2 * case))), 1 << 20)
        assert np.array_equal(intersect.run_2set(a, b, name), oracle.branchless_merge(a, b)), case
