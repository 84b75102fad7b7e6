"""GPU parity of the int32 mode (SETOPS_SIGNED_I32): sorted int32 sets with negative values, as the
reference's harness passes them (benchmark/src/timer.rs:19) and its property tests draw them
(setops/tests/testlib/mod.rs:11-30).  Every plain entry, single pair, k sets, CSR and gather
batches (indexed or not), against numpy's int32 intersect1d (the closed form: the ascending
intersection in signed order)."""
import ctypes
import sys

import numpy as np
import pytest

from simd_set_operations_amd import _native, intersect
from simd_set_operations_amd.batch import DeviceBatch, GatherBatch

pytestmark = pytest.mark.gpu

PLAIN = ["hip_branchless_merge", "hip_galloping", "hip_binary_search", "hip_shuffling", "hip_qfilter",
         "hip_small_adaptive", "hip_simd_galloping", "hip_broadcast"]
I32_MIN, I32_MAX = -(1 << 31), (1 << 31) - 1


def announce(msg):
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    yield


def i32_set(rng, n, lo=I32_MIN, hi=I32_MAX):
    return np.unique(rng.integers(lo, hi, size=n, endpoint=True, dtype=np.int64)).astype(np.int32)


def expect(a, b):
    return np.intersect1d(a, b).astype(np.int32)


EDGE = {
    "empty": (np.zeros(0, np.int32), np.array([-5, 0, 5], np.int32)),
    "extremes": (np.array([I32_MIN, -1, 0, 1, I32_MAX], np.int32), np.array([I32_MIN, 0, I32_MAX], np.int32)),
    "sign_boundary": (np.arange(-40, 40, dtype=np.int32), np.arange(-3, 70, 3, dtype=np.int32)),
    "all_negative": (np.arange(-1000, -1, 2, dtype=np.int32), np.arange(-999, 0, 3, dtype=np.int32)),
}


@pytest.mark.parametrize("name", PLAIN)
def test_int32_single_pairs(name):
    announce(f"int32 single pairs {name}")
    rng = np.random.Generator(np.random.PCG64(41))
    for case, (a, b) in EDGE.items():
        got = intersect.run_2set(a, b, name)
        assert got.dtype == np.int32, case
        assert np.array_equal(got, expect(a, b)), case
        assert np.array_equal(intersect.run_2set_c(a, b, name), expect(a, b)), case
    for i in range(12):
        # around zero, so both signs interleave; sizes cross tile boundaries
        na, nb = int(rng.integers(0, 20000)), int(rng.integers(0, 20000))
        if i % 3 == 2:
            na = int(rng.integers(1, 40))  # skewed: the search entries probe the small side
        span = int(rng.integers(1000, 1 << 20))
        a, b = i32_set(rng, na, -span, span), i32_set(rng, nb, -span, span)
        got = intersect.run_2set(a, b, name)
        assert np.array_equal(got, expect(a, b)), (i, a.size, b.size)


def test_int32_typed_entry_point(hip):
    announce("int32 typed entry point")
    rng = np.random.Generator(np.random.PCG64(43))
    a, b = i32_set(rng, 30000, -50000, 50000), i32_set(rng, 30000, -50000, 50000)
    out = np.zeros(min(a.size, b.size), np.int32)
    n = ctypes.c_size_t()
    vp = ctypes.c_void_p
    rc = hip.setops_hip_intersect2_i32(intersect.algo_id("hip_branchless_merge"), vp(a.ctypes.data), a.size,
                                       vp(b.ctypes.data), b.size, vp(out.ctypes.data), out.size, ctypes.byref(n))
    assert rc == 0, hip.setops_hip_last_error()
    assert np.array_equal(out[:n.value], expect(a, b))
    # the same bits sorted as u32 sets give the same elements, in u32 order
    au, bu = np.sort(a.view(np.uint32)), np.sort(b.view(np.uint32))
    got_u = intersect.run_2set(au, bu, "hip_branchless_merge")
    assert np.array_equal(np.sort(got_u.view(np.int32)), expect(a, b))


@pytest.mark.parametrize("name", ["hip_branchless_merge", "hip_galloping", "hip_small_adaptive"])
def test_int32_kset(name):
    announce(f"int32 k sets {name}")
    rng = np.random.Generator(np.random.PCG64(47))
    for k in (2, 3, 5, 11):
        sets = [i32_set(rng, int(rng.integers(2000, 40000)), -30000, 30000) for _ in range(k)]
        sets.sort(key=len)
        exp = sets[0]
        for s in sets[1:]:
            exp = np.intersect1d(exp, s)
        got = intersect.run_kset(sets, name)
        assert got.dtype == np.int32
        assert np.array_equal(got, exp.astype(np.int32)), k
        assert np.array_equal(intersect.run_svs(sets, name), exp.astype(np.int32)), k


def _csr(sets):
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum([s.size for s in sets])
    pool = np.concatenate(sets) if sets else np.zeros(0, np.int32)
    return pool.astype(np.int32), off


@pytest.mark.parametrize("name", PLAIN)
def test_int32_csr_batch(name):
    announce(f"int32 CSR batch {name}")
    rng = np.random.Generator(np.random.PCG64(53))
    sa, sb = [], []
    for p in range(700):
        n_l = int(2 ** rng.uniform(0, 13))
        n_s = max(0, int(n_l / 2 ** rng.uniform(0, 8)))
        span = max(4 * n_l, 64)
        sa.append(i32_set(rng, n_s, -span, span))
        sb.append(i32_set(rng, n_l, -span, span))
    A, a_off = _csr(sa)
    B, b_off = _csr(sb)
    db = DeviceBatch(name, A, a_off, B, b_off)
    assert db.signed
    db.run()
    off, cnt, data = db.results()
    db.close()
    assert data.dtype == np.int32
    for p in range(len(sa)):
        e = expect(sa[p], sb[p])
        assert cnt[p] == e.size and np.array_equal(data[int(off[p]):int(off[p + 1])], e), p


@pytest.mark.parametrize("indexed", [False, True])
@pytest.mark.parametrize("name", ["hip_galloping", "hip_binary_search", "hip_simd_galloping", "hip_branchless_merge",
                                  "hip_qfilter"])
def test_int32_gather_resident_set(name, indexed):
    """A resident int32 set with both signs (2^20 elements) against 600 small sets, read in place."""
    announce(f"int32 gather {name} indexed={indexed}")
    rng = np.random.Generator(np.random.PCG64(59))
    L = i32_set(rng, 1 << 20)
    small = []
    for p in range(600):
        m = int(rng.integers(0, 300))
        hits = rng.choice(L, size=m // 4, replace=False)
        small.append(np.unique(np.concatenate([hits, i32_set(rng, m - m // 4)])).astype(np.int32))
    pool, off = _csr(small)
    P = len(small)
    gb = GatherBatch(name, pool, off[:-1], off[1:], L, np.zeros(P, np.uint64), np.full(P, L.size, np.uint64))
    if indexed:
        gb.index_b()
    gb.run()
    o, cnt, data = gb.results()
    gb.close()
    for p in range(P):
        e = expect(small[p], L)
        assert cnt[p] == e.size and np.array_equal(data[int(o[p]):int(o[p + 1])], e), p
