"""GPU parity of batched k-set groups (setops_hip_batch_set_further / KsetBatch): about 10,000
queries of 2 to 12 sets over one pool answered in one batch, every group's result against the
oracle's small_adaptive (setops/src/intersect/adaptive.rs:107-147) and svs fold
(setops/src/intersect/svs.rs:50-73), for each search entry, with and without the B-pool index."""
import sys

import numpy as np
import pytest

import oracle
from simd_set_operations_amd import _native
from simd_set_operations_amd.batch import KsetBatch

pytestmark = pytest.mark.gpu


def announce(msg):
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    yield


def make_pool(rng, nsets=500, universe=1 << 22, signed=False):
    """Sets that share a core, so k-way intersections are often non-empty: each set keeps a random
    share of 3,000 core values and adds its own values (0 to 60,000 of them)."""
    core = np.unique(rng.integers(0, universe, size=3000))
    sets = []
    for i in range(nsets):
        keep = core[rng.random(core.size) < rng.uniform(0.3, 1.0)]
        own = rng.integers(0, universe, size=int(2 ** rng.uniform(0, 16)) if i % 7 else int(rng.integers(0, 5)))
        s = np.unique(np.concatenate([keep, own]))
        if i % 11 == 0:
            s = s[: int(rng.integers(0, 4))]  # empty and tiny sets
        sets.append(s)
    if signed:
        sets = [(s.astype(np.int64) - (universe // 2)).astype(np.int32) for s in sets]
    dtype = np.int32 if signed else np.uint32
    lens = np.array([s.size for s in sets], np.uint64)
    pad = (lens + 3) // 4 * 4  # every set starts 16-byte aligned
    beg = np.zeros(nsets, np.uint64)
    beg[1:] = np.cumsum(pad)[:-1]
    pool = np.zeros(int(pad.sum()) + 4, dtype)
    for s, b in zip(sets, beg):
        pool[int(b):int(b) + s.size] = s.astype(dtype)
    return sets, pool, beg, beg + lens


def make_groups(rng, sets, ngroups, kmax=6):
    ks = rng.integers(2, kmax + 1, size=ngroups)
    ks[:: 97] = 12  # a few long groups
    ids = []
    for k in ks:
        g = rng.choice(len(sets), size=int(k), replace=False)
        g = sorted(g.tolist(), key=lambda i: (sets[i].size, i))  # smallest first
        ids.append(g)
    off = np.zeros(ngroups + 1, np.int64)
    off[1:] = np.cumsum(ks)
    return ids, off


@pytest.mark.parametrize("indexed", [False, True])
@pytest.mark.parametrize("name", ["hip_small_adaptive", "hip_galloping", "hip_binary_search", "hip_simd_galloping"])
def test_kset_groups_in_one_batch(name, indexed):
    rng = np.random.Generator(np.random.PCG64(311))
    sets, pool, beg, end = make_pool(rng)
    ids, off = make_groups(rng, sets, 10_000)
    flat = np.concatenate([np.array(g, np.int64) for g in ids])
    announce(f"kset groups {name} indexed={indexed} G={len(ids)}")
    kb = KsetBatch(name, pool, beg[flat], end[flat], off, index=indexed)
    for rep in range(2):
        kb.run()
        g_off, g_cnt, g_data = kb.results()
        for q, g in enumerate(ids):
            qs = [sets[i] for i in g]
            exp = oracle.small_adaptive(qs)
            got = g_data[int(g_off[q]):int(g_off[q + 1])]
            assert g_cnt[q] == exp.size and np.array_equal(got, exp), (rep, q, len(g))
            if q % 50 == 0:
                assert np.array_equal(got, oracle.svs("branchless_merge", qs)), (rep, q)
    # count-only
    kb.run(count_only=True)
    assert kb.totals()[0] == int(g_cnt.sum())
    # detached: the same batch answers the pairs alone
    kb.set_further(None)
    kb.run()
    g_off, g_cnt, g_data = kb.results()
    for q in range(0, len(ids), 37):
        exp = oracle.branchless_merge(sets[ids[q][0]], sets[ids[q][1]])
        assert np.array_equal(g_data[int(g_off[q]):int(g_off[q + 1])], exp), q
    kb.close()


def test_kset_groups_int32():
    rng = np.random.Generator(np.random.PCG64(313))
    sets, pool, beg, end = make_pool(rng, nsets=200, signed=True)
    ids, off = make_groups(rng, sets, 2000)
    flat = np.concatenate([np.array(g, np.int64) for g in ids])
    announce("kset groups int32")
    kb = KsetBatch("hip_small_adaptive", pool, beg[flat], end[flat], off, index=True)
    assert kb.signed
    kb.run()
    g_off, g_cnt, g_data = kb.results()
    kb.close()
    for q, g in enumerate(ids):
        exp = sets[g[0]]
        for i in g[1:]:
            exp = np.intersect1d(exp, sets[i])
        assert np.array_equal(g_data[int(g_off[q]):int(g_off[q + 1])], exp.astype(np.int32)), q


def test_further_sets_refused_by_merge_entries():
    rng = np.random.Generator(np.random.PCG64(317))
    sets, pool, beg, end = make_pool(rng, nsets=20)
    with pytest.raises(_native.SetopsError) as e:
        KsetBatch("hip_branchless_merge", pool, beg[:3], end[:3], np.array([0, 3]))
    assert e.value.code == _native.SETOPS_ERR_UNSUPPORTED
