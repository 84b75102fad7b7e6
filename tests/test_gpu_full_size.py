"""GPU parity at the full sizes of the resident-set workloads, against the oracle and exact
properties:

* skewed: one resident set of 2^24 over [0, 2^28) against 4,096 sets of 256 (26 planted matches
  each), every search entry with and without the B-pool index, every pair's output against the
  oracle's galloping (setops/src/intersect/galloping.rs:7-30);
* graph: R-MAT at 2^22 vertices, edge factor 16, symmetrized: sum over edges u < v of
  |N(u) & N(v)| = 3T with T from the oriented lists, the BSR popcount sum and the skew-split sum
  equal to it, and per-edge counts against the oracle's branchless_merge (merge.rs:39-60,
  merge.rs:62-87 for BSR) on a seeded 1% sample of edges;
* served: the 2^20-list corpus (i.i.d. bounded-Zipf(1.1) lengths in [2^4, 2^18], 18.6 G elements,
  a pool longer than 2^32 elements), 2,000 queries arriving at 800K QPS served with results kept,
  each against the oracle's svs fold over host copies of its lists (svs.rs:50-73), and a sample
  served alone against the same results.
"""
import sys

import numpy as np
import pytest

import oracle
from simd_set_operations_amd import datagen
from simd_set_operations_amd.batch import GatherBatch

pytestmark = pytest.mark.gpu


def announce(msg):
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    yield


# ---- skewed resident set -------------------------------------------------------------------------
@pytest.fixture(scope="module")
def skewed():
    L, S = datagen.skewed()
    P, m = S.shape
    pool = S.reshape(-1)
    a_beg = np.arange(P, dtype=np.uint64) * m
    a_end = a_beg + m
    b_beg = np.zeros(P, np.uint64)
    b_end = np.full(P, L.size, np.uint64)
    cnt, cap_off, strided = oracle.batch_intersect_gather("galloping", pool, a_beg, a_end, L, b_beg, b_end,
                                                          nthreads=8, with_output=True)
    off, data = oracle.compact(cnt, cap_off, strided)
    return dict(L=L, pool=pool, a_beg=a_beg, a_end=a_end, b_beg=b_beg, b_end=b_end, cnt=cnt, off=off, data=data)


@pytest.mark.parametrize("indexed", [False, True])
@pytest.mark.parametrize("name", ["hip_galloping", "hip_binary_search", "hip_simd_galloping"])
def test_skewed_full_size(skewed, name, indexed):
    announce(f"skewed 2^24 x 4096x256 {name} indexed={indexed}")
    import torch

    k = skewed
    assert k["L"].size == 1 << 24 and k["pool"].size == 4096 * 256
    assert np.all(k["cnt"] == datagen.SKEWED["shared"])  # the planted matches
    dL = torch.from_numpy(k["L"].view(np.int32)).cuda()
    gb = GatherBatch(name, k["pool"], k["a_beg"], k["a_end"], dL, k["b_beg"], k["b_end"])
    if indexed:
        gb.index_b()
    for _ in range(2):  # the second run replays the captured launch sequence
        gb.run()
        off, cnt, data = gb.results()
        assert np.array_equal(cnt, k["cnt"])
        assert np.array_equal(off, k["off"])
        assert np.array_equal(data, k["data"])
    gb.close()


# ---- R-MAT graph at scale 22 --------------------------------------------------------------------
def test_graph_full_size():
    announce("graph R-MAT scale 22")
    import torch

    from simd_set_operations_amd.graph import EdgeIntersections, csr_to_bsr, split_by_skew

    rp, col = datagen.rmat_graph(scale=22)
    u, v = datagen.graph_edges(rp, col)
    up = datagen.upper_starts(rp, col)
    brp, bases, states = csr_to_bsr(rp, col)
    deg = np.diff(rp).astype(np.int64)
    assert rp.size - 1 == 1 << 22 and u.size > 60_000_000
    dcol = torch.from_numpy(col.view(np.int32)).cuda()
    dbases = torch.from_numpy(bases.view(np.int32)).cuda()
    dstates = torch.from_numpy(states.view(np.int32)).cuda()

    def run(e):
        e.run()
        torch.cuda.synchronize()
        e.batch.check()
        return e.totals()

    merge = EdgeIntersections("hip_branchless_merge", dcol, rp[:-1], rp[1:], u, v)
    s3, _ = run(merge)
    # per-edge counts on a seeded 1% sample against the oracle
    rng = np.random.Generator(np.random.PCG64(2024))
    sel = np.sort(rng.choice(u.size, size=u.size // 100, replace=False))
    got = merge.batch.d_counts[torch.from_numpy(sel).cuda()].cpu().numpy().view(np.uint32)
    su, sv = u[sel], v[sel]
    exp, _, _ = oracle.batch_intersect_gather("branchless_merge", col, rp[su], rp[su + 1], col, rp[sv], rp[sv + 1],
                                              nthreads=8, with_output=False)
    assert np.array_equal(got, exp)
    merge.close()
    oriented = EdgeIntersections("hip_branchless_merge", dcol, up, rp[1:], u, v)
    t, _ = run(oriented)
    oriented.close()
    assert s3 == 3 * t and t > 0
    bsr = EdgeIntersections("hip_branchless_merge_bsr", dbases, brp[:-1], brp[1:], u, v, states=dstates)
    _, s3b = run(bsr)
    bsr.close()
    assert s3b == s3
    sk = split_by_skew(deg[u], deg[v], 32.0)
    a = EdgeIntersections("hip_galloping", dcol, rp[:-1], rp[1:], u, v, select=sk, index=True)
    b = EdgeIntersections("hip_branchless_merge", dcol, rp[:-1], rp[1:], u, v, select=~sk)
    sa, _ = run(a)
    sb, _ = run(b)
    a.close()
    b.close()
    assert sa + sb == s3
    del dcol, dbases, dstates
    torch.cuda.empty_cache()


# ---- served k-set queries over the 2^20-list corpus -------------------------------------------------
def test_served_full_size_corpus():
    announce("served corpus 2^20 lists, 2000 queries at 800K QPS")
    import torch

    from simd_set_operations_amd import serving

    corpus = serving.Corpus()
    try:
        assert corpus.nlists == 1 << 20
        assert corpus.total > 1 << 32  # pool offsets beyond 32 bits are exercised
        eng = serving.Engine(corpus, "hip_galloping")
        nq = 2000
        ids, ks = serving.make_queries(corpus, nq, seed=77)
        assert np.any(corpus.beg[np.maximum(ids, 0)] > (1 << 32))
        arr = serving.poisson_arrivals(nq, 8e5, seed=78)
        lat, results, st = eng.serve(ids, ks, arr, keep_results=True)
        assert st["served"] == nq
        cache = {}

        def host_list(i):
            if i not in cache:
                cache[i] = corpus.list(i)
            return cache[i]

        for q in range(nq):
            lists = [host_list(int(i)) for i in ids[q, :ks[q]]]
            exp = oracle.svs("branchless_merge", lists)
            assert np.array_equal(results[q], exp), q
        # a sample served alone gives the same results
        for q in range(0, nq, 50):
            _, alone, _ = eng.serve(ids[q:q + 1], ks[q:q + 1], np.zeros(1), keep_results=True)
            assert np.array_equal(alone[0], results[q]), q
        eng.close()
    finally:
        del corpus
        torch.cuda.empty_cache()
