"""The no-allocation pair call (setops_hip_intersect2_dev_noalloc) against setops_hip_intersect2_dev:
the same results and totals, bit for bit, on two calls in a row (the second with a smaller pair,
so it runs over what the first left in the scratch), for merge-path, search and BSR entries, with
outputs and count-only; and pairs beyond the stated limit are refused before anything runs."""
import ctypes
import sys

import numpy as np
import pytest

from simd_set_operations_amd import _native, intersect

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


def announce(msg):
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


def sorted_set(rng, n, universe):
    return np.unique(rng.integers(0, universe, size=n, dtype=np.int64)).astype(np.uint32)


def _dev(torch, arr):
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    t = torch.zeros(max(a.size, 4), dtype=torch.int32, device="cuda")
    if a.size:
        t[:a.size].copy_(torch.from_numpy(a.view(np.int32)))
    return t


def _call(lib, fn, algo, a, b, a_s, b_s, count_only, torch):
    """One call of `fn`; returns (totals, keys, states) on the host."""
    cap = max(min(a.numel(), b.numel()), 4)
    out = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    out_s = torch.full((cap,), -1, dtype=torch.int32, device="cuda") if a_s is not None else None
    tot = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    rc = getattr(lib, fn)(
        algo, a.data_ptr(), a_s.data_ptr() if a_s is not None else None, a.numel(), b.data_ptr(),
        b_s.data_ptr() if b_s is not None else None, b.numel(), None if count_only else out.data_ptr(),
        None if (count_only or out_s is None) else out_s.data_ptr(), cap, tot.data_ptr(), s.cuda_stream)
    _native.check(rc)
    torch.cuda.synchronize()
    t = tot.cpu().numpy()
    n = int(t[0])
    keys = out[:n].cpu().numpy().view(np.uint32) if not count_only else None
    sts = out_s[:n].cpu().numpy().view(np.uint32) if (out_s is not None and not count_only) else None
    return (int(t[0]), int(t[1])), keys, sts


@pytest.fixture(scope="module")
def lib(hip):
    return hip


@pytest.mark.parametrize("name", ["hip_branchless_merge", "hip_galloping", "hip_shuffling", "hip_binary_search",
                                  "hip_branchless_merge_bsr", "hip_galloping_bsr"])
def test_noalloc_matches_dev_entry(lib, name):
    import torch

    from simd_set_operations_amd.bsr import BsrVec

    rng = np.random.Generator(np.random.PCG64(211))
    algo = intersect.algo_id(name)
    bsr = name.endswith("_bsr")
    # first a pair of many tiles (near the merge limit for merge-path entries), then a small one-tile pair
    sizes = [(1_500_000, 2_000_000), (3000, 2500)] if "merge" in name or "shuffling" in name else \
        [(600_000, 3_000_000), (700, 90_000)]
    for call, (na, nb) in enumerate(sizes):
        announce(f"noalloc {name} call {call} na={na} nb={nb}")
        u = 4 * max(na, nb)
        a, b = sorted_set(rng, na, u), sorted_set(rng, nb, u)
        if bsr:
            ba, bb = BsrVec.from_sorted(a), BsrVec.from_sorted(b)
            da, db = _dev(torch, ba.bases), _dev(torch, bb.bases)
            das, dbs = _dev(torch, ba.states), _dev(torch, bb.states)
            da, db = da[:ba.bases.size], db[:bb.bases.size]
            das, dbs = das[:ba.bases.size], dbs[:bb.bases.size]
        else:
            da, db = _dev(torch, a)[:a.size], _dev(torch, b)[:b.size]
            das = dbs = None
        for count_only in (False, True):
            ref = _call(lib, "setops_hip_intersect2_dev", algo, da, db, das, dbs, count_only, torch)
            got = _call(lib, "setops_hip_intersect2_dev_noalloc", algo, da, db, das, dbs, count_only, torch)
            assert got[0] == ref[0], (call, count_only)
            if not count_only:
                assert np.array_equal(got[1], ref[1]), call
                if bsr:
                    assert np.array_equal(got[2], ref[2]), call
                else:  # and numpy's closed form
                    assert np.array_equal(got[1], np.intersect1d(a, b)), call


def test_noalloc_int32_mode(lib):
    import torch

    rng = np.random.Generator(np.random.PCG64(223))
    a = np.unique(rng.integers(-(1 << 31), 1 << 31, size=200_000, dtype=np.int64)).astype(np.int32)
    b = np.unique(np.concatenate([rng.choice(a, 50_000, replace=False),
                                  rng.integers(-(1 << 31), 1 << 31, size=150_000, dtype=np.int64).astype(np.int32)]))
    algo = intersect.algo_id("hip_branchless_merge") | _native.SIGNED_I32
    announce("noalloc int32")
    da, db = _dev(torch, a.view(np.uint32))[:a.size], _dev(torch, b.view(np.uint32))[:b.size]
    ref = _call(lib, "setops_hip_intersect2_dev", algo, da, db, None, None, False, torch)
    got = _call(lib, "setops_hip_intersect2_dev_noalloc", algo, da, db, None, None, False, torch)
    assert got[0] == ref[0]
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[1].view(np.int32), np.intersect1d(a, b))


def test_noalloc_refuses_pairs_beyond_the_limit(lib):
    """Refused on the host, before any launch (the pointers are never read)."""
    import torch

    d = torch.zeros(16, dtype=torch.int32, device="cuda")
    tot = torch.zeros(2, dtype=torch.int64, device="cuda")
    merge = intersect.algo_id("hip_branchless_merge")
    gallop = intersect.algo_id("hip_galloping")
    p = d.data_ptr()
    rc = lib.setops_hip_intersect2_dev_noalloc(merge, p, None, (1 << 21) + 1, p, None, 1 << 21, p, None, 16,
                                               tot.data_ptr(), None)
    assert rc == _native.SETOPS_ERR_INVALID
    assert b"no-allocation limit" in lib.setops_hip_last_error()
    rc = lib.setops_hip_intersect2_dev_noalloc(gallop, p, None, (1 << 20) + 1, p, None, 1 << 24, p, None, 16,
                                               tot.data_ptr(), None)
    assert rc == _native.SETOPS_ERR_INVALID
