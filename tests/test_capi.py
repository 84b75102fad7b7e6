"""C ABI checks that need no GPU: the library loads, exports every symbol include/setops_hip.h
declares, the registry answers, and argument errors are reported (not undefined behaviour)."""
import ctypes
import os
import re

import numpy as np
import pytest

from simd_set_operations_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "setops_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(setops_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "setops_hip_intersect2" in names
    assert "setops_hip_svs" in names
    assert "setops_hip_batch_run" in names
    assert len(names) >= 17


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.HIP_LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) == set(_native.HIP_SIGNATURES)


def test_registry_without_device():
    lib = _native.hip()
    n = lib.setops_hip_algo_count()
    names = [lib.setops_hip_algo_name(i).decode() for i in range(n)]
    assert "hip_branchless_merge" in names
    assert "hip_branchless_merge_bsr" in names
    for i, name in enumerate(names):
        assert lib.setops_hip_algo_lookup(name.encode()) == i
        assert name.startswith("hip_")
    assert lib.setops_hip_algo_flags(lib.setops_hip_algo_lookup(b"hip_branchless_merge")) & _native.ALGO_TWOSET
    assert lib.setops_hip_algo_flags(lib.setops_hip_algo_lookup(b"hip_branchless_merge_bsr")) & _native.ALGO_BSR
    assert lib.setops_hip_algo_name(n) is None
    assert lib.setops_hip_algo_flags(-1) == 0


# Every 2-set entry of the reference's registry that this path covers has a GPU twin
# (benchmark/src/timer.rs:80-167 plain names, :212-254 BSR names).
REFERENCE_TWINS = {
    "branchless_merge": _native.ALGO_TWOSET | _native.ALGO_KSET,
    "galloping": _native.ALGO_TWOSET | _native.ALGO_KSET,
    "binary_search": _native.ALGO_TWOSET | _native.ALGO_KSET,
    "shuffling": _native.ALGO_TWOSET | _native.ALGO_KSET,
    "qfilter": _native.ALGO_TWOSET | _native.ALGO_KSET,
    "branchless_merge_bsr": _native.ALGO_BSR,
    "galloping_bsr": _native.ALGO_BSR,
    "shuffling_bsr": _native.ALGO_BSR,
    "qfilter_bsr": _native.ALGO_BSR,
}


def test_registry_covers_reference_entries():
    lib = _native.hip()
    for ref, flags in REFERENCE_TWINS.items():
        i = lib.setops_hip_algo_lookup(("hip_" + ref).encode())
        assert i >= 0, ref
        assert lib.setops_hip_algo_flags(i) == flags, ref


def test_unknown_algorithm_is_an_error():
    lib = _native.hip()
    assert lib.setops_hip_algo_lookup(b"branchless_merge_but_wrong") == _native.SETOPS_ERR_UNKNOWN_ALGO
    assert b"unknown algorithm" in lib.setops_hip_last_error()
    with pytest.raises(_native.SetopsError) as e:
        _native.check(lib.setops_hip_algo_lookup(b"nope"))
    assert e.value.code == _native.SETOPS_ERR_UNKNOWN_ALGO


def test_argument_errors_before_device_use():
    """Errors that the library reports without touching the GPU."""
    lib = _native.hip()
    n = ctypes.c_size_t(0)
    algo = lib.setops_hip_algo_lookup(b"hip_branchless_merge")
    bsr = lib.setops_hip_algo_lookup(b"hip_branchless_merge_bsr")
    # NULL input with non-zero length
    assert lib.setops_hip_intersect2(algo, None, 5, None, 0, None, 0, ctypes.byref(n)) == _native.SETOPS_ERR_INVALID
    # a BSR entry through the plain call
    a = np.arange(4, dtype=np.uint32)
    p = a.ctypes.data_as(ctypes.c_void_p)
    assert lib.setops_hip_intersect2(bsr, p, 4, p, 4, None, 0, ctypes.byref(n)) == _native.SETOPS_ERR_UNSUPPORTED
    # unknown id
    assert lib.setops_hip_intersect2(99, p, 4, p, 4, None, 0, ctypes.byref(n)) == _native.SETOPS_ERR_UNKNOWN_ALGO
    # svs with fewer than 2 sets (svs.rs:61 asserts)
    arr = (ctypes.c_void_p * 1)(a.ctypes.data)
    lens = (ctypes.c_size_t * 1)(4)
    rc = lib.setops_hip_svs(algo, ctypes.cast(arr, ctypes.c_void_p), ctypes.cast(lens, ctypes.c_void_p), 1, None, 0,
                            ctypes.byref(n))
    assert rc == _native.SETOPS_ERR_INVALID
    # batch handle NULL
    assert lib.setops_hip_batch_run(None, None, None, 0, None, None, None, None) == _native.SETOPS_ERR_INVALID
    assert lib.setops_hip_batch_check(None) == _native.SETOPS_ERR_INVALID
    assert lib.setops_hip_batch_reserve(None, 1, 1) == _native.SETOPS_ERR_INVALID
    assert lib.setops_hip_batch_step(None, None, None, 1, -1, None, None, None, 0, None, None, None, None, None, 0,
                                     None) == _native.SETOPS_ERR_INVALID
    # the int32 mode is for plain entries: the reference builds BSR from u32 values
    w, el = ctypes.c_size_t(), ctypes.c_uint64()
    rc = lib.setops_hip_intersect2_bsr(bsr | _native.SIGNED_I32, p, p, 4, p, p, 4, None, None, 0, ctypes.byref(w),
                                       ctypes.byref(el))
    assert rc == _native.SETOPS_ERR_UNSUPPORTED
    assert b"int32" in lib.setops_hip_last_error()
    # tile shapes: the plain merge entry has four, the block-compare entries one
    shuf = lib.setops_hip_algo_lookup(b"hip_shuffling")
    assert lib.setops_hip_intersect2(shuf | _native.shape_flag(1), p, 4, p, 4, None, 0,
                                     ctypes.byref(n)) == _native.SETOPS_ERR_UNSUPPORTED
    assert b"tile shape" in lib.setops_hip_last_error()
    assert lib.setops_hip_intersect2(algo | _native.shape_flag(4), p, 4, p, 4, None, 0,
                                     ctypes.byref(n)) == _native.SETOPS_ERR_UNSUPPORTED
    # an unknown id stays unknown with the int32 flag
    assert lib.setops_hip_intersect2(99 | _native.SIGNED_I32, p, 4, p, 4, None, 0,
                                     ctypes.byref(n)) == _native.SETOPS_ERR_UNKNOWN_ALGO


def test_serve_loop_argument_errors_before_device_use():
    """The native request loop checks its slots and its query stream before any device work."""
    lib = _native.hip()
    fake = 16  # never dereferenced: every case below fails validation first
    sl = _native.ServeSlot(fake, 64, 128, fake, 4, None)
    slots = (_native.ServeSlot * 2)(sl, sl)
    pairs = np.zeros(4, np.uint64)
    na = np.ones(1, np.uint64)
    xoff = np.zeros(2, np.int64)
    need = np.array([0, 8], np.int64)
    cost = np.array([40], np.int64)
    arr = np.zeros(1, np.float64)
    q = _native.ServeQueries(1, pairs.ctypes.data, na.ctypes.data, na.ctypes.data, xoff.ctypes.data, None, None,
                             need.ctypes.data, cost.ctypes.data, arr.ctypes.data)
    done = np.zeros(1)
    st = np.zeros(4, np.uint64)

    def serve(nslots, budget=1000, slots_=slots, q_=q):
        return lib.setops_hip_serve(ctypes.byref(slots_), nslots, ctypes.byref(q_), budget, 4, 1.0, done.ctypes.data,
                                    None, None, st.ctypes.data)

    assert serve(1) == _native.SETOPS_ERR_INVALID  # two slots at least
    assert b"two slots" in lib.setops_hip_last_error()
    assert lib.setops_hip_serve(None, 2, ctypes.byref(q), 1000, 4, 1.0, done.ctypes.data, None, None,
                                st.ctypes.data) == _native.SETOPS_ERR_INVALID
    assert serve(2, budget=39) == _native.SETOPS_ERR_CAPACITY  # the query's cost exceeds the budget
    need[1] = 1 << 20  # a result bound larger than the slots' result buffers
    assert serve(2) == _native.SETOPS_ERR_CAPACITY
    need[1] = 8
    bad = _native.ServeSlot(None, 64, 128, fake, 4, None)
    assert serve(2, slots_=(_native.ServeSlot * 2)(sl, bad)) == _native.SETOPS_ERR_INVALID
    arr2 = np.array([1.0, 0.5])  # arrivals must ascend
    pairs2, na2, xoff2 = np.zeros(8, np.uint64), np.ones(2, np.uint64), np.zeros(3, np.int64)
    need2, cost2 = np.array([0, 8, 16], np.int64), np.array([40, 40], np.int64)
    q2 = _native.ServeQueries(2, pairs2.ctypes.data, na2.ctypes.data, na2.ctypes.data, xoff2.ctypes.data, None, None,
                              need2.ctypes.data, cost2.ctypes.data, arr2.ctypes.data)
    done2 = np.zeros(2)
    assert lib.setops_hip_serve(ctypes.byref(slots), 2, ctypes.byref(q2), 1000, 4, 1.0, done2.ctypes.data, None, None,
                                st.ctypes.data) == _native.SETOPS_ERR_INVALID
    assert b"ascending" in lib.setops_hip_last_error()


def test_version_string():
    assert b"gfx950" in _native.hip().setops_hip_version()
