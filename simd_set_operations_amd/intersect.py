"""The reference's intersection API over the GPU entries (mirrors setops/src/intersect.rs).

2-set entries have the Intersect2 shape ``f(set_a, set_b, visitor)`` (intersect.rs:40): they
visit the ascending intersection.  BSR entries have the Intersect2Bsr shape
``f(bsr_a, bsr_b, visitor)`` (bsr.rs:16).  The runners ``run_2set``, ``run_2set_c``,
``run_kset``, ``run_2set_bsr``, ``svs_generic`` and ``run_svs`` follow intersect.rs:44-90 and
svs.rs:50-116.

Every GPU entry calls the C ABI synchronously (one launch sequence per call: the plain per-pair
drop-in).  Batched, device-resident work goes through :mod:`simd_set_operations_amd.batch`.
"""
import ctypes

import numpy as np

from . import _native
from .bsr import BsrVec
from .visitor import VecWriter, Counter, BsrWriter, Clearable

_c = ctypes


def _u32(a):
    arr = np.ascontiguousarray(a)
    if arr.dtype != np.uint32:
        if arr.size and (arr.min() < 0 or arr.max() > 0xFFFFFFFF):
            raise ValueError("set elements must fit in u32")
        arr = arr.astype(np.uint32)
    return arr


def _is_int32(*sets):
    return any(np.asarray(x).dtype == np.int32 for x in sets)


def _keys(*sets):
    """The sets as u32 bit patterns for the C ABI, and whether they are int32 sets.

    An int32 input selects the int32 mode (SETOPS_SIGNED_I32: signed order, the element type of
    the reference's harness, benchmark/src/timer.rs:19); every set must then fit int32.  Otherwise
    the sets must fit u32."""
    signed = _is_int32(*sets)
    out = []
    for x in sets:
        arr = np.ascontiguousarray(x)
        if signed:
            if arr.dtype != np.int32:
                if arr.size and (arr.min() < -(1 << 31) or arr.max() > (1 << 31) - 1):
                    raise ValueError("int32 sets: every element must fit in int32")
                arr = arr.astype(np.int32)
            out.append(arr.view(np.uint32))
        else:
            out.append(_u32(arr))
    return out, signed


def _result(arr, signed):
    return arr.view(np.int32) if signed else arr


def _ptr(arr):
    return arr.ctypes.data_as(_c.c_void_p) if arr.size else None


def init(device=0):
    _native.check(_native.hip().setops_hip_init(device))


def algo_id(name):
    """Id of a registered entry; ``"name@i"`` selects tile shape i (SETOPS_SHAPE)."""
    base, _, shape = name.partition("@")
    aid = _native.check(_native.hip().setops_hip_algo_lookup(base.encode()))
    return aid | (_native.shape_flag(int(shape)) if shape else 0)


def registered():
    """All registered GPU entry names with their SETOPS_ALGO_* flags."""
    lib = _native.hip()
    return {lib.setops_hip_algo_name(i).decode(): lib.setops_hip_algo_flags(i) for i in range(lib.setops_hip_algo_count())}


def intersect2_c(name, set_a, set_b, out=None):
    """Intersect2C (intersect.rs:41): writes into ``out`` (capacity >= result) and returns the
    count.  ``out=None`` is the count-only call.  int32 sets (and an int32 ``out``) use the int32
    mode."""
    lib = _native.hip()
    (a, b), signed = _keys(set_a, set_b)
    algo = algo_id(name) | (_native.SIGNED_I32 if signed else 0)
    n = _c.c_size_t(0)
    if out is None:
        rc = lib.setops_hip_intersect2(algo, _ptr(a), a.size, _ptr(b), b.size, None, 0, _c.byref(n))
    else:
        want = np.int32 if signed else np.uint32
        if not (isinstance(out, np.ndarray) and out.dtype == want and out.flags.c_contiguous):
            raise TypeError(f"out must be a contiguous {np.dtype(want).name} array")
        rc = lib.setops_hip_intersect2(
            algo, _ptr(a), a.size, _ptr(b), b.size, _ptr(out) if out.size else None, out.size, _c.byref(n)
        )
    _native.check(rc)
    return n.value


def intersect2_bsr(name, bsr_a, bsr_b, count_only=False):
    """BSR 2-set call; returns (BsrVec result or None, element count)."""
    lib = _native.hip()
    ba, sa = _u32(bsr_a.bases), _u32(bsr_a.states)
    bb, sb = _u32(bsr_b.bases), _u32(bsr_b.states)
    words = _c.c_size_t(0)
    elems = _c.c_uint64(0)
    cap = min(ba.size, bb.size)
    if count_only:
        rc = lib.setops_hip_intersect2_bsr(
            algo_id(name), _ptr(ba), _ptr(sa), ba.size, _ptr(bb), _ptr(sb), bb.size, None, None, 0,
            _c.byref(words), _c.byref(elems),
        )
        _native.check(rc)
        return None, elems.value
    ob = np.zeros(max(cap, 1), np.uint32)
    os_ = np.zeros(max(cap, 1), np.uint32)
    rc = lib.setops_hip_intersect2_bsr(
        algo_id(name), _ptr(ba), _ptr(sa), ba.size, _ptr(bb), _ptr(sb), bb.size, _ptr(ob), _ptr(os_), cap,
        _c.byref(words), _c.byref(elems),
    )
    _native.check(rc)
    n = words.value
    return BsrVec(ob[:n], os_[:n]), elems.value


def _make_twoset(name):
    def entry(set_a, set_b, visitor):
        if isinstance(visitor, Counter):
            visitor.add(intersect2_c(name, set_a, set_b, None))
            return
        cap = min(len(set_a), len(set_b))
        out = np.zeros(max(cap, 1), np.int32 if _is_int32(set_a, set_b) else np.uint32)
        n = intersect2_c(name, set_a, set_b, out)
        visitor.visit_many(out[:n])

    entry.__name__ = name
    entry.__doc__ = f"GPU 2-set entry '{name}' (Intersect2 shape: set_a, set_b, visitor)."
    entry.setops_name = name
    return entry


def _make_bsr(name):
    def entry(bsr_a, bsr_b, visitor):
        if isinstance(visitor, Counter):
            _, elems = intersect2_bsr(name, bsr_a, bsr_b, count_only=True)
            visitor.add(elems)
            return
        res, _ = intersect2_bsr(name, bsr_a, bsr_b)
        visitor.visit_bsr_many(res.bases, res.states)

    entry.__name__ = name
    entry.__doc__ = f"GPU BSR entry '{name}' (Intersect2Bsr shape: bsr_a, bsr_b, visitor)."
    entry.setops_name = name
    return entry


_ENTRIES = {}


def lookup(name):
    """Registry lookup by name (benchmark/src/timer.rs:75-276): returns the callable entry."""
    if name in _ENTRIES:
        return _ENTRIES[name]
    flags = _native.hip().setops_hip_algo_flags(algo_id(name))
    fn = _make_bsr(name) if flags & _native.ALGO_BSR else _make_twoset(name)
    fn.setops_flags = flags
    _ENTRIES[name] = fn
    return fn


def __getattr__(attr):
    # module-level access to registered entries: intersect.hip_branchless_merge(a, b, v)
    if attr.startswith("hip_") and "@" not in attr:
        try:
            return lookup(attr)
        except _native.SetopsError as e:
            raise AttributeError(attr) from e
    raise AttributeError(attr)


def _resolve(f):
    return lookup(f) if isinstance(f, str) else f


def run_2set(set_a, set_b, intersect):
    """intersect.rs:44-52: run an Intersect2 entry into a VecWriter and return the vector."""
    w = VecWriter()
    _resolve(intersect)(set_a, set_b, w)
    return w.as_array()


def run_2set_c(set_a, set_b, intersect):
    """intersect.rs:54-68: allocate min(|a|,|b|), call the Intersect2C entry, truncate."""
    name = intersect if isinstance(intersect, str) else intersect.setops_name
    out = np.zeros(max(min(len(set_a), len(set_b)), 1), np.int32 if _is_int32(set_a, set_b) else np.uint32)
    n = intersect2_c(name, set_a, set_b, out)
    return out[:n].copy()


def run_2set_bsr(bsr_a, bsr_b, intersect):
    """intersect.rs:82-90: run a BSR entry into a BsrVec."""
    w = BsrWriter()
    _resolve(intersect)(bsr_a, bsr_b, w)
    return w.into_bsr()


def svs_generic(sets, left, right, intersect):
    """svs.rs:50-73: sets[0] ∩ sets[1] into ``left``; for each further set swap the sinks, clear
    the new ``left`` and intersect the previous result with the set.  Returns the final sink."""
    if len(sets) < 2:
        raise ValueError("svs_generic needs at least 2 sets")
    f = _resolve(intersect)
    f(sets[0], sets[1], left)
    for s in sets[2:]:
        left, right = right, left
        left.clear()
        f(right.as_array(), s, left)
    return left


def run_svs(sets, intersect):
    """svs.rs:103-116: svs_generic with two VecWriters."""
    return svs_generic(sets, VecWriter(), VecWriter(), intersect).as_array()


def run_kset(sets, intersect):
    """intersect.rs:70-80 for the GPU: the k-set call of the C ABI (svs_generic on the device
    library) with a fresh output vector."""
    if len(sets) < 2:
        raise ValueError("k-set intersection needs at least 2 sets")
    name = intersect if isinstance(intersect, str) else intersect.setops_name
    lib = _native.hip()
    arrs, signed = _keys(*sets)
    k = len(arrs)
    ptrs = (_c.c_void_p * k)(*[a.ctypes.data if a.size else None for a in arrs])
    lens = (_c.c_size_t * k)(*[a.size for a in arrs])
    cap = min(a.size for a in arrs)
    out = np.zeros(max(cap, 1), np.uint32)
    n = _c.c_size_t(0)
    _native.check(
        lib.setops_hip_svs(algo_id(name) | (_native.SIGNED_I32 if signed else 0), _c.cast(ptrs, _c.c_void_p),
                           _c.cast(lens, _c.c_void_p), k, _ptr(out), cap, _c.byref(n))
    )
    return _result(out[: n.value].copy(), signed)
