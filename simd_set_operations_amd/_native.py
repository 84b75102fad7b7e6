"""Loader for the native libraries built in this package folder.

``libsetops_hip.so`` is the product (HIP kernels for gfx950 behind the C ABI of
``include/setops_hip.h``).  There is no fallback: if the library is missing every entry point
raises, so a GPU run can never silently take another path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libsetops_hip.so")
DATAGEN_LIB_PATH = os.path.join(_HERE, "libsetops_datagen.so")
DATAGEN_HIP_LIB_PATH = os.path.join(_HERE, "libsetops_datagen_hip.so")

SETOPS_OK = 0
SETOPS_ERR_INVALID = -1
SETOPS_ERR_CAPACITY = -2
SETOPS_ERR_UNKNOWN_ALGO = -3
SETOPS_ERR_HIP = -4
SETOPS_ERR_NOMEM = -5
SETOPS_ERR_UNSUPPORTED = -6
SETOPS_ERR_NO_DEVICE = -7

ALGO_TWOSET = 1
ALGO_BSR = 2
ALGO_KSET = 4
SIGNED_I32 = 0x10000  # OR'ed into an algorithm id: int32 sets, signed order
SHAPE_SHIFT = 20  # SETOPS_SHAPE(i) = i << SHAPE_SHIFT: the tile shape of a call


def shape_flag(i):
    return int(i) << SHAPE_SHIFT

_ERR_NAMES = {
    SETOPS_ERR_INVALID: "SETOPS_ERR_INVALID",
    SETOPS_ERR_CAPACITY: "SETOPS_ERR_CAPACITY",
    SETOPS_ERR_UNKNOWN_ALGO: "SETOPS_ERR_UNKNOWN_ALGO",
    SETOPS_ERR_HIP: "SETOPS_ERR_HIP",
    SETOPS_ERR_NOMEM: "SETOPS_ERR_NOMEM",
    SETOPS_ERR_UNSUPPORTED: "SETOPS_ERR_UNSUPPORTED",
    SETOPS_ERR_NO_DEVICE: "SETOPS_ERR_NO_DEVICE",
}


class SetopsError(RuntimeError):
    """A negative status from the C ABI, with the library's message."""

    def __init__(self, code, message):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {message}")
        self.code = code


class ServeSlot(ctypes.Structure):
    """setops_hip_serve_slot: a step slot of the native request loop."""

    _fields_ = [("handle", ctypes.c_void_p), ("idx_words", ctypes.c_uint64), ("out_words", ctypes.c_uint64),
                ("d_totals", ctypes.c_void_p), ("max_pairs", ctypes.c_uint64), ("stream", ctypes.c_void_p)]


class ServeQueries(ctypes.Structure):
    """setops_hip_serve_queries: a query stream in arrival order (host arrays)."""

    _fields_ = [("nq", ctypes.c_uint64), ("pairs", ctypes.c_void_p), ("na", ctypes.c_void_p),
                ("nb", ctypes.c_void_p), ("xoff", ctypes.c_void_p), ("xb", ctypes.c_void_p),
                ("xe", ctypes.c_void_p), ("need_pfx", ctypes.c_void_p), ("cost", ctypes.c_void_p),
                ("arrivals", ctypes.c_void_p)]


# Every exported symbol of include/setops_hip.h with its ctypes signature.
_c = ctypes
_u32p = _c.POINTER(_c.c_uint32)
_u64p = _c.POINTER(_c.c_uint64)
_szp = _c.POINTER(_c.c_size_t)
_vp = _c.c_void_p
HIP_SIGNATURES = {
    "setops_hip_init": (_c.c_int, [_c.c_int]),
    "setops_hip_last_error": (_c.c_char_p, []),
    "setops_hip_version": (_c.c_char_p, []),
    "setops_hip_algo_count": (_c.c_int, []),
    "setops_hip_algo_lookup": (_c.c_int, [_c.c_char_p]),
    "setops_hip_algo_name": (_c.c_char_p, [_c.c_int]),
    "setops_hip_algo_flags": (_c.c_uint, [_c.c_int]),
    "setops_hip_intersect2": (_c.c_int, [_c.c_int, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _c.c_size_t, _szp]),
    "setops_hip_intersect2_dev": (
        _c.c_int,
        [_c.c_int, _vp, _vp, _c.c_size_t, _vp, _vp, _c.c_size_t, _vp, _vp, _c.c_size_t, _vp, _vp],
    ),
    "setops_hip_intersect2_dev_noalloc": (
        _c.c_int,
        [_c.c_int, _vp, _vp, _c.c_size_t, _vp, _vp, _c.c_size_t, _vp, _vp, _c.c_size_t, _vp, _vp],
    ),
    "setops_hip_intersect2_bsr": (
        _c.c_int,
        [_c.c_int, _vp, _vp, _c.c_size_t, _vp, _vp, _c.c_size_t, _vp, _vp, _c.c_size_t, _szp, _u64p],
    ),
    "setops_hip_svs": (_c.c_int, [_c.c_int, _vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _szp]),
    "setops_hip_intersect2_i32": (_c.c_int, [_c.c_int, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _c.c_size_t, _szp]),
    "setops_hip_svs_i32": (_c.c_int, [_c.c_int, _vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _szp]),
    "setops_hip_batch_create": (
        _c.c_int,
        [_c.POINTER(_vp), _c.c_int, _vp, _vp, _c.c_uint64, _vp, _vp, _vp, _c.c_uint64, _vp, _c.c_uint64],
    ),
    "setops_hip_batch_create_gather": (
        _c.c_int,
        [_c.POINTER(_vp), _c.c_int, _vp, _vp, _c.c_uint64, _vp, _vp, _vp, _vp, _c.c_uint64, _vp, _vp, _c.c_uint64],
    ),
    "setops_hip_batch_index_b": (_c.c_int, [_vp]),
    "setops_hip_batch_set_further": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "setops_hip_batch_rebind": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _c.c_uint64, _vp, _vp]),
    "setops_hip_batch_reserve": (_c.c_int, [_vp, _c.c_uint64, _c.c_uint64]),
    "setops_hip_batch_step": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _c.c_int64, _vp, _vp, _vp, _c.c_uint64, _vp,
                                         _vp, _vp, _vp, _vp, _c.c_uint64, _vp]),
    "setops_hip_serve": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_int64, _c.c_uint64, _c.c_double, _vp, _vp, _vp, _vp]),
    "setops_hip_batch_run": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _vp, _vp, _vp, _vp]),
    "setops_hip_batch_check": (_c.c_int, [_vp]),
    "setops_hip_batch_set_timing": (_c.c_int, [_vp, _c.c_int]),
    "setops_hip_batch_kernel_time": (_c.c_int, [_vp, _c.POINTER(_c.c_double), _u64p]),
    "setops_hip_batch_info": (_c.c_int, [_vp, _u64p, _u64p, _u64p]),
    "setops_hip_batch_destroy": (_c.c_int, [_vp]),
}

DATAGEN_SIGNATURES = {
    "setops_gen_splitmix64": (_c.c_uint64, [_c.c_uint64, _vp, _c.c_uint64]),
    "setops_gen_planted_pair": (
        _c.c_int,
        [_c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _vp, _vp],
    ),
    "setops_gen_uneven_sizes": (
        _c.c_int,
        [_c.c_uint64, _c.c_uint64, _c.c_double, _c.c_double, _c.c_double, _c.c_double, _vp, _vp],
    ),
    "setops_gen_skewed": (
        _c.c_int,
        [_c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _vp, _vp],
    ),
    "setops_gen_rmat": (
        _c.c_int64,
        [_c.c_uint64, _c.c_uint32, _c.c_uint32, _c.c_double, _c.c_double, _c.c_double, _vp, _vp],
    ),
    "setops_gen_fixed_batch": (
        _c.c_int,
        [_c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _c.c_uint64, _vp, _vp],
    ),
    "setops_gen_uneven_fill": (
        _c.c_int,
        [_c.c_uint64, _c.c_uint64, _c.c_double, _c.c_double, _c.c_double, _c.c_double, _vp, _vp, _vp, _vp, _vp],
    ),
}

DATAGEN_HIP_SIGNATURES = {
    "setops_gen_sorted_lists_dev": (_c.c_int, [_c.c_uint64, _c.c_uint64, _vp, _vp, _c.c_uint64, _vp, _vp]),
}

_hip = None
_datagen = None
_datagen_hip = None


def _bind(path, sigs):
    if not os.path.exists(path):
        raise OSError(
            f"{os.path.basename(path)} is not built ({path}); run __graft_entry__.build() "
            "or `make -C simd_set_operations_amd/csrc`"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def hip():
    """The HIP library (loaded on first use; raises if it is not built).

    PyTorch ships its own libamdhip64 (same SONAME as the system one).  Importing torch first
    makes the dynamic linker bind this library to that already-loaded runtime, so the process
    holds one HIP runtime and device pointers / streams from torch are valid here."""
    global _hip
    if _hip is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _hip = _bind(HIP_LIB_PATH, HIP_SIGNATURES)
    return _hip


def datagen():
    global _datagen
    if _datagen is None:
        _datagen = _bind(DATAGEN_LIB_PATH, DATAGEN_SIGNATURES)
    return _datagen


def check(rc):
    """Raise SetopsError for a negative status code; return rc otherwise."""
    if rc < 0:
        msg = hip().setops_hip_last_error()
        raise SetopsError(rc, msg.decode() if msg else "")
    return rc


def datagen_hip():
    """Device-side workload generator (imports torch first, like hip())."""
    global _datagen_hip
    if _datagen_hip is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _datagen_hip = _bind(DATAGEN_HIP_LIB_PATH, DATAGEN_HIP_SIGNATURES)
    return _datagen_hip
