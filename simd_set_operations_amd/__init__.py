"""MI355X-native (gfx950) sorted-u32 set intersection.

A from-scratch GPU implementation of the sorted-array intersection path of
UNSW-database/simd_set_operations: 2-set entries, their BSR variants and svs_generic for k
sets, as HIP kernels behind the C ABI in ``include/setops_hip.h``.

Modules
-------
intersect  the reference API (Intersect2 entries, run_2set, svs_generic, run_svs, registry)
visitor    output sinks (Counter, VecWriter, SliceWriter, BsrWriter)
bsr        BsrVec encode / decode
batch      device-resident batches (one persistent launch for thousands of pairs)
datagen    seeded synthetic workloads
"""
from ._native import SetopsError  # noqa: F401

__all__ = ["SetopsError", "intersect", "visitor", "bsr", "batch", "datagen"]
