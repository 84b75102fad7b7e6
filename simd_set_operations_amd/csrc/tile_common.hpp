// tile_common.hpp -- wave/block primitives shared by the intersection kernels (gfx950, wave64).
//
// Nothing here is specific to one algorithm: group-cooperative partition search over global
// memory and block-wide exclusive scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace setops {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p)
{
    return __ballot(p);
}

// First index f in [lo, hi] such that pred(i) is false for every i in [f, hi), for a predicate
// that is monotone (true ... true, false ... false) on [lo, hi).  Executed cooperatively by a
// group of G consecutive lanes starting at lane `gbase`; `g` is the caller's index in the group.
// Each round issues one load per lane and shrinks the range by a factor G, so a 2^14-long range
// needs three dependent round trips with G = 32 instead of fourteen for a scalar binary search.
template <int G, class Pred>
__device__ __forceinline__ uint64_t group_partition_point(uint64_t lo, uint64_t hi, uint32_t g, uint32_t gbase,
                                                          Pred pred)
{
    const uint64_t gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    while (hi - lo > (uint64_t)G) {
        const uint64_t n = hi - lo;
        const uint64_t stride = (n + G - 1) / G;
        const uint64_t x = lo + (uint64_t)g * stride;
        const bool p = (g > 0) && (x < hi) && pred(x);
        const uint32_t c = (uint32_t)__popcll((ballot(p) >> gbase) & gmask);
        const uint64_t nlo = lo + (uint64_t)c * stride;
        uint64_t nhi = nlo + stride;
        if (nhi > hi) nhi = hi;
        lo = nlo;
        hi = nhi;
    }
    const bool p = ((uint64_t)g < hi - lo) && pred(lo + g);
    return lo + (uint64_t)__popcll((ballot(p) >> gbase) & gmask);
}

// Wave-wide inclusive scan of a u32 with DPP: shifts by 1, 2, 4, 8 inside each row of 16 lanes,
// then row 0's total into row 1 (and row 2's into row 3) by row_bcast:15, then the first half's
// total into rows 2 and 3 by row_bcast:31.  A lane with no source adds 0.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    constexpr int kRowShr = 0x110, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 1, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 2, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 4, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 8, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast15, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast31, 0xc, 0xf, false);
    return v;
}

// Wave-wide maximum of a u32 (the same DPP pattern as the scan; lane 63 ends with the maximum).
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
    constexpr int kRowShr = 0x110, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 1, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 2, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 4, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 8, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast15, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast31, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, kWave - 1);
}

// Block-wide exclusive scan; returns the exclusive prefix and writes the block total to *total.
// `scratch` needs NT/64 words of LDS.  Contains __syncthreads().
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *scratch, uint32_t *total)
{
    constexpr int NW = NT / kWave;
    const uint32_t w = threadIdx.x / kWave;
    const uint32_t incl = wave_inclusive_scan(v);
    if (lane_id() == kWave - 1) scratch[w] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        const uint32_t s = scratch[k];
        base += (k < (int)w) ? s : 0u;
        tot += s;
    }
    *total = tot;
    return base + incl - v;
}

} // namespace setops
