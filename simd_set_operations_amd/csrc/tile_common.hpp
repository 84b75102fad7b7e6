// tile_common.hpp -- wave/block primitives shared by the intersection kernels (gfx950, wave64).
//
// Nothing here is specific to one algorithm: group-cooperative partition search over global
// memory, block-wide exclusive scan, and the decoupled look-back that turns per-tile match
// counts into ordered output positions in a single pass.
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

// Wave-wide inclusive scan of a u32 (Hillis-Steele over DPP-capable shuffles).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    const uint32_t l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t t = __shfl_up(v, d, kWave);
        if (l >= (uint32_t)d) v += t;
    }
    return v;
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

// ---------------------------------------------------------------------------------------------
// Decoupled look-back over tile status words.  Each word packs a 2-bit flag and a 62-bit value,
// so a single 64-bit atomic publishes both; no fence is needed because nothing else is
// communicated.  Every tile publishes its aggregate right after its merge, before any waiting,
// so a waiter only ever waits for tiles that are unclaimed or still merging.
// ---------------------------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPfx = 2ull << 62;
constexpr uint64_t kFlagMask = 3ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1ull;
constexpr uint32_t kSpinForever = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t status_load(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------------------
// Tile claiming over Q counters.  Tile t belongs to queue t % Q and queue q hands out its tiles
// in order (tile = q + Q * counter).  One counter serialises device-scope atomics (~11 ns each
// on MI355X), so spreading claims over Q addresses multiplies the claim rate by Q.
//
// Rule: a workgroup claims from its home queue unless that queue's next tile is more than
// `slack` tiles beyond the globally smallest unclaimed tile; then it claims from the queue that
// holds that smallest tile.  So the lowest unclaimed tile is always claimed by whichever
// workgroup asks next once the other heads have moved `slack` past it: progress never depends
// on a particular workgroup being resident, which keeps the look-back waits finite.
// Executed by all lanes of one wave; returns the tile (>= ntiles when everything is claimed).
// ---------------------------------------------------------------------------------------------
// claim_heads issues the head reads (lanes < Q) so their latency can overlap other work;
// claim_finish decides and takes a tile.  Both are executed by all lanes of one wave.
template <int Q>
__device__ __forceinline__ uint32_t claim_heads(uint32_t *ctr, uint32_t ctr_stride)
{
    static_assert(Q >= 1 && Q <= kWave, "queues fit in a wave");
    const uint32_t l = lane_id();
    return (l < (uint32_t)Q) ? __hip_atomic_load(&ctr[l * ctr_stride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : 0u;
}

template <int Q>
__device__ __forceinline__ uint32_t claim_finish(uint32_t *ctr, uint32_t ctr_stride, uint32_t ntiles, uint32_t home,
                                                 uint32_t slack, uint32_t k_head)
{
    const uint32_t l = lane_id();
    for (int attempt = 0; attempt <= Q; attempt++) {
        if (attempt > 0) k_head = claim_heads<Q>(ctr, ctr_stride);
        uint32_t head = 0xFFFFFFFFu;
        if (l < (uint32_t)Q) {
            const uint64_t ht = (uint64_t)k_head * Q + l;
            head = ht < ntiles ? (uint32_t)ht : 0xFFFFFFFFu;
        }
        uint32_t mn = head;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, d, kWave));
        if (mn == 0xFFFFFFFFu) return ntiles;  // every queue exhausted
        const uint32_t home_head = (uint32_t)__shfl((int)head, (int)home, kWave);
        const uint32_t target = (home_head != 0xFFFFFFFFu && home_head - mn <= slack) ? home : (mn % Q);
        uint32_t t = 0;
        if (l == 0) {
            const uint32_t k = atomicAdd(&ctr[target * ctr_stride], 1u);
            const uint64_t tt = (uint64_t)k * Q + target;
            t = tt < ntiles ? (uint32_t)tt : ntiles;
        }
        t = (uint32_t)__shfl((int)t, 0, kWave);
        if (t < ntiles) return t;
        // the target queue ran out between the head read and the claim: it is now exhausted
    }
    return ntiles;
}

// Lane 0 publishes tile t's aggregate (tile 0 publishes its inclusive prefix directly).
__device__ __forceinline__ void lookback_publish(uint64_t *status, uint32_t t, uint64_t agg)
{
    if (lane_id() == 0) status_store(&status[t], (t == 0 ? kFlagPfx : kFlagAgg) | agg);
}

// Called by all 64 lanes of one wave after lookback_publish(t).  Sums predecessor words back to
// the nearest published prefix.  While a needed word is unpublished it sleeps and reloads, at
// most `spin_limit` times (kSpinForever: until published).  On success publishes tile t's
// inclusive prefix, stores the exclusive prefix in *excl and returns true; returns false when
// the spin limit ran out.
__device__ __forceinline__ bool lookback_resolve(uint64_t *status, uint32_t t, uint64_t agg, uint32_t spin_limit,
                                                 uint64_t *excl_out)
{
    const uint32_t l = lane_id();
    if (t == 0) {
        *excl_out = 0;
        return true;
    }
    uint64_t excl = 0;
    int64_t base = (int64_t)t - 1;
    uint32_t spins = 0;
    for (;;) {
        const int64_t idx = base - (int64_t)l;
        uint64_t v = (idx >= 0) ? status_load(&status[idx]) : kFlagPfx;
        uint64_t pfx;
        uint32_t stop;
        // Wait only for the words up to (and including) the nearest published prefix.
        for (;;) {
            pfx = ballot((v & kFlagMask) == kFlagPfx);
            stop = pfx ? (uint32_t)__builtin_ctzll(pfx) : (uint32_t)(kWave - 1);
            const uint64_t need = (stop >= kWave - 1) ? ~0ull : ((2ull << stop) - 1ull);
            if ((ballot((v & kFlagMask) == 0) & need) == 0) break;
            if (spin_limit != kSpinForever && spins >= spin_limit) return false;
            spins++;
            __builtin_amdgcn_s_sleep(2);
            if ((v & kFlagMask) == 0) v = status_load(&status[idx]);
        }
        if (!pfx) stop = kWave;
        // Lanes up to (and including) the nearest prefix word contribute.
        uint64_t contrib = (l <= stop) ? (v & kValMask) : 0ull;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) contrib += __shfl_xor(contrib, d, kWave);
        excl += contrib;
        if (pfx) break;
        base -= kWave;
    }
    if (l == 0) status_store(&status[t], kFlagPfx | (excl + agg));
    *excl_out = excl;
    return true;
}

// Claims the globally lowest unclaimed tile with a compare-and-swap on its queue head, so every
// tile below the returned one is already claimed (and will publish its aggregate without
// waiting): a blocking look-back on the returned tile always ends.  All lanes of one wave.
template <int Q>
__device__ __forceinline__ uint32_t claim_lowest(uint32_t *ctr, uint32_t ctr_stride, uint32_t ntiles)
{
    const uint32_t l = lane_id();
    for (;;) {
        const uint32_t k_head = claim_heads<Q>(ctr, ctr_stride);
        uint32_t head = 0xFFFFFFFFu;
        if (l < (uint32_t)Q) {
            const uint64_t ht = (uint64_t)k_head * Q + l;
            head = ht < ntiles ? (uint32_t)ht : 0xFFFFFFFFu;
        }
        uint32_t mn = head;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, d, kWave));
        if (mn == 0xFFFFFFFFu) return ntiles;
        uint32_t ok = 0;
        if (l == 0) {
            const uint32_t q = mn % Q, k = mn / Q;
            ok = (atomicCAS(&ctr[q * ctr_stride], k, k + 1u) == k) ? 1u : 0u;
        }
        if (__shfl((int)ok, 0, kWave)) return mn;
        // another workgroup took it: that is progress; read the heads again
    }
}

} // namespace setops
