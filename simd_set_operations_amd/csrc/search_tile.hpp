// search_tile.hpp -- probe-per-element tile kernel for skewed pairs (gfx950).
//
// The search policy of the reference's galloping / binary_search_intersect
// (setops/src/intersect/galloping.rs:7-49): every element of the smaller set ("probe") is looked
// up in the larger set.  Units are probes (tile_plan.hpp, POLICY_SEARCH), so a tile carries
// U = NT*KS probes whatever the sizes of the larger sets.  Per tile:
//   1. segment table; stage the probe windows in LDS (the smaller side of each pair, which may
//      be A or B -- orientation does not change the result, A∩B ascending),
//   2. narrow each pair's larger set to [lo, hi): lower bound of the segment's first probe and
//      upper bound of its last probe, by a wave-cooperative 64-way partition search,
//   3. stage the narrowed larger windows in LDS while they fit (a prefix of the segments); the
//      rest are searched in global memory,
//   4. each lane looks up its KS probes, in order, from where the previous probe landed:
//      galloping (exponential then binary, galloping.rs:7-30), binary search over the rest of
//      the window (binary_search_intersect, galloping.rs:32-49) or block galloping with a full
//      compare of the last block (simd_galloping.rs:49-107),
//   5. block scan of hits, results (probe values; BSR: base and state AND) staged in order,
//      tile and per-pair counts -- the same outputs as the merge tile kernel.
#pragma once

#include "tile_plan.hpp"

namespace setops {

template <int NT, int KS, int MAXSEG, int CAPB, bool BSR>
struct SearchSmem {
    static constexpr int U = NT * KS;
    static constexpr int PKEYS = U + 8 * MAXSEG + 16;
    static constexpr int LKEYS = CAPB + 16;
    uint32_t pkey[PKEYS];               // probe windows
    uint32_t pst[BSR ? PKEYS : 4];
    uint32_t lkey[LKEYS];               // staged larger windows
    uint32_t lst[BSR ? LKEYS : 4];
    uint32_t seg_u[MAXSEG + 1];
    uint32_t seg_la[MAXSEG];            // LDS index of the segment's first probe
    uint32_t seg_wa[MAXSEG];            // probes in the window
    uint32_t seg_lo[MAXSEG];            // narrowed larger window [lo, hi), pair-relative
    uint32_t seg_hi[MAXSEG];
    uint32_t seg_lb[MAXSEG];            // LDS index of lkey[lo], or kNotStaged
    uint32_t seg_flags[MAXSEG];         // bit0: pair starts here, bit1: ends here, bit2: larger side is A
    uint64_t seg_lbase[MAXSEG];         // pool index of the larger set's first element
    uint32_t lane_mask[NT];
    uint32_t lane_pfx[NT];
    uint32_t scan_scratch[NT / kWave];
    uint32_t wave_el[NT / kWave];
    uint32_t win_c[MAXSEG + 1];
    uint64_t win_g[MAXSEG];
    uint32_t win_src[MAXSEG];
    uint32_t xlo[kMaxExtraSets];  // further sets (single pair): the tile's probe range in each
    uint32_t xhi[kMaxExtraSets];
    // top level of the B-pool index over the tile's larger B sets: top[j - top_j0] = B[kTopStride j]
    alignas(16) uint32_t top[kTopCap];
    unsigned long long red64[NT / kWave][2];
};

constexpr uint32_t kNotStaged = 0xFFFFFFFFu;
constexpr uint32_t kNoNarrow = 16 * 3072;  // larger sets searched through the B-pool index unnarrowed

// Lower bound of x in arr[lo, hi) starting from lo: galloping (probe lo+1, lo+3, lo+7, ... then
// binary search in the last gap) or plain binary search.  Works on LDS or global arrays.
template <bool GALLOP, class Arr>
__device__ __forceinline__ uint32_t lower_bound_from(const Arr &arr, uint32_t lo, uint32_t hi, uint32_t x)
{
    if constexpr (GALLOP) {
        uint32_t step = 1, prev = lo;
        uint32_t cur = lo;
        while (cur < hi && arr(cur) < x) {
            prev = cur + 1;
            cur = (step >= hi - lo) ? hi : lo + step;
            step <<= 1;
        }
        lo = prev;
        hi = cur;  // arr[prev-1] < x <= arr[cur] (or cur == hi)
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (arr(mid) < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Search modes of the probes: binary search, galloping, or galloping over blocks with a full
// compare of the last block (simd_galloping).
enum : int { SM_BINARY = 0, SM_GALLOP = 1, SM_BLOCK = 2 };
constexpr uint32_t kBlockLds = 8;    // block of a staged (LDS) window: 8 reads
constexpr uint32_t kBlockPool = 32;  // block of a pool window: one 128-byte span, nine 16-byte loads

// Blocks [lo + W b, lo + W b + W) of arr[lo, hi): the first block whose last element is >= x, found
// by galloping over the blocks' last elements (blocks 0, 1, 2, 4, 8, ...) and a binary search of the
// blocks the gallop leaves -- gallop_wide and binary_search_wide of simd_galloping.rs:196-243 with
// GPU-sized blocks.  Returns the block, or the block count when every element is < x.
template <uint32_t W, class Arr>
__device__ __forceinline__ uint32_t gallop_blocks(const Arr &arr, uint32_t lo, uint32_t hi, uint32_t x)
{
    const uint32_t nbk = (hi - lo + W - 1) / W;
    auto last = [&](uint32_t b) {
        const uint32_t e = lo + W * b + W;
        return arr((e < hi ? e : hi) - 1);
    };
    uint32_t bl = 0, bh = 0, step = 1;
    while (bh < nbk && last(bh) < x) {
        bl = bh + 1;
        bh = (step >= nbk) ? nbk : step;
        step <<= 1;
    }
    while (bl < bh) {  // blocks < bl end below x; block bh (if any) ends at >= x
        const uint32_t mid = (bl + bh) >> 1;
        if (last(mid) < x) bl = mid + 1; else bh = mid;
    }
    return bl;
}

// Lower bound of x in arr[lo, hi) by block galloping; the block that can hold x is counted in
// full, W independent reads in place of the last log2(W) dependent steps (block_compare,
// simd_galloping.rs:272-296).  For LDS arrays.
template <uint32_t W, class Arr>
__device__ __forceinline__ uint32_t lower_bound_block(const Arr &arr, uint32_t lo, uint32_t hi, uint32_t x)
{
    const uint32_t b = gallop_blocks<W>(arr, lo, hi, x);
    const uint32_t s0 = lo + W * b;
    if (s0 >= hi) return hi;
    const uint32_t n = (hi - s0 < W) ? hi - s0 : W;
    uint32_t lt = 0;
#pragma unroll
    for (uint32_t j = 0; j < W; j++) lt += (j < n && arr(s0 + j) < x) ? 1u : 0u;
    return s0 + lt;
}

// lower_bound_block over a window of a resident pool, L = P + g, P readable up to pool elements
// and 16-byte aligned: the last block is read as the aligned 16-byte chunks that cover it (one
// round trip).  Also reports whether the window holds x.  Pool values are compared XOR f (x is
// already in that space).
__device__ __forceinline__ uint32_t lower_bound_block_pool(const uint32_t *__restrict__ P, uint64_t pool, uint64_t g,
                                                           uint32_t lo, uint32_t hi, uint32_t x, uint32_t f, bool &found)
{
    constexpr uint32_t W = kBlockPool;
    const uint32_t *__restrict__ L = P + g;
    auto arr = [&](uint32_t i) { return L[i] ^ f; };
    const uint32_t b = gallop_blocks<W>(arr, lo, hi, x);
    const uint32_t s0 = lo + W * b;
    found = false;
    if (s0 >= hi) return hi;
    const uint64_t q0 = g + s0, q1 = g + ((hi - s0 < W) ? hi : s0 + W);
    const uint64_t c0 = q0 & ~3ull;
    uint32_t lt = 0;
    bool eq = false;
#pragma unroll
    for (int c = 0; c < (int)(W / 4) + 1; c++) {
        const uint64_t e = c0 + 4ull * (uint64_t)c;
        if (e < q1) {
            const uint4 v = load_chunk(P, pool, e);
            const uint32_t w[4] = {v.x ^ f, v.y ^ f, v.z ^ f, v.w ^ f};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint64_t q = e + (uint64_t)j;
                const bool in = q >= q0 && q < q1;
                lt += (in && w[j] < x) ? 1u : 0u;
                eq |= in && w[j] == x;
            }
        }
    }
    found = eq;
    return s0 + lt;
}

constexpr uint64_t kNoPos = ~0ull;

// Entries <= x among arr[lo, hi) of a sorted array of n entries, hi - lo <= 16: the aligned 16-byte
// chunks that cover the range load together (one round trip).
__device__ __forceinline__ uint32_t count_le16(const uint32_t *__restrict__ arr, uint64_t n, uint64_t lo, uint64_t hi,
                                               uint32_t x)
{
    uint32_t cnt = 0;
    const uint64_t c0 = lo & ~3ull;
#pragma unroll
    for (int c = 0; c < 5; c++) {
        const uint64_t e = c0 + 4ull * (uint64_t)c;
        if (e < hi) {
            const uint4 v = load_chunk(arr, n, e);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint64_t q = e + (uint64_t)j;
                cnt += (q >= lo && q < hi && w[j] <= x) ? 1u : 0u;
            }
        }
    }
    return cnt;
}

// Pool index of x in B[g0, g1), or kNoPos, using idx[k] = B[32k] ^ f (every 32nd pool element, in
// the compare space) and, when given, the middle level mid[m] = idx[16 m]: a binary search over
// the middle-level entries inside the range, then the 16 index entries under the one found (one
// round trip), or, without the middle level, a binary search over the index entries; then the
// only 32-element line that can hold x, compared in full (eight 16-byte loads).  The range is
// sorted (one set), so the entries inside it are too.
__device__ __forceinline__ uint64_t index_probe(const uint32_t *__restrict__ B, uint64_t pool,
                                                const uint32_t *__restrict__ idx, const uint32_t *__restrict__ mid,
                                                uint64_t g0, uint64_t g1, uint32_t x, uint32_t f)
{
    if (g0 >= g1) return kNoPos;
    const uint64_t kb = (g0 + 31) / 32, ke = (g1 + 31) / 32;  // entries with 32k in [g0, g1)
    uint64_t k;                                                // first entry > x
    if (mid) {
        const uint64_t mb = (kb + kMidIdx - 1) / kMidIdx, me = (ke + kMidIdx - 1) / kMidIdx;
        uint64_t lo = mb, hi = me;  // first middle entry > x
        while (lo < hi) {
            const uint64_t m = (lo + hi) >> 1;
            if (mid[m] <= x) lo = m + 1; else hi = m;
        }
        // the first index entry > x lies in [L, H], at most 16 entries
        const uint64_t L = (lo > mb) ? kMidIdx * (lo - 1) + 1 : kb;
        const uint64_t H = (lo < me) ? kMidIdx * lo : ke;
        k = L + count_le16(idx, (pool + 31) / 32, L, H, x);
    } else {
        uint64_t lo = kb, hi = ke;
        while (lo < hi) {
            const uint64_t m = (lo + hi) >> 1;
            if (idx[m] <= x) lo = m + 1; else hi = m;
        }
        k = lo;
    }
    if (k == 0) return kNoPos;  // kb == 0 and x < B[0] = B[g0]
    const uint64_t base = 32 * (k - 1);  // the line [base, base+32) holds every candidate
    uint64_t hit = kNoPos;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint64_t e = base + 4ull * c;
        const uint4 v = load_chunk(B, pool, e);
        const uint32_t w[4] = {v.x ^ f, v.y ^ f, v.z ^ f, v.w ^ f};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t q = e + (uint64_t)j;
            if (q >= g0 && q < g1 && q < pool && w[j] == x) hit = q;
        }
    }
    return hit;
}

// ---- wave-cooperative index lookups --------------------------------------------------------------
// A lookup through the B-pool index with its top level in LDS (every kTopStride-th element) searches
// the top level in LDS, then makes three dependent global reads: a 64-byte window of the middle
// level (every kMidStride-th element), a 64-byte window of the index (every 32nd element) and the
// 128-byte line of B that can hold x.  Read by the probe's own lane, each window is four or eight
// 16-byte loads to one line, and each load instruction of a wave touches 64 different lines: the
// vector memory pipeline looks every one of them up separately, and on resident pools that, not
// HBM, bounds the search tile.  Here a group of lanes reads one probe's window, one 16-byte chunk
// per lane, so an instruction reads 16 whole 64-byte windows (groups of 4) or 8 whole 128-byte lines
// (groups of 8); the probes' arguments reach the groups, and the counts come back, through
// cross-lane permutes.  Every lane of the wave must call these together.

__device__ __forceinline__ uint32_t lane_perm(uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}

// For the NP probes of every lane (probe k of lane p: window blk[k], lh[k], x[k]): entries <= x among
// positions [lo, hi) of the 16-entry block blk of arr (entries 16 blk .. 16 blk + 15, arr holding
// n entries); lh = lo | hi << 8, 0 for no window.  Round r serves probe k = r / 4 of lanes
// 16 (r % 4) .. 16 (r % 4) + 15, one per group of 4 lanes; all 4 NP loads are in flight together.
template <int NP>
__device__ __forceinline__ void coop_count16(const uint32_t *__restrict__ arr, uint64_t n, const uint32_t (&blk)[NP],
                                             const uint32_t (&lh)[NP], const uint32_t (&x)[NP], uint32_t (&res)[NP])
{
    constexpr int R = 4 * NP;
    const uint32_t lane = lane_id();
    const uint32_t c = lane & 3u;  // the chunk this lane reads of the block it serves
    uint4 v[R];
    uint32_t plh[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t p = 16u * (uint32_t)(r & 3) + (lane >> 2);  // the lane whose probe r / 4 is served
        const uint32_t pb = lane_perm(blk[r >> 2], p);
        plh[r] = lane_perm(lh[r >> 2], p);
        v[r] = plh[r] ? load_chunk(arr, n, 16ull * pb + 4u * c) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < NP; k++) res[k] = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t p = 16u * (uint32_t)(r & 3) + (lane >> 2);
        const uint32_t px = lane_perm(x[r >> 2], p);
        const uint32_t lo = plh[r] & 0xFFu, hi = plh[r] >> 8;
        const uint32_t w[4] = {v[r].x, v[r].y, v[r].z, v[r].w};
        uint32_t k = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t e = 4u * c + j;
            k += (e >= lo && e < hi && w[j] <= px) ? 1u : 0u;
        }
        k += __shfl_xor(k, 1, kWave);
        k += __shfl_xor(k, 2, kWave);
        const uint32_t t = lane_perm(k, 4u * (lane & 15u));  // lane 16 (r % 4) + i reads group i
        if ((lane >> 4) == (uint32_t)(r & 3)) res[r >> 2] = t;
    }
}

// For the NP probes of every lane: the position of x (compared XOR f) among positions [lo, hi) of
// the 32-element line `line` of B (elements 32 line .. 32 line + 31), or 32 when absent; lh as
// above.  Round r serves probe k = r / 8 of lanes 8 (r % 8) .. 8 (r % 8) + 7, one per group of 8.
template <int NP>
__device__ __forceinline__ void coop_find32(const uint32_t *__restrict__ B, uint64_t pool, const uint32_t (&line)[NP],
                                            const uint32_t (&lh)[NP], const uint32_t (&x)[NP], uint32_t f,
                                            uint32_t (&res)[NP])
{
    constexpr int R = 8 * NP;
    const uint32_t lane = lane_id();
    const uint32_t c = lane & 7u;
    uint4 v[R];
    uint32_t plh[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t p = 8u * (uint32_t)(r & 7) + (lane >> 3);
        const uint32_t pl = lane_perm(line[r >> 3], p);
        plh[r] = lane_perm(lh[r >> 3], p);
        v[r] = plh[r] ? load_chunk(B, pool, 32ull * pl + 4u * c) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < NP; k++) res[k] = 32;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t p = 8u * (uint32_t)(r & 7) + (lane >> 3);
        const uint32_t px = lane_perm(x[r >> 3], p);
        const uint32_t lo = plh[r] & 0xFFu, hi = plh[r] >> 8;
        const uint32_t w[4] = {v[r].x ^ f, v[r].y ^ f, v[r].z ^ f, v[r].w ^ f};
        uint32_t h = 0;  // position + 1 of the match (a set holds x at most once)
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t e = 4u * c + j;
            h = (e >= lo && e < hi && w[j] == px) ? e + 1u : h;
        }
        h |= __shfl_xor(h, 1, kWave);
        h |= __shfl_xor(h, 2, kWave);
        h |= __shfl_xor(h, 4, kWave);
        const uint32_t t = lane_perm(h, 8u * (lane & 7u));  // lane 8 (r % 8) + i reads group i
        if ((lane >> 3) == (uint32_t)(r & 7)) res[r >> 3] = t ? t - 1u : 32u;
    }
}

// Pool index of x in B[g0, g1), or kNoPos, for the NP probes of every lane of the wave (on = false:
// no probe), through the top level top[j - tj0] = B[kTopStride j] in LDS, the middle level
// mid[m] = idx[kMidIdx m] and the index idx[k] = B[32 k] (values in the compare space), with the
// global reads made by groups of lanes.  The windows it reads each lie in one aligned block:
// the middle-level entries over one top-level interval are those of one 16-entry block (top
// entries are every 16th middle entry), likewise the index entries under one middle-level
// interval, and the candidates of B one 32-element line.
template <int NP>
__device__ __forceinline__ void coop_index_probe_top(const bool (&on_)[NP], const uint32_t *__restrict__ B, uint64_t pool,
                                                     const uint32_t *__restrict__ idx, const uint32_t *__restrict__ mid,
                                                     const uint32_t *top, uint64_t tj0, const uint64_t (&g0)[NP],
                                                     const uint64_t (&g1)[NP], const uint32_t (&x)[NP], uint32_t f,
                                                     uint64_t (&q)[NP])
{
    bool on[NP];
    uint64_t L[NP], H[NP], mb[NP], me[NP];
    uint32_t mblk[NP], mlh[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        on[k] = on_[k] && g0[k] < g1[k];
        L[k] = H[k] = mb[k] = me[k] = 0;
        if (on[k]) {
            const uint64_t kb = (g0[k] + 31) / 32, ke = (g1[k] + 31) / 32;
            const uint64_t jb = (kb + kTopIdx - 1) / kTopIdx, je = (ke + kTopIdx - 1) / kTopIdx;
            uint64_t lo = jb, hi = je;  // first top entry > x
            while (lo < hi) {
                const uint64_t m = (lo + hi) >> 1;
                if (top[m - tj0] <= x[k]) lo = m + 1; else hi = m;
            }
            L[k] = (lo > jb) ? kTopIdx * (lo - 1) + 1 : kb;
            H[k] = (lo < je) ? kTopIdx * lo : ke;
            mb[k] = (L[k] + kMidIdx - 1) / kMidIdx;
            me[k] = (H[k] + kMidIdx - 1) / kMidIdx;
        }
        // middle level: [mb, me) inside block mb / 16
        mblk[k] = (uint32_t)(mb[k] >> 4);
        mlh[k] = (on[k] && me[k] > mb[k]) ? (uint32_t)(mb[k] & 15u) | ((uint32_t)(me[k] - 16ull * mblk[k]) << 8) : 0u;
    }
    uint32_t mc[NP];
    coop_count16<NP>(mid, (pool + kMidStride - 1) / kMidStride, mblk, mlh, x, mc);
    uint64_t L2[NP];
    uint32_t iblk[NP], ilh[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        uint64_t H2 = 0;
        L2[k] = 0;
        if (on[k]) {
            const uint64_t ms = mb[k] + mc[k];
            L2[k] = (ms > mb[k]) ? kMidIdx * (ms - 1) + 1 : L[k];
            H2 = (ms < me[k]) ? kMidIdx * ms : H[k];
        }
        // index: [L2, H2) inside block L2 / 16; the first entry > x is L2 + count
        iblk[k] = (uint32_t)(L2[k] >> 4);
        ilh[k] = (on[k] && H2 > L2[k]) ? (uint32_t)(L2[k] & 15u) | ((uint32_t)(H2 - 16ull * iblk[k]) << 8) : 0u;
    }
    uint32_t ic[NP];
    coop_count16<NP>(idx, (pool + 31) / 32, iblk, ilh, x, ic);
    uint32_t line[NP], blh[NP];
    bool bon[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        // B: the line [32(kk-1), 32 kk) holds every candidate; positions inside [g0, g1) and the pool
        const uint64_t kk = L2[k] + ic[k];
        bon[k] = on[k] && kk > 0;
        const uint64_t base = bon[k] ? 32 * (kk - 1) : 0;
        line[k] = (uint32_t)(base >> 5);
        blh[k] = 0;
        if (bon[k]) {
            const uint64_t e1 = g1[k] < pool ? g1[k] : pool;
            const uint32_t lo = g0[k] > base ? (uint32_t)min(g0[k] - base, (uint64_t)32) : 0u;
            const uint32_t hi = e1 > base ? (uint32_t)min(e1 - base, (uint64_t)32) : 0u;
            blh[k] = lo < hi ? lo | (hi << 8) : 0u;
        }
    }
    uint32_t pos[NP];
    coop_find32<NP>(B, pool, line, blh, x, f, pos);
#pragma unroll
    for (int k = 0; k < NP; k++) q[k] = (bon[k] && pos[k] < 32) ? 32ull * line[k] + pos[k] : kNoPos;
}

template <int NT, int KS, int MAXSEG, int CAPB, bool BSR, bool COUNT_ONLY, int SMODE, bool DIRECT = false>
__global__ __launch_bounds__(NT, 4) void search_tile_kernel(PairBatch in, TilePlan plan, TileOut out, uint32_t tile0)
{
    static_assert(KS >= 1 && KS <= 31, "lane masks hold KS bits");
    constexpr int U = NT * KS;
    constexpr int NW = NT / kWave;
    using Smem = SearchSmem<NT, KS, MAXSEG, CAPB, BSR>;
    __shared__ Smem sm;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / kWave;

    {
        // one tile per workgroup (a loop over tiles here costs ~20% on the merge kernel)
        const uint32_t t = tile0 + blockIdx.x;
        if (t >= plan.ntiles) return;
        // ---- 1. segment table, probe windows -------------------------------------------------------
        // DIRECT (the launch for plan.direct: one pair, or a batch of at most NT pairs): nothing
        // planned the tiles.  A batch's pair p owns the units (probes and padding) from the sum over
        // the pairs before it (a block scan, one pair per thread), and a tile's pairs are those its
        // unit range meets.  plan.single is uniform: every thread takes the branch or none.
        static_assert(!DIRECT || !BSR, "a direct tile holds plain pairs");
        constexpr bool direct = DIRECT;
        const bool dbatch = direct && !plan.single;
        if (direct && t == 0 && tid == 0 && out.error) *out.error = 0;  // before the flags below
        const uint64_t u0 = (uint64_t)t * U;
        uint32_t ulen = (uint32_t)((u0 + U < plan.utotal) ? U : plan.utotal - u0);
        uint32_t d_first = 0, d_last = 0;
        if constexpr (DIRECT) {
            if (!plan.single) {
                // pair tid's first unit and units go to LDS (the probe-key space, staged later)
                // for the segment table; the tile's first and last pairs after them
                if (tid == 0) {
                    sm.pkey[2 * NT] = ~0u;
                    sm.pkey[2 * NT + 1] = ~0u;
                }
                const bool own = tid < in.P;
                const uint32_t v = own ? (uint32_t)pair_units(in, plan.policy, plan.c0, tid) : 0u;
                uint32_t tu = 0;
                const uint32_t pu = block_exclusive_scan<NT>(v, sm.scan_scratch, &tu);
                // the device sizes must give the units the host planned for (as the plan checks)
                if (t == 0 && tid == 0 && out.error && (uint64_t)tu != plan.utotal) atomicOr(out.error, 4u);
                if (own) {
                    sm.pkey[2 * tid] = pu;
                    sm.pkey[2 * tid + 1] = v;
                    const uint64_t ue = u0 + ulen - 1;  // the tile's last unit
                    if (pu <= u0 && u0 < (uint64_t)pu + v) sm.pkey[2 * NT] = tid;
                    if (pu <= ue && ue < (uint64_t)pu + v) sm.pkey[2 * NT + 1] = tid;
                }
                __syncthreads();  // also: the scratch is scanned again below
                d_first = sm.pkey[2 * NT];
                d_last = sm.pkey[2 * NT + 1];
            }
        }
        const uint32_t p_first = direct ? d_first : plan.tile_first[t];
        const uint32_t p_last = direct ? d_last : plan.tile_last[t];
        uint32_t nseg = p_last - p_first + 1;
        bool bad = p_last < p_first || nseg > (uint32_t)MAXSEG || p_last >= in.P;
        if (bad) nseg = 0;
        if (t == 0 && tid == 0 && out.error && !direct && plan.uoff[in.P] != plan.utotal) atomicOr(out.error, 4u);
        if (dbatch && tid == 0 && !bad) {  // the tile maps the placement reads
            plan.tile_first[t] = p_first;
            plan.tile_last[t] = p_last;
        }
        uint32_t seg_pu = 0;  // direct batch: the segment's pair's first unit
        uint32_t cp = 0;
        uint64_t gp = 0;
        uint32_t swap = 0;
        // the span of the tile's larger B sets, for the top level of the B-pool index (1b)
        unsigned long long gmin = ~0ull, gmax = 0;
        if (tid < nseg) {
            SegWin w;
            if (dbatch) {
                const uint32_t p = p_first + tid;
                seg_pu = sm.pkey[2 * p];
                w = search_direct_window(in, p, seg_pu, sm.pkey[2 * p + 1], u0, ulen);
            } else {
                w = direct ? direct_window<true>(plan) : seg_window(in, plan, t, p_first, nseg, tid, ulen);
            }
            swap = (w.nb < w.na) ? 1u : 0u;  // probes come from the smaller side
            gp = (swap ? w.b : w.a) + w.i0;
            const uint32_t n = (uint32_t)(w.i1 - w.i0);
            cp = n ? ((uint32_t)(gp & 3) + n + 3) / 4 : 0u;
            sm.seg_u[tid] = w.u;
            sm.seg_wa[tid] = n;
            sm.seg_flags[tid] = w.flags | (swap << 2);
            sm.seg_lbase[tid] = swap ? w.a : w.b;
            sm.seg_hi[tid] = (uint32_t)(swap ? w.na : w.nb);  // larger size until narrowed
            if (!swap && w.nb > 0) {
                gmin = w.b;
                gmax = w.b + w.nb;
            }
        }
        if (in.b_top) {
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                gmin = min(gmin, (unsigned long long)__shfl_xor(gmin, d, kWave));
                gmax = max(gmax, (unsigned long long)__shfl_xor(gmax, d, kWave));
            }
            if (lane == 0) {
                sm.red64[wave][0] = gmin;
                sm.red64[wave][1] = gmax;
            }
        }
        uint32_t tot = 0;
        const uint32_t pfx = block_exclusive_scan<NT>(cp, sm.scan_scratch, &tot);
        if (4u * tot + 8u > (uint32_t)Smem::PKEYS) bad = true;  // not reachable by the unit bound
        if (tid < nseg) {
            sm.win_c[tid] = pfx;
            sm.win_g[tid] = gp & ~3ull;
            sm.win_src[tid] = swap;
            sm.seg_la[tid] = 4u * pfx + (uint32_t)(gp & 3);
        }
        if (tid == 0) {
            sm.win_c[nseg] = tot;
            sm.seg_u[nseg] = ulen;
        }
        __syncthreads();
        // ---- 1b. top level of the B-pool index over the tile's larger B sets, loaded together with
        // the probe windows (one round trip for both) ---------------------------------------------------
        bool top_ok = false;
        uint64_t tj0 = 0;
        uint64_t tn = 0;
        if (in.b_top && !bad) {
            gmin = ~0ull;
            gmax = 0;
#pragma unroll
            for (int k = 0; k < NW; k++) {
                gmin = min(gmin, sm.red64[k][0]);
                gmax = max(gmax, sm.red64[k][1]);
            }
            if (gmax > gmin) {
                // entries j with kTopStride j in [gmin, gmax)
                tj0 = (gmin + kTopStride - 1) / kTopStride;
                tn = (gmax + kTopStride - 1) / kTopStride - tj0;
                top_ok = tn <= (uint64_t)kTopCap;
            }
        }
        if (bad) {
            nseg = 0;
            ulen = 0;
            if (tid == 0 && out.error) atomicOr(out.error, 2u);
        } else {
            // the top level goes straight to LDS (4 bytes per lane, a wave's row of 64 entries per
            // load), in flight with the windows' loads and holding no registers meanwhile
            if (top_ok) {
                for (uint32_t r = tid - lane; r < (uint32_t)tn; r += NT)  // r: the wave's row (uniform)
                    if (r + lane < (uint32_t)tn)
                        __builtin_amdgcn_global_load_lds(in.b_top + tj0 + r + lane, sm.top + r, 4, 0, 0);
            }
            constexpr int QP = (Smem::PKEYS + 4 * NT - 1) / (4 * NT);
            stage_windows<NT, QP>(in.A, in.a_pool, in.B, in.b_pool, sm.win_c, sm.win_g, sm.win_src, nseg, sm.pkey,
                                  in.sflip);
            if constexpr (BSR)
                stage_windows<NT, QP>(in.As, in.a_pool, in.Bs, in.b_pool, sm.win_c, sm.win_g, sm.win_src, nseg, sm.pst,
                                      0u);
            if (top_ok) __builtin_amdgcn_s_waitcnt(0);  // the direct-to-LDS loads have landed
        }
        __syncthreads();

        // ---- 2. narrow the larger sets: one wave per segment ----------------------------------------
        // A larger set in an indexed B pool longer than kNoNarrow elements is not narrowed: its
        // window could be staged only if the probes clustered, and the two 64-way searches would
        // cost more HBM lines than the probes themselves; every probe finds its line through the
        // index instead.
        for (uint32_t s = wave; s < nseg; s += NW) {
            const uint32_t n = sm.seg_wa[s];
            const uint32_t nl = sm.seg_hi[s];
            uint32_t lo = 0, hi = 0;
            if (n && in.b_idx && !((sm.seg_flags[s] >> 2) & 1u) && nl > kNoNarrow) {
                hi = nl;
            } else if (n) {
                const bool large_is_a = (sm.seg_flags[s] >> 2) & 1u;
                const uint32_t *__restrict__ L = (large_is_a ? in.A : in.B) + sm.seg_lbase[s];
                const uint32_t vf = sm.pkey[sm.seg_la[s]], vl = sm.pkey[sm.seg_la[s] + n - 1];
                const uint32_t f = in.sflip;
                lo = (uint32_t)group_partition_point<kWave>(0, nl, lane, 0, [&](uint64_t x) { return (L[x] ^ f) < vf; });
                hi = (uint32_t)group_partition_point<kWave>(lo, nl, lane, 0, [&](uint64_t x) { return (L[x] ^ f) <= vl; });
            }
            if (lane == 0) {
                sm.seg_lo[s] = lo;
                sm.seg_hi[s] = hi;
            }
        }
        __syncthreads();

        // ---- 3. stage the narrowed larger windows while they fit --------------------------------------
        {
            uint32_t cl = 0;
            uint64_t gl = 0;
            uint32_t src = 0;
            if (tid < nseg) {
                const uint32_t n = sm.seg_hi[tid] - sm.seg_lo[tid];
                gl = sm.seg_lbase[tid] + sm.seg_lo[tid];
                src = ((sm.seg_flags[tid] >> 2) & 1u) ? 0u : 1u;  // pool of the larger side
                cl = n ? ((uint32_t)(gl & 3) + n + 3) / 4 : 0u;
            }
            uint32_t ltot = 0;
            const uint32_t lpfx = block_exclusive_scan<NT>(cl, sm.scan_scratch, &ltot);
            const bool fits = lpfx + cl <= (uint32_t)(CAPB / 4);
            // staged segments form a prefix; their count is the number of fitting threads
            uint32_t nst = 0;
            {
                const uint64_t bal = ballot(tid < nseg && fits);
                if (lane == 0) sm.wave_el[wave] = (uint32_t)__popcll(bal);
            }
            __syncthreads();
            for (int k = 0; k < NW; k++) nst += sm.wave_el[k];
            if (tid < nseg) {
                if (fits) {
                    sm.win_c[tid] = lpfx;
                    sm.win_g[tid] = gl & ~3ull;
                    sm.win_src[tid] = src;
                    sm.seg_lb[tid] = 4u * lpfx + (uint32_t)(gl & 3);
                } else {
                    sm.seg_lb[tid] = kNotStaged;
                }
                if (tid + 1 == nst) sm.win_c[nst] = lpfx + cl;
            }
            if (tid == 0 && nst == 0) sm.win_c[0] = 0;
            __syncthreads();
            if (nst) {
                constexpr int QL = (Smem::LKEYS + 4 * NT - 1) / (4 * NT);
                stage_windows<NT, QL>(in.A, in.a_pool, in.B, in.b_pool, sm.win_c, sm.win_g, sm.win_src, nst, sm.lkey,
                                      in.sflip);
                if constexpr (BSR)
                    stage_windows<NT, QL>(in.As, in.a_pool, in.Bs, in.b_pool, sm.win_c, sm.win_g, sm.win_src, nst,
                                          sm.lst, 0u);
            }
            __syncthreads();
        }

        // ---- 3b. further sets (a single pair, plain sets: small_adaptive over k sets) -----------------
        // Every probe found in the pair's larger set is also looked up in each further set in turn
        // (adaptive.rs:107-147); each further set is narrowed to the tile's probe range here, one
        // wave per set, and each lane gallops forward through it from there.
        uint32_t xcur[kMaxExtraSets];
        const uint32_t nx = (!BSR && nseg == 1 && sm.seg_wa[0] > 0) ? in.nx : 0u;
        if (nx) {
            const uint32_t vf = sm.pkey[sm.seg_la[0]], vl = sm.pkey[sm.seg_la[0] + sm.seg_wa[0] - 1];
            for (uint32_t j = wave; j < nx; j += NW) {
                const uint32_t *__restrict__ X = in.xsets[j];
                const uint64_t nxj = in.xlens[j];
                const uint32_t f = in.sflip;
                const uint32_t lo = (uint32_t)group_partition_point<kWave>(0, nxj, lane, 0,
                                                                           [&](uint64_t i) { return (X[i] ^ f) < vf; });
                const uint32_t hi = (uint32_t)group_partition_point<kWave>(
                    lo, nxj, lane, 0, [&](uint64_t i) { return (X[i] ^ f) <= vl; });
                if (lane == 0) {
                    sm.xlo[j] = lo;
                    sm.xhi[j] = hi;
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)kMaxExtraSets; j++) xcur[j] = (j < nx) ? sm.xlo[j] : 0u;
        }

        // ---- 4. lookups --------------------------------------------------------------------------
        const uint32_t ub = tid * KS;
        uint32_t hmask = 0, nelems = 0;
        uint32_t svals[BSR ? KS : 1];
        uint32_t s0 = 0;
        if (ub < ulen) {
            uint32_t lo = 0, hi = nseg;  // largest s with seg_u[s] <= ub
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sm.seg_u[mid] <= ub) lo = mid; else hi = mid;
            }
            s0 = lo;
        }
        // 4a. probes answered through the indexed B pool with its top level staged: looked up
        // first, the whole wave together (coop_index_probe_top), the k-th probe of every lane at
        // once; qrel[k] is the found position relative to the larger set, or ~0.  A probe's range
        // starts at its narrowed window rather than at the previous hit: a set holds x once, so the
        // position found is the same.
        uint32_t qrel[KS];
#pragma unroll
        for (int k = 0; k < KS; k++) qrel[k] = ~0u;
        if (top_ok && in.b_idx) {
            // one probe per lane at a time (two at once, NP = 2, measured no faster at two probes
            // per lane and slower at four, where the registers of the second cost occupancy)
            constexpr int NP = 1;
            uint32_t s = s0;
#pragma unroll
            for (int k0 = 0; k0 < KS; k0 += NP) {
                bool on[NP];
                uint32_t x[NP];
                uint64_t g[NP], g0[NP], g1[NP], q[NP];
                bool any = false;
#pragma unroll
                for (int j = 0; j < NP; j++) {
                    const uint32_t u = ub + (uint32_t)(k0 + j);
                    on[j] = k0 + j < KS && u < ulen;
                    x[j] = 0;
                    g[j] = g0[j] = g1[j] = 0;
                    if (on[j]) {
                        while (u >= sm.seg_u[s + 1]) s++;
                        const uint32_t e = u - sm.seg_u[s];
                        on[j] = e < sm.seg_wa[s] && sm.seg_lb[s] == kNotStaged && !((sm.seg_flags[s] >> 2) & 1u);
                        if (on[j]) {
                            x[j] = sm.pkey[sm.seg_la[s] + e];
                            g[j] = sm.seg_lbase[s];
                            g0[j] = g[j] + sm.seg_lo[s];
                            g1[j] = g[j] + sm.seg_hi[s];
                        }
                    }
                    any = any || on[j];
                }
                if (ballot(any) != 0) {  // wave-uniform
                    coop_index_probe_top<NP>(on, in.B, in.b_pool, in.b_idx, in.b_mid, sm.top, tj0, g0, g1, x, in.sflip,
                                             q);
#pragma unroll
                    for (int j = 0; j < NP; j++)
                        if (k0 + j < KS && q[j] != kNoPos) qrel[k0 + j] = (uint32_t)(q[j] - g[j]);
                }
            }
        }
        if (ub < ulen) {
            uint32_t s = s0;
            uint32_t cur = sm.seg_lo[s];  // pair-relative position in the larger set
#pragma unroll
            for (int k = 0; k < KS; k++) {
                const uint32_t u = ub + (uint32_t)k;
                if (u >= ulen) break;
                if (u >= sm.seg_u[s + 1]) {
                    s++;
                    cur = sm.seg_lo[s];
                }
                const uint32_t e = u - sm.seg_u[s];
                if (e >= sm.seg_wa[s]) continue;  // padding unit
                const uint32_t x = sm.pkey[sm.seg_la[s] + e];
                const uint32_t lo_s = sm.seg_lo[s], hi_s = sm.seg_hi[s], lb = sm.seg_lb[s];
                const bool large_is_a = (sm.seg_flags[s] >> 2) & 1u;
                uint32_t pos;
                bool found;
                uint32_t lstate = 0;
                if (lb != kNotStaged) {
                    auto arr = [&](uint32_t i) { return sm.lkey[lb + (i - lo_s)]; };
                    if constexpr (SMODE == SM_BLOCK) pos = lower_bound_block<kBlockLds>(arr, cur, hi_s, x);
                    else pos = lower_bound_from<SMODE == SM_GALLOP>(arr, cur, hi_s, x);
                    found = pos < hi_s && arr(pos) == x;
                    if constexpr (BSR)
                        if (found) lstate = sm.lst[lb + (pos - lo_s)];
                } else if (in.b_idx && !large_is_a) {
                    // resident B pool with an index of every 32nd element: binary search the index,
                    // then compare one 128-byte line of B (with the top level staged: looked up in 4a)
                    const uint64_t g = sm.seg_lbase[s];
                    const uint64_t q = top_ok ? (qrel[k] != ~0u ? g + qrel[k] : kNoPos)
                                              : index_probe(in.B, in.b_pool, in.b_idx, in.b_mid, g + cur, g + hi_s, x,
                                                            in.sflip);
                    found = q != kNoPos;
                    pos = found ? (uint32_t)(q - g) : cur;
                    if constexpr (BSR)
                        if (found) lstate = in.Bs[q];
                } else {
                    if constexpr (SMODE == SM_BLOCK) {
                        pos = lower_bound_block_pool(large_is_a ? in.A : in.B, large_is_a ? in.a_pool : in.b_pool,
                                                     sm.seg_lbase[s], cur, hi_s, x, in.sflip, found);
                    } else {
                        const uint32_t *__restrict__ L = (large_is_a ? in.A : in.B) + sm.seg_lbase[s];
                        const uint32_t f = in.sflip;
                        auto arr = [&](uint32_t i) { return L[i] ^ f; };
                        pos = lower_bound_from<SMODE == SM_GALLOP>(arr, cur, hi_s, x);
                        found = pos < hi_s && arr(pos) == x;
                    }
                    if constexpr (BSR)
                        if (found) lstate = (large_is_a ? in.As : in.Bs)[sm.seg_lbase[s] + pos];
                }
                cur = pos;
                if constexpr (!BSR) {
                    // further sets in order; the first one that lacks x ends the lookup
                    for (uint32_t j = 0; found && j < nx; j++) {
                        const uint32_t *__restrict__ X = in.xsets[j];
                        const uint32_t f = in.sflip;
                        auto xarr = [&](uint32_t i) { return X[i] ^ f; };
                        const uint32_t xh = sm.xhi[j];
                        const uint32_t xp = lower_bound_from<SMODE != SM_BINARY>(xarr, xcur[j], xh, x);
                        xcur[j] = xp;
                        found = xp < xh && xarr(xp) == x;
                    }
                }
                if constexpr (BSR) {
                    const uint32_t stv = found ? (sm.pst[sm.seg_la[s] + e] & lstate) : 0u;
                    found = stv != 0;
                    svals[k] = stv;
                    nelems += __popc(stv);
                }
                hmask |= found ? (1u << k) : 0u;
            }
        }
        if constexpr (!BSR) {
            // k-set groups: each hit is looked up in its pair's further sets in B, in order; the
            // first that lacks it drops it (small_adaptive's walk, adaptive.rs:107-147, one element
            // at a time: each lookup searches the whole further set).  A pass of its own after the
            // lookups, so its state is not live beside theirs.
            if (in.x_off && hmask) {
                uint32_t s = s0;
                for (int k = 0; k < KS; k++) {
                    if (!((hmask >> k) & 1u)) continue;
                    const uint32_t u = ub + (uint32_t)k;
                    while (u >= sm.seg_u[s + 1]) s++;
                    const uint32_t x = sm.pkey[sm.seg_la[s] + (u - sm.seg_u[s])];
                    const uint64_t p = (uint64_t)p_first + s;
                    const uint64_t j1 = in.x_off[p + 1];
                    bool found = true;
                    for (uint64_t j = in.x_off[p]; found && j < j1; j++) {
                        // a range is read only inside the pool (and its index)
                        const uint64_t g1 = min(in.x_end[j], in.b_pool), g0 = min(in.x_beg[j], g1);
                        if (in.b_idx) {
                            found = index_probe(in.B, in.b_pool, in.b_idx, in.b_mid, g0, g1, x, in.sflip) != kNoPos;
                        } else if (g1 > g0) {
                            const uint32_t n = (uint32_t)(g1 - g0);
                            if constexpr (SMODE == SM_BLOCK) {
                                (void)lower_bound_block_pool(in.B, in.b_pool, g0, 0u, n, x, in.sflip, found);
                            } else {
                                const uint32_t *__restrict__ X = in.B + g0;
                                const uint32_t f = in.sflip;
                                auto xarr = [&](uint32_t i) { return X[i] ^ f; };
                                const uint32_t xp = lower_bound_from<false>(xarr, 0u, n, x);
                                found = xp < n && xarr(xp) == x;
                            }
                        } else {
                            found = false;
                        }
                    }
                    if (!found) hmask &= ~(1u << k);
                }
            }
        }
        const uint32_t nwords = __popc(hmask);
        if constexpr (!BSR) nelems = nwords;

        // ---- 5. tile scan, staged outputs, counts ------------------------------------------------------
        const uint32_t pfx2 = block_exclusive_scan<NT>(nwords, sm.scan_scratch, &tot);
        sm.lane_mask[tid] = hmask;
        sm.lane_pfx[tid] = pfx2;
        if constexpr (BSR) {
            uint32_t el = nelems;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) el += __shfl_xor(el, dd, kWave);
            if (lane == 0) sm.wave_el[wave] = el;
        }
        __syncthreads();
        if constexpr (!COUNT_ONLY) {
            uint32_t *dk = plan.stage + (uint64_t)t * plan.stage_words;
            uint32_t *ds = dk + plan.stage_words / 2;
            uint32_t pos = pfx2;
            uint32_t s = s0;
#pragma unroll
            for (int k = 0; k < KS; k++) {
                const uint32_t u = ub + (uint32_t)k;
                if (hmask & (1u << k)) {
                    while (u >= sm.seg_u[s + 1]) s++;
                    dk[pos] = sm.pkey[sm.seg_la[s] + (u - sm.seg_u[s])] ^ in.sflip;  // the value as given
                    if constexpr (BSR) ds[pos] = svals[k];
                    pos++;
                }
            }
        }
        if (tid == 0) {
            uint32_t el = tot;
            if constexpr (BSR) {
                el = 0;
                for (int k = 0; k < NW; k++) el += sm.wave_el[k];
            }
            plan.tile_count[t] = tot;
            plan.tile_elems[t] = el;
            if (plan.one_tile) {  // one tile: its counts are the totals, its results the output
                if (out.totals) {
                    out.totals[0] = tot;
                    out.totals[1] = el;
                }
                if (out.out_off) out.out_off[in.P] = tot;
            }
        }
        if (tid < nseg) {
            auto before = [&](uint32_t x) -> uint32_t {
                if (x >= ulen) return tot;
                const uint32_t l = x / KS, k = x % KS;
                return sm.lane_pfx[l] + __popc(sm.lane_mask[l] & ((1u << k) - 1u));
            };
            const uint32_t c0v = before(sm.seg_u[tid]);
            const uint32_t c = before(sm.seg_u[tid + 1]) - c0v;
            const uint32_t f = sm.seg_flags[tid];
            const uint64_t p = (uint64_t)p_first + tid;
            if (f & 1u) {
                plan.pair_local[p] = c0v;
                if (dbatch) plan.uoff[p] = seg_pu;  // for the placement (no plan wrote it)
            }
            if (plan.one_tile && out.out_off) out.out_off[p] = c0v;  // every pair lies in the one tile
            if (out.counts) {
                if ((f & 3u) == 3u) out.counts[p] = c;
                else if (c) atomicAdd(&out.counts[p], c);
            }
        }
    }
}

} // namespace setops
