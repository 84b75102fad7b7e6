// merge_tile.hpp -- batched sorted-set intersection over a global unit space (gfx950).
//
// Work model.  Pair p of a batch owns units [uoff[p], uoff[p+1]).  For the merge policy a pair
// has |A|+|B| real units (one per merge-path diagonal step) plus c0 padding units; the padding
// bounds how many pairs one tile can touch (<= U/c0 + 2), which sizes the LDS segment table.
// Tile t covers units [t*U, (t+1)*U) of the concatenated unit space, so tiny pairs are packed
// many to a tile (several per lane), medium pairs span lanes of one tile, and huge pairs are
// split across tiles by merge-path -- one kernel, no per-size launches, perfectly even tiles.
//
// Pipeline (count -> scan -> write), no communication between workgroups:
//   plan_csr_kernel     unit prefix, tile -> pair maps, zeroed per-pair counts
//   plan_split_kernel   merge-path split of every tile boundary (one thread each)
//   merge_tile_kernel   one workgroup per tile: count and stage the tile's results
//   scan_tiles_kernel   exclusive scan of the tile counts
//   place_tiles_kernel  move staged results to their final positions, per-pair offsets
//
// Per tile (merge_tile_kernel):
//   1. read the tile's precomputed boundaries and the covered pairs' descriptors,
//   3. stage the tile's A range and B range (each contiguous in the CSR pools) in LDS with
//      16-byte loads; the B range carries one element past the tile for the boundary compare,
//   4. every lane finds its K-unit slice by an LDS merge-path search and walks K merge steps,
//      emitting on an A step whose value equals the current B value (exactly once per match,
//      independent of where tile and lane boundaries fall),
//   5. block scan of per-lane match counts; results written in order to the tile's staging slot
//      straight from LDS (each wave's results form one contiguous run),
//   6. tile count, per-pair counts and in-tile offsets.
//
// The same kernel handles the BSR representation (key = base, payload = 32-bit state): a base
// match emits (base, state_a & state_b) when the AND is non-zero.
#pragma once

#include "tile_common.hpp"

namespace setops {

constexpr uint32_t kScanGroup = 1024;  // tiles per group in the two-level tile scan

struct PairBatch {
    const uint32_t *A;   // A keys (values, or BSR bases)
    const uint32_t *As;  // A BSR states (null for plain sets)
    const uint32_t *B;
    const uint32_t *Bs;
    const uint64_t *a_off;  // [P+1] CSR offsets into A
    const uint64_t *b_off;  // [P+1] CSR offsets into B
    uint64_t a_pool;        // elements readable at A (loads never go past)
    uint64_t b_pool;
    uint64_t P;
};

struct TilePlan {
    uint64_t *uoff;        // [P+1] unit prefix
    uint32_t *tile_first;  // [ntiles] pair containing the tile's first unit
    uint32_t *tile_last;   // [ntiles] pair containing the tile's last unit
    uint64_t *tile_ga;     // [ntiles+1] A pool index where tile t starts
    uint64_t *tile_gb;     // [ntiles+1] B pool index where tile t starts
    uint32_t *tile_count;  // [ntiles] result words of tile t
    uint32_t *tile_elems;  // [ntiles] result elements of tile t (popcount for BSR)
    uint64_t *tile_off;    // [ntiles+1] exclusive scan of tile_count
    unsigned long long *group_cnt;  // [ngroups] sum of tile_count per group of kScanGroup tiles (reduce kernel)
    unsigned long long *group_el;   // [ngroups] sum of tile_elems per group
    uint32_t *stage;       // [ntiles * stage_words] per-tile staged outputs (keys, then BSR states)
    uint32_t *pair_local;  // [P] offset of a pair's first result inside the tile where it starts
    uint64_t utotal;
    uint32_t ntiles;
    uint32_t c0;           // padding units per pair
    uint32_t U;            // units per tile
    uint32_t stage_words;  // words per tile in `stage` (keys; states follow at +stage_words/2)
    // single-pair mode: the plan kernel writes a_off = {0, na}, b_off = {0, nb} to single_off
    uint32_t single;
    uint64_t *single_off;
    uint64_t single_na, single_nb;
};

struct TileOut {
    uint32_t *out;          // compacted keys (null in count-only mode)
    uint32_t *out_s;        // compacted BSR states
    uint64_t cap;           // capacity of out / out_s in elements
    uint64_t *out_off;      // [P+1] output offset of each pair (may be null)
    uint32_t *counts;       // [P] per-pair output count (may be null)
    unsigned long long *totals;  // [2]: output words, output elements (popcount for BSR)
    uint32_t *error;        // bit 0: output capacity exceeded
};

// Prologue: unit prefix (closed form for CSR pools), tile -> pair maps, zeroed look-back words,
// claim counter, per-pair counts and totals.  One thread per pair plus grid-stride loops.
__global__ void plan_csr_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gsz = (uint64_t)gridDim.x * blockDim.x;
    auto aoff = [&](uint64_t p) -> uint64_t { return plan.single ? (p ? plan.single_na : 0) : in.a_off[p]; };
    auto boff = [&](uint64_t p) -> uint64_t { return plan.single ? (p ? plan.single_nb : 0) : in.b_off[p]; };
    if (plan.single && gid == 0) {
        plan.single_off[0] = 0;
        plan.single_off[1] = plan.single_na;
        plan.single_off[2] = 0;
        plan.single_off[3] = plan.single_nb;
    }
    const uint64_t a0 = aoff(0), b0 = boff(0);
    const uint64_t U = plan.U;
    for (uint64_t p = gid; p <= in.P; p += gsz) {
        const uint64_t u = (aoff(p) - a0) + (boff(p) - b0) + (uint64_t)plan.c0 * p;
        plan.uoff[p] = u;
        if (p < in.P) {
            const uint64_t un = (aoff(p + 1) - a0) + (boff(p + 1) - b0) + (uint64_t)plan.c0 * (p + 1);
            // tiles whose first unit t*U lies in [u, un)
            for (uint64_t t = (u + U - 1) / U; t * U < un && t < plan.ntiles; t++) plan.tile_first[t] = (uint32_t)p;
            // tiles whose last unit min((t+1)U, utotal)-1 lies in [u, un)
            uint64_t tl = u / U;  // tile containing unit u
            for (uint64_t t = tl; t < plan.ntiles; t++) {
                uint64_t last = (t + 1) * U;
                if (last > plan.utotal) last = plan.utotal;
                last -= 1;
                if (last >= un) break;
                if (last >= u) plan.tile_last[t] = (uint32_t)p;
            }
            if (out.counts) out.counts[p] = 0;
        }
    }
    if (gid == 0) {
        if (out.totals) {
            out.totals[0] = 0;
            out.totals[1] = 0;
        }
        if (out.error) *out.error = 0;
    }
}

// Tile boundaries: for every t in [0, ntiles], the global A / B element index where tile t starts
// (the merge-path split of unit t*U inside its pair).  Tile t stages A[ga(t) .. ga(t+1)) and
// B[gb(t) .. gb(t+1)] without searching anything itself.
__global__ void plan_split_kernel(PairBatch in, TilePlan plan)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > plan.ntiles) return;
    if (t == plan.ntiles) {
        plan.tile_ga[t] = in.a_off[in.P];
        plan.tile_gb[t] = in.b_off[in.P];
        return;
    }
    const uint64_t p = plan.tile_first[t];
    const uint64_t a = in.a_off[p], b = in.b_off[p];
    const uint64_t na = in.a_off[p + 1] - a, nb = in.b_off[p + 1] - b;
    const uint64_t rel = (uint64_t)t * plan.U - plan.uoff[p];
    const uint64_t d = rel < na + nb ? rel : na + nb;
    uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    const uint32_t *__restrict__ A = in.A + a;
    const uint32_t *__restrict__ B = in.B + b;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - 1 - mid]) lo = mid + 1; else hi = mid;
    }
    plan.tile_ga[t] = a + lo;
    plan.tile_gb[t] = b + (d - lo);
}

template <int NT, int K, int MAXSEG, bool BSR>
struct MergeSmem {
    static constexpr int U = NT * K;
    uint32_t key[U + 32];
    uint32_t st[BSR ? U + 32 : 4];
    uint32_t seg_u[MAXSEG + 1];  // tile-relative unit where segment s starts; [nseg] = tile length
    uint32_t seg_la[MAXSEG];     // LDS index of the segment's first A element in the window
    uint32_t seg_lb[MAXSEG];     // LDS index of the segment's first B element in the window
    uint32_t seg_wa[MAXSEG];     // A elements in the window
    uint32_t seg_wb[MAXSEG];     // B elements in the window (the extra boundary element excluded)
    uint32_t seg_nb[MAXSEG];     // B elements of the pair from the window start (bounds the match check)
    uint32_t seg_flags[MAXSEG];  // bit0: pair starts in this tile, bit1: pair ends in this tile
    uint32_t lane_mask[NT];      // per-lane step mask of matches
    uint32_t lane_pfx[NT];       // per-lane exclusive prefix of matches
    uint32_t scan_scratch[NT / kWave];
    uint32_t wave_el[NT / kWave];
};

// A tile's boundaries: A range [ga0, ga1), B range [gb0, gb1] (+ the boundary element), pairs
// p_first .. p_last.  Uniform across the workgroup (scalar loads).
struct TileDesc {
    uint64_t ga0, ga1, gb0, gb1;
    uint32_t p_first, p_last;
};

__device__ __forceinline__ TileDesc load_desc(const TilePlan &plan, uint32_t t)
{
    TileDesc d;
    d.ga0 = plan.tile_ga[t];
    d.ga1 = plan.tile_ga[t + 1];
    d.gb0 = plan.tile_gb[t];
    d.gb1 = plan.tile_gb[t + 1];
    d.p_first = plan.tile_first[t];
    d.p_last = plan.tile_last[t];
    return d;
}

// Geometry of a tile's LDS image, derived from its descriptor (uniform).
struct TileGeom {
    uint32_t ulen, nseg, n_a, n_b, shift_a, shift_b, boff;
    bool bad;
};

template <int U, int MAXSEG>
__device__ __forceinline__ TileGeom tile_geom(const PairBatch &in, const TilePlan &plan, const TileDesc &d, uint32_t t)
{
    TileGeom g;
    const uint64_t u0 = (uint64_t)t * U;
    const uint64_t u1 = (u0 + U < plan.utotal) ? u0 + U : plan.utotal;
    g.ulen = (uint32_t)(u1 - u0);
    g.nseg = d.p_last - d.p_first + 1;
    g.bad = g.nseg == 0 || g.nseg > (uint32_t)MAXSEG || d.ga1 < d.ga0 || d.gb1 < d.gb0 ||
            (d.ga1 - d.ga0) + (d.gb1 - d.gb0) > (uint64_t)U;
    if (g.bad) {
        g.nseg = 0;
        g.ulen = 0;
        g.n_a = g.n_b = 0;
    } else {
        g.n_a = (uint32_t)(d.ga1 - d.ga0);
        g.n_b = (uint32_t)(d.gb1 - d.gb0) + ((d.gb1 < in.b_pool) ? 1u : 0u);  // + boundary element
    }
    g.shift_a = (uint32_t)(d.ga0 & 3);
    g.shift_b = (uint32_t)(d.gb0 & 3);
    g.boff = 4u * ((g.shift_a + g.n_a + 3) / 4);
    return g;
}

// Copy `n` elements starting at global element index g0 of `src` (pool length `pool`) into LDS
// at dst[shift ...], where shift = g0 & 3: 16-byte loads of the aligned chunks covering the range,
// one chunk per thread per iteration (coalesced), the pool's last partial chunk element-wise.
template <int NT>
__device__ __forceinline__ void stage_range(const uint32_t *__restrict__ src, uint64_t pool, uint64_t g0, uint32_t n,
                                            uint32_t *dst)
{
    const uint64_t base = g0 & ~3ull;
    const uint32_t span = (uint32_t)(g0 - base) + n;
    const uint32_t chunks = (span + 3) / 4;
    for (uint32_t c = threadIdx.x; c < chunks; c += NT) {
        const uint64_t e = base + 4ull * c;
        uint4 v;
        if (e + 4 <= pool) {
            v = *reinterpret_cast<const uint4 *>(src + e);
        } else {
            v.x = (e + 0 < pool) ? src[e + 0] : 0u;
            v.y = (e + 1 < pool) ? src[e + 1] : 0u;
            v.z = (e + 2 < pool) ? src[e + 2] : 0u;
            v.w = (e + 3 < pool) ? src[e + 3] : 0u;
        }
        *reinterpret_cast<uint4 *>(dst + 4 * c) = v;
    }
}

// Segment table of a tile: one entry per covered pair (thread `s` handles pair p_first + s).
template <int NT, int K, int MAXSEG, bool BSR>
__device__ __forceinline__ void segment_table(MergeSmem<NT, K, MAXSEG, BSR> &sm, const PairBatch &in,
                                              const TilePlan &plan, const TileDesc &d, const TileGeom &g, uint32_t t)
{
    constexpr int U = NT * K;
    const uint32_t tid = threadIdx.x;
    if (tid < g.nseg) {
        const uint64_t p = (uint64_t)d.p_first + tid;
        const uint64_t up = plan.uoff[p], upn = plan.uoff[p + 1];
        const uint64_t a = in.a_off[p], b = in.b_off[p];
        const uint64_t na = in.a_off[p + 1] - a, nb = in.b_off[p + 1] - b;
        const uint64_t u0 = (uint64_t)t * U;
        const uint64_t u1 = u0 + g.ulen;
        const uint64_t i0 = (tid == 0) ? d.ga0 - a : 0, j0 = (tid == 0) ? d.gb0 - b : 0;
        const uint64_t i1 = (tid == g.nseg - 1) ? d.ga1 - a : na, j1 = (tid == g.nseg - 1) ? d.gb1 - b : nb;
        sm.seg_u[tid] = (up > u0) ? (uint32_t)(up - u0) : 0u;
        sm.seg_flags[tid] = (up >= u0 ? 1u : 0u) | (upn <= u1 ? 2u : 0u);
        sm.seg_la[tid] = g.shift_a + (uint32_t)(a + i0 - d.ga0);
        sm.seg_lb[tid] = g.boff + g.shift_b + (uint32_t)(b + j0 - d.gb0);
        sm.seg_wa[tid] = (uint32_t)(i1 - i0);
        sm.seg_wb[tid] = (uint32_t)(j1 - j0);
        sm.seg_nb[tid] = (uint32_t)(nb - j0);
    }
    if (tid == 0) sm.seg_u[g.nseg] = g.ulen;
}

// Count + stage: one workgroup per tile, no communication between workgroups (nothing ever waits
// on another workgroup).  Each tile's results go to its own staging slot (or nowhere in
// count-only mode); its counts go to tile_count / tile_elems; per-pair counts and in-tile offsets
// are recorded for the placement kernel.
template <int NT, int K, int MAXSEG, bool BSR, bool COUNT_ONLY>
__global__ __launch_bounds__(NT) void merge_tile_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    static_assert(K >= 1 && K <= 32, "lane masks hold K bits");
    constexpr int U = NT * K;
    __shared__ MergeSmem<NT, K, MAXSEG, BSR> sm;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / kWave;
    const uint32_t t = blockIdx.x;
    if (t >= plan.ntiles) return;

    // ---- 1-2. descriptor, inputs and segment table into LDS -------------------------------------
    const TileDesc d0 = load_desc(plan, t);
    const TileGeom g0 = tile_geom<U, MAXSEG>(in, plan, d0, t);
    if (g0.bad && tid == 0 && out.error) atomicOr(out.error, 2u);
    if (!g0.bad) {
        stage_range<NT>(in.A, in.a_pool, d0.ga0, g0.n_a, sm.key);
        stage_range<NT>(in.B, in.b_pool, d0.gb0, g0.n_b, sm.key + g0.boff);
        if constexpr (BSR) {
            stage_range<NT>(in.As, in.a_pool, d0.ga0, g0.n_a, sm.st);
            stage_range<NT>(in.Bs, in.b_pool, d0.gb0, g0.n_b, sm.st + g0.boff);
        }
    }
    segment_table<NT, K, MAXSEG, BSR>(sm, in, plan, d0, g0, t);
    __syncthreads();
    {
        const uint32_t ulen = g0.ulen, nseg = g0.nseg;
        const uint32_t p_first = d0.p_first;

        // ---- 3. per-lane merge --------------------------------------------------------------------
        // smask: bit k set when step k emitted; tmask: bit k set when step k consumed an A element.
        // The emitted values are A elements, re-read from LDS at write time: the A offset of the
        // emit at step k is popcount(tmask & ((1 << k) - 1)).
        const uint32_t ub = tid * K;
        uint32_t smask = 0, tmask = 0, pa0 = 0;
        uint32_t svals[BSR ? K : 1];
        uint32_t nelems = 0;
        uint32_t kmax = 0, kt = K;  // steps to run; step at which the lane enters the next pair
        uint32_t pa = 0, pb = 0, ea = 0, eb = 0, ebp = 0, s = 0;
        if (ub < ulen) {
            uint32_t lo = 0, hi = nseg;  // largest s with seg_u[s] <= ub (seg_u[0] == 0)
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sm.seg_u[mid] <= ub) lo = mid; else hi = mid;
            }
            s = lo;
            const uint32_t la = sm.seg_la[s], lb = sm.seg_lb[s];
            ea = la + sm.seg_wa[s];
            eb = lb + sm.seg_wb[s];
            ebp = lb + sm.seg_nb[s];
            const uint32_t e = ub - sm.seg_u[s];
            const uint32_t wa = ea - la, wb = eb - lb;
            if (e >= wa + wb) {
                pa = ea;
                pb = eb;
            } else {
                uint32_t xlo = (e > wb) ? e - wb : 0u, xhi = (e < wa) ? e : wa;
                while (xlo < xhi) {
                    const uint32_t mid = (xlo + xhi) >> 1;
                    if (sm.key[la + mid] <= sm.key[lb + e - 1 - mid]) xlo = mid + 1; else xhi = mid;
                }
                pa = la + xlo;
                pb = lb + (e - xlo);
            }
            kmax = (ulen - ub < (uint32_t)K) ? ulen - ub : (uint32_t)K;
            const uint32_t send = sm.seg_u[s + 1];  // > ub
            if (send - ub < kmax) kt = send - ub;   // every pair owns >= c0 >= K units: at most one
        }
        pa0 = pa;
        uint32_t a = sm.key[pa], b = sm.key[pb];
        // One merge step, branch-free: consume A when it is not exhausted and not larger than the
        // current B (or B is exhausted); emit when the consumed A equals the current B.  One LDS
        // read per step (the side that advanced).
        auto step = [&](uint32_t k) {
            const bool aok = pa < ea, bok = pb < eb;
            const bool takeA = aok && (!bok || a <= b);
            bool hit = takeA && pb < ebp && a == b;
            if constexpr (BSR) {
                uint32_t stv = 0;
                if (hit) stv = sm.st[pa] & sm.st[pb];
                hit = hit && stv != 0;
                svals[k] = stv;
                nelems += __popc(stv);
            }
            smask |= hit ? (1u << k) : 0u;
            tmask |= takeA ? (1u << k) : 0u;
            pa += takeA ? 1u : 0u;
            pb += (!takeA && bok) ? 1u : 0u;
            const uint32_t v = sm.key[takeA ? pa : pb];
            a = takeA ? v : a;
            b = takeA ? b : v;
        };
        const bool uniform = ballot(kmax != (uint32_t)K || kt != (uint32_t)K) == 0;
        if (uniform) {
            // no lane of this wave ends its tile early or crosses into another pair
#pragma unroll
            for (int k = 0; k < K; k++) step((uint32_t)k);
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) {
                if ((uint32_t)k == kt) {
                    // next pair: its window starts at the pair's first elements (contiguous in LDS)
                    s++;
                    pa = sm.seg_la[s];
                    pb = sm.seg_lb[s];
                    ea = pa + sm.seg_wa[s];
                    eb = pb + sm.seg_wb[s];
                    ebp = pb + sm.seg_nb[s];
                    a = sm.key[pa];
                    b = sm.key[pb];
                }
                if ((uint32_t)k < kmax) step((uint32_t)k);
            }
        }
        const uint32_t nwords = __popc(smask);
        if constexpr (!BSR) nelems = nwords;

        // ---- 4. tile scan -----------------------------------------------------------------------
        uint32_t tot = 0;
        const uint32_t pfx = block_exclusive_scan<NT>(nwords, sm.scan_scratch, &tot);
        sm.lane_mask[tid] = smask;
        sm.lane_pfx[tid] = pfx;
        if constexpr (BSR) {
            uint32_t el = nelems;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) el += __shfl_xor(el, dd, kWave);
            if (lane == 0) sm.wave_el[wave] = el;
        }
        __syncthreads();

        // ---- 5. staged outputs (in order inside the tile's slot) ----------------------------------
        if constexpr (!COUNT_ONLY) {
            uint32_t *dk = plan.stage + (uint64_t)t * plan.stage_words;
            uint32_t *ds = dk + plan.stage_words / 2;
            // the A offset of the emit at step k counts the A steps before it; after a move to
            // the next pair the A window continues contiguously in LDS, so pa0 + offset still holds
            uint32_t pos = pfx;
            if constexpr (BSR) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (smask & (1u << k)) {
                        dk[pos] = sm.key[pa0 + __popc(tmask & ((1u << k) - 1u))];
                        ds[pos] = svals[k];
                        pos++;
                    }
                }
            } else {
                uint32_t sm_ = smask;
                while (sm_) {
                    const uint32_t k = __builtin_ctz(sm_);
                    sm_ &= sm_ - 1;
                    dk[pos++] = sm.key[pa0 + __popc(tmask & ((1u << k) - 1u))];
                }
            }
        }

        // ---- 6. tile counts, group sums, per-pair counts and in-tile offsets ----------------------
        if (tid == 0) {
            uint32_t el = tot;
            if constexpr (BSR) {
                el = 0;
                for (int k = 0; k < NT / kWave; k++) el += sm.wave_el[k];
            }
            plan.tile_count[t] = tot;
            plan.tile_elems[t] = el;
        }
        if (tid < nseg) {
            auto before = [&](uint32_t x) -> uint32_t {
                if (x >= ulen) return tot;
                const uint32_t l = x / K, k = x % K;
                return sm.lane_pfx[l] + __popc(sm.lane_mask[l] & ((1u << k) - 1u));
            };
            const uint32_t c0v = before(sm.seg_u[tid]);
            const uint32_t c = before(sm.seg_u[tid + 1]) - c0v;
            const uint32_t f = sm.seg_flags[tid];
            const uint64_t p = (uint64_t)p_first + tid;
            if (f & 1u) plan.pair_local[p] = c0v;
            if (out.counts) {
                if ((f & 3u) == 3u) out.counts[p] = c;
                else if (c) atomicAdd(&out.counts[p], c);
            }
        }
    }
}

// Group sums of the tile counts (one workgroup per group of kScanGroup tiles, no atomics).
template <int NT>
__global__ __launch_bounds__(NT) void reduce_tiles_kernel(TilePlan plan)
{
    static_assert(NT == (int)kScanGroup, "one thread per tile of a group");
    __shared__ unsigned long long red[NT / kWave][2];
    const uint32_t i = blockIdx.x * kScanGroup + threadIdx.x;
    unsigned long long c = 0, e = 0;
    if (i < plan.ntiles) {
        c = plan.tile_count[i];
        e = plan.tile_elems[i];
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        c += __shfl_xor(c, d, kWave);
        e += __shfl_xor(e, d, kWave);
    }
    if (lane_id() == 0) {
        red[threadIdx.x / kWave][0] = c;
        red[threadIdx.x / kWave][1] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sc = 0, se = 0;
        for (int w = 0; w < NT / kWave; w++) {
            sc += red[w][0];
            se += red[w][1];
        }
        plan.group_cnt[blockIdx.x] = sc;
        plan.group_el[blockIdx.x] = se;
    }
}

// Scan: one workgroup per group of kScanGroup tiles.  The group's base is the sum of the earlier
// group sums; a block scan of the group's tile counts gives tile_off.  The last group writes the
// total (and the totals).
template <int NT>
__global__ __launch_bounds__(NT) void scan_tiles_kernel(TilePlan plan, TileOut out)
{
    static_assert(NT == (int)kScanGroup, "one thread per tile of a group");
    __shared__ uint32_t scratch[NT / kWave];
    __shared__ unsigned long long red[NT / kWave][2];
    const uint32_t g = blockIdx.x;
    const uint32_t ngroups = (plan.ntiles + kScanGroup - 1) / kScanGroup;
    // base = sum of group_cnt[0 .. g), and (last group) the totals
    const bool last = g == ngroups - 1;
    unsigned long long base = 0, el = 0;
    for (uint32_t k = threadIdx.x; k < (last ? ngroups : g); k += NT) {
        const unsigned long long c = plan.group_cnt[k];
        if (k < g) base += c;
        if (last) el += plan.group_el[k];
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        base += __shfl_xor(base, d, kWave);
        el += __shfl_xor(el, d, kWave);
    }
    if (lane_id() == 0) {
        red[threadIdx.x / kWave][0] = base;
        red[threadIdx.x / kWave][1] = el;
    }
    __syncthreads();
    base = 0;
    el = 0;
    for (int w = 0; w < NT / kWave; w++) {
        base += red[w][0];
        el += red[w][1];
    }
    const uint32_t i = g * kScanGroup + threadIdx.x;
    const uint32_t c = (i < plan.ntiles) ? plan.tile_count[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<NT>(c, scratch, &tot);
    if (i < plan.ntiles) plan.tile_off[i] = base + ex;
    if (last && threadIdx.x == 0) {
        const unsigned long long total = base + tot;
        plan.tile_off[plan.ntiles] = total;
        if (out.totals) {
            out.totals[0] = total;
            out.totals[1] = el;
        }
        if (out.out && total > out.cap && out.error) atomicOr(out.error, 1u);
    }
}

// Write: moves each tile's staged results to its final position and turns in-tile pair offsets
// into output offsets.  One wave per tile, NT/64 tiles per workgroup.
template <int NT, bool BSR>
__global__ __launch_bounds__(NT) void place_tiles_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    const uint32_t t = blockIdx.x * (NT / kWave) + threadIdx.x / kWave;
    const uint32_t l = lane_id();
    if (t >= plan.ntiles) return;
    const uint64_t base = plan.tile_off[t];
    const uint32_t n = plan.tile_count[t];
    if (out.out) {
        const uint32_t *sk = plan.stage + (uint64_t)t * plan.stage_words;
        const uint32_t *ss = sk + plan.stage_words / 2;
        uint64_t lim = n;
        if (base + n > out.cap) lim = (base < out.cap) ? out.cap - base : 0;
        for (uint32_t x = l; x < lim; x += kWave) {
            out.out[base + x] = sk[x];
            if constexpr (BSR) out.out_s[base + x] = ss[x];
        }
    }
    if (out.out_off) {
        const uint64_t u0 = (uint64_t)t * plan.U;
        const uint32_t pf = plan.tile_first[t], pl = plan.tile_last[t];
        for (uint32_t p = pf + l; p <= pl; p += kWave)
            if (plan.uoff[p] >= u0) out.out_off[p] = base + plan.pair_local[p];
        if (t == plan.ntiles - 1 && l == 0) out.out_off[in.P] = plan.tile_off[plan.ntiles];
    }
}

} // namespace setops
