// tile_plan.hpp -- batch description, unit-space planning, tile scan and result placement (gfx950).
//
// A batch is P pairs.  Pair p is A[a_beg[p] .. a_end[p]) against B[b_beg[p] .. b_end[p]) in two
// element pools.  CSR batches (consecutive pairs adjacent in both pools) pass a_end = a_beg + 1;
// "contig" marks them and lets a tile stage its whole A range and B range as two contiguous runs.
// Gather batches (pairs anywhere in the pools: graph edges, corpus lists, intermediate results)
// stage each pair's window separately.
//
// Work model.  Pair p owns units [uoff[p], uoff[p+1]) of one global unit space:
//   merge policy:  |A| + |B| + c0 units (one per merge-path step, plus padding),
//   search policy: min(|A|, |B|) + c0 units (one per probe of the smaller side, plus padding).
// The c0 padding bounds how many pairs one tile can touch (<= U/c0 + 2), which sizes the LDS
// segment table.  Tile t covers units [t*U, (t+1)*U): tiny pairs are packed many to a tile, large
// pairs are split across tiles, every tile carries the same work.
//
// Pipeline (count -> scan -> write), no communication between workgroups:
//   plan     unit prefix (closed form for CSR, a three-kernel scan for gather), tile -> pair maps
//   split    where each tile starts inside its first pair (merge-path split or probe index)
//   tile     one workgroup per tile: count and stage the tile's results (merge_tile / search_tile)
//   scan     scan of the tile counts (one kernel when the tiles are few)
//   place    move staged results to their final positions, per-pair output offsets
#pragma once

#include "tile_common.hpp"

namespace setops {

constexpr uint32_t kScanGroup = 1024;  // elements per group in the three-kernel scans

struct PairBatch {
    const uint32_t *A;   // A keys (values, or BSR bases)
    const uint32_t *As;  // A BSR states (null for plain sets)
    const uint32_t *B;
    const uint32_t *Bs;
    const uint64_t *a_beg;  // [P] first A element of pair p
    const uint64_t *a_end;  // [P] one past the last A element of pair p
    const uint64_t *b_beg;
    const uint64_t *b_end;
    uint64_t a_pool;  // elements readable at A (loads never go past)
    uint64_t b_pool;
    uint64_t P;
    uint32_t contig;  // 1: a_end[p] == a_beg[p+1] and b_end[p] == b_beg[p+1] for every p (CSR)
    // 0x80000000 for sorted int32 sets (plain sets only), else 0.  Every value comparison is made
    // on values XOR sflip, where unsigned order is signed order; values are written out unchanged.
    uint32_t sflip;
    const uint32_t *b_idx;  // optional: b_idx[k] = B[32k] (search policy: resident B pools)
    const uint32_t *b_top;  // optional, with b_idx: b_top[j] = B[kTopStride j] (staged in LDS per tile)
    const uint32_t *b_mid;  // with b_top: b_mid[m] = B[kMidStride m]
    // optional, one pair, search policy, plain sets: nx further sets (device pointers and
    // lengths); a probe's result needs it in every one of them as well
    const uint32_t *const *xsets;
    const uint64_t *xlens;
    uint32_t nx;
    // optional, any number of pairs, search policy, plain sets: pair p's further sets are
    // B[x_beg[j] .. x_end[j]) for j in [x_off[p], x_off[p+1]) (k-set groups: a probe found in the
    // pair's larger set is a result only if every further set holds it too)
    const uint64_t *x_off;
    const uint64_t *x_beg;
    const uint64_t *x_end;
};

constexpr int kMaxExtraSets = 8;  // further sets per launch (k-set queries run in groups of 2 + 8)
constexpr uint64_t kTopIdx = 256;              // index entries per top-level entry
constexpr uint64_t kTopStride = 32 * kTopIdx;  // pool elements per top-level entry (8,192)
constexpr int kTopCap = 2048;                  // top-level entries a search tile stages (8 KB of LDS)
constexpr uint64_t kMidIdx = 16;               // index entries per middle-level entry
constexpr uint64_t kMidStride = 32 * kMidIdx;  // pool elements per middle-level entry (512)
static_assert(kTopIdx / kMidIdx <= 16 && kMidIdx <= 16, "a level spans at most 16 entries of the next");

// Index of a resident pool in the compare space: idx[k] = pool[stride * k] ^ flip, k < ceil(n / stride).
__global__ void build_index_kernel(const uint32_t *__restrict__ pool, uint64_t n, uint32_t *__restrict__ idx,
                                   uint64_t stride, uint32_t flip)
{
    const uint64_t m = (n + stride - 1) / stride;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
        idx[k] = pool[stride * k] ^ flip;
}

enum : uint32_t { POLICY_MERGE = 0, POLICY_SEARCH = 1 };

struct TilePlan {
    uint64_t *uoff;        // [P+1] unit prefix
    uint32_t *tile_first;  // [ntiles] pair containing the tile's first unit
    uint32_t *tile_last;   // [ntiles] pair containing the tile's last unit
    uint32_t *tile_sa;     // [ntiles] pair-relative A index where the tile starts (probe index: search)
    uint32_t *tile_sb;     // [ntiles] pair-relative B index where the tile starts (merge only)
    uint64_t *tile_ga;     // [ntiles+1] pool index of tile t's first A element (merge, CSR batches)
    uint64_t *tile_gb;     // [ntiles+1] pool index of tile t's first B element (merge, CSR batches)
    uint32_t *tile_count;  // [ntiles] result words of tile t
    uint32_t *tile_elems;  // [ntiles] result elements of tile t (popcount for BSR)
    uint64_t *tile_off;    // [ntiles+1] exclusive scan of tile_count
    unsigned long long *gsum;   // [ngroups] group sums (pair units, then tile counts)
    unsigned long long *gsum2;  // [ngroups] second group sums (tile elements)
    uint32_t *stage;       // [ntiles * stage_words] per-tile staged outputs (keys, then BSR states)
    uint32_t *pair_local;  // [P] offset of a pair's first result inside the tile where it starts
    uint64_t utotal;
    uint32_t ntiles;
    uint32_t c0;           // padding units per pair
    uint32_t U;            // units per tile
    uint32_t stage_words;  // words per tile in `stage` (keys; states follow at +stage_words/2)
    uint32_t policy;       // POLICY_MERGE or POLICY_SEARCH (how units are counted)
    // single-pair mode: the plan kernel writes {0, na, 0, nb} to single_off, which serves as
    // a_beg / a_end / b_beg / b_end
    uint32_t single;
    uint64_t *single_off;
    uint64_t single_na, single_nb;
    // single pair in one tile, plain sets, output capacity >= min(na, nb): `stage` is the caller's
    // output, the tile kernel writes the totals, and the split, scan and placement are skipped
    uint32_t direct;
    // the one tile's outputs are final (a direct plan, or a plain search-policy batch of one tile
    // with output room for every unit): `stage` is the caller's output and the search tile writes
    // the per-pair offsets and the totals; the scan and placement are skipped
    uint32_t one_tile;
};

struct TileOut {
    uint32_t *out;          // compacted keys (null in count-only mode)
    uint32_t *out_s;        // compacted BSR states
    uint64_t cap;           // capacity of out / out_s in elements
    uint64_t *out_off;      // [P+1] output offset of each pair (may be null)
    uint32_t *counts;       // [P] per-pair output count (may be null)
    unsigned long long *totals;  // [2]: output words, output elements (popcount for BSR)
    uint32_t *error;        // bit 0: output capacity exceeded, bit 1: tile bound violated,
                            // bit 2: the device pair sizes disagree with the plan's unit total
};

__device__ __forceinline__ uint64_t pair_units(const PairBatch &in, uint32_t policy, uint32_t c0, uint64_t p)
{
    const uint64_t na = in.a_end[p] - in.a_beg[p], nb = in.b_end[p] - in.b_beg[p];
    return (policy == POLICY_SEARCH ? (na < nb ? na : nb) : na + nb) + c0;
}

// Records, for the pair owning units [u, un), the tiles it starts and ends.
// Search policy: a tile starting inside the pair starts at probe min(t*U - u, probes) of it (the
// split needs no search), recorded here too, so the split kernel does not run for such batches.
__device__ __forceinline__ void map_tiles(const TilePlan &plan, uint64_t p, uint64_t u, uint64_t un)
{
    const uint64_t U = plan.U;
    const uint64_t own = un - u - plan.c0;  // the pair's probes (search policy)
    for (uint64_t t = (u + U - 1) / U; t * U < un && t < plan.ntiles; t++) {
        plan.tile_first[t] = (uint32_t)p;
        if (plan.policy == POLICY_SEARCH) {
            const uint64_t rel = t * U - u;
            plan.tile_sa[t] = (uint32_t)(rel < own ? rel : own);
            plan.tile_sb[t] = 0;
        }
    }
    for (uint64_t t = u / U; t < plan.ntiles; t++) {
        uint64_t last = (t + 1) * U;
        if (last > plan.utotal) last = plan.utotal;
        last -= 1;
        if (last >= un) break;
        if (last >= u) plan.tile_last[t] = (uint32_t)p;
    }
}

__device__ __forceinline__ void plan_reset(const TileOut &out)
{
    if (out.totals) {
        out.totals[0] = 0;
        out.totals[1] = 0;
    }
    if (out.error) *out.error = 0;
}

// CSR batches: unit prefix in closed form, tile maps, zeroed per-pair counts.  Grid-stride over
// pairs.  In single-pair mode it also writes the pair's offsets.
__global__ void plan_csr_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t gsz = (uint64_t)gridDim.x * blockDim.x;
    auto abeg = [&](uint64_t p) -> uint64_t { return plan.single ? 0 : in.a_beg[p]; };
    auto bbeg = [&](uint64_t p) -> uint64_t { return plan.single ? 0 : in.b_beg[p]; };
    auto aend = [&](uint64_t p) -> uint64_t { return plan.single ? plan.single_na : in.a_end[p]; };
    auto bend = [&](uint64_t p) -> uint64_t { return plan.single ? plan.single_nb : in.b_end[p]; };
    if (plan.single && gid == 0) {
        plan.single_off[0] = 0;
        plan.single_off[1] = plan.single_na;
        plan.single_off[2] = 0;
        plan.single_off[3] = plan.single_nb;
        if (plan.direct) {
            // the one tile starts at the pair's start (the split of unit 0); its runs are the sets
            plan.tile_sa[0] = 0;
            plan.tile_sb[0] = 0;
            plan.tile_ga[0] = 0;
            plan.tile_ga[1] = plan.single_na;
            plan.tile_gb[0] = 0;
            plan.tile_gb[1] = plan.single_nb;
        }
    }
    const uint64_t a0 = abeg(0), b0 = bbeg(0);
    for (uint64_t p = gid; p < in.P; p += gsz) {
        const uint64_t ab = abeg(p) - a0, bb = bbeg(p) - b0;
        const uint64_t ae = aend(p) - a0, be = bend(p) - b0;
        uint64_t u, un;
        if (plan.policy == POLICY_SEARCH) {
            // min() does not telescope: only single-pair plans take this branch
            u = 0;
            un = ((ae - ab) < (be - bb) ? (ae - ab) : (be - bb)) + plan.c0;
        } else {
            u = ab + bb + (uint64_t)plan.c0 * p;
            un = ae + be + (uint64_t)plan.c0 * (p + 1);
        }
        plan.uoff[p] = u;
        if (p + 1 == in.P) plan.uoff[in.P] = un;
        map_tiles(plan, p, u, un);
        if (out.counts) out.counts[p] = 0;
    }
    if (gid == 0) plan_reset(out);
}

// ---- three-kernel scan of pair units (gather batches and the search policy) -------------------

// Group sums: one workgroup per kScanGroup pairs.
template <int NT>
__global__ __launch_bounds__(NT) void units_reduce_kernel(PairBatch in, TilePlan plan, uint32_t small_below)
{
    static_assert(NT == (int)kScanGroup, "one thread per element of a group");
    __shared__ unsigned long long red[NT / kWave][2];
    const uint64_t p = (uint64_t)blockIdx.x * kScanGroup + threadIdx.x;
    unsigned long long v = (p < in.P) ? pair_units(in, plan.policy, plan.c0, p) : 0ull;
    // with small_below: also count the pairs owning fewer than small_below units of their own
    unsigned long long sm = (small_below && p < in.P && v - plan.c0 < small_below) ? 1ull : 0ull;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        v += __shfl_xor(v, d, kWave);
        sm += __shfl_xor(sm, d, kWave);
    }
    if (lane_id() == 0) {
        red[threadIdx.x / kWave][0] = v;
        red[threadIdx.x / kWave][1] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0, s2 = 0;
        for (int w = 0; w < NT / kWave; w++) {
            s += red[w][0];
            s2 += red[w][1];
        }
        plan.gsum[blockIdx.x] = s;
        if (small_below) plan.gsum2[blockIdx.x] = s2;
    }
}

// Exclusive scan of n group sums in place by one workgroup (and of gsum2 when `two`); the
// totals go to tot[0] (and tot[1]).
template <int NT>
__global__ __launch_bounds__(NT) void groups_scan_kernel(unsigned long long *gsum, unsigned long long *gsum2,
                                                         uint64_t n, int two, unsigned long long *tot)
{
    __shared__ unsigned long long ws[NT / kWave][2];
    unsigned long long carry = 0, carry2 = 0;
    const uint32_t w = threadIdx.x / kWave, l = lane_id();
    for (uint64_t base = 0; base < n; base += NT) {
        const uint64_t i = base + threadIdx.x;
        const unsigned long long v = (i < n) ? gsum[i] : 0ull;
        const unsigned long long v2 = (two && i < n) ? gsum2[i] : 0ull;
        unsigned long long x = v, x2 = v2;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const unsigned long long y = __shfl_up(x, d, kWave), y2 = __shfl_up(x2, d, kWave);
            if (l >= (uint32_t)d) {
                x += y;
                x2 += y2;
            }
        }
        if (l == kWave - 1) {
            ws[w][0] = x;
            ws[w][1] = x2;
        }
        __syncthreads();
        unsigned long long pre = 0, pre2 = 0, all = 0, all2 = 0;
        for (int k = 0; k < NT / kWave; k++) {
            if (k < (int)w) {
                pre += ws[k][0];
                pre2 += ws[k][1];
            }
            all += ws[k][0];
            all2 += ws[k][1];
        }
        if (i < n) {
            gsum[i] = carry + pre + x - v;
            if (two) gsum2[i] = carry2 + pre2 + x2 - v2;
        }
        carry += all;
        carry2 += all2;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot[0] = carry;
        if (two) tot[1] = carry2;
    }
}

// Per-group scan of pair units plus the group base: uoff, tile maps, zeroed counts.  With
// `one_group` (a batch of at most NT * MAXCHUNK pairs) one workgroup scans the whole batch, the
// base is 0 and the group-sum kernels do not run.
template <int NT, int MAXCHUNK>
__global__ __launch_bounds__(NT) void units_scan_kernel(PairBatch in, TilePlan plan, TileOut out, int one_group)
{
    static_assert(NT == (int)kScanGroup, "one thread per pair of a group");
    __shared__ unsigned long long wsum[NT / kWave];
    // one_group: a single workgroup scans all pairs, thread t the C = ceil(P / NT) <= MAXCHUNK
    // consecutive pairs from t*C (their sizes load together, one block scan of the thread sums);
    // otherwise group blockIdx.x, one pair per thread, after the group sums' scan
    const uint32_t C = one_group ? (uint32_t)((in.P + NT - 1) / NT) : 1u;
    const uint64_t p0 = one_group ? (uint64_t)threadIdx.x * C : (uint64_t)blockIdx.x * kScanGroup + threadIdx.x;
    unsigned long long vs[MAXCHUNK], local = 0;
#pragma unroll
    for (int c = 0; c < MAXCHUNK; c++) {
        const uint64_t p = p0 + (uint64_t)c;
        vs[c] = ((uint32_t)c < C && p < in.P) ? pair_units(in, plan.policy, plan.c0, p) : 0ull;
        local += vs[c];
    }
    // 64-bit block scan of the thread sums: wave scan, then wave totals
    unsigned long long x = local;
    const uint32_t l = lane_id(), w = threadIdx.x / kWave;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned long long y = __shfl_up(x, d, kWave);
        if (l >= (uint32_t)d) x += y;
    }
    if (l == kWave - 1) wsum[w] = x;
    __syncthreads();
    unsigned long long u = one_group ? 0ull : plan.gsum[blockIdx.x];
    for (int k = 0; k < (int)w; k++) u += wsum[k];
    u += x - local;  // the thread's first pair starts here
#pragma unroll
    for (int c = 0; c < MAXCHUNK; c++) {
        const uint64_t p = p0 + (uint64_t)c;
        if ((uint32_t)c < C && p < in.P) {
            const uint64_t un = u + vs[c];
            plan.uoff[p] = u;
            if (p + 1 == in.P) plan.uoff[in.P] = un;
            map_tiles(plan, p, u, un);
            if (out.counts) out.counts[p] = 0;
            u = un;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) plan_reset(out);
}

// One pair over several tiles (single mode, merge policy): the plan and the split in one launch.
// Pair 0 owns units [0, na + nb + c0) -- every tile lies in it -- so thread t writes tile t's maps
// and finds its merge-path split of unit t*U directly (the search of plan_split_kernel).
__global__ void plan_split_single_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t na = plan.single_na, nb = plan.single_nb;
    if (t == 0) {
        plan.single_off[0] = 0;
        plan.single_off[1] = na;
        plan.single_off[2] = 0;
        plan.single_off[3] = nb;
        plan.uoff[0] = 0;
        plan.uoff[1] = na + nb + plan.c0;
        plan.tile_ga[plan.ntiles] = na;
        plan.tile_gb[plan.ntiles] = nb;
        if (out.counts) out.counts[0] = 0;
        plan_reset(out);
    }
    if (t >= plan.ntiles) return;
    plan.tile_first[t] = 0;
    plan.tile_last[t] = 0;
    const uint64_t rel = t * plan.U;
    const uint64_t d = rel < na + nb ? rel : na + nb;
    uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    const uint32_t *__restrict__ A = in.A;
    const uint32_t *__restrict__ B = in.B;
    const uint32_t f = in.sflip;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((A[mid] ^ f) <= (B[d - 1 - mid] ^ f)) lo = mid + 1; else hi = mid;
    }
    plan.tile_sa[t] = (uint32_t)lo;
    plan.tile_sb[t] = (uint32_t)(d - lo);
    plan.tile_ga[t] = lo;
    plan.tile_gb[t] = d - lo;
}

// Where each tile starts inside its first pair: the merge-path split of unit t*U (merge policy:
// A index i and B index d - i with every A[x < i] <= B[d-1-...]), or the probe index (search).
// One thread per tile (a cooperative search costs more loads in total than it saves latency).
__global__ void plan_split_kernel(PairBatch in, TilePlan plan)
{
    const uint64_t gsz = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < plan.ntiles; t += gsz) {
        const uint64_t p = plan.tile_first[t];
        if (p >= in.P) {  // a tile no pair reached (device sizes differ from the plan): flagged by the tile kernel
            plan.tile_sa[t] = 0;
            plan.tile_sb[t] = 0;
            continue;
        }
        const uint64_t a = in.a_beg[p], b = in.b_beg[p];
        const uint64_t na = in.a_end[p] - a, nb = in.b_end[p] - b;
        const uint64_t rel = (uint64_t)t * plan.U - plan.uoff[p];
        if (plan.policy == POLICY_SEARCH) {
            const uint64_t np = na < nb ? na : nb;
            plan.tile_sa[t] = (uint32_t)(rel < np ? rel : np);
            plan.tile_sb[t] = 0;
            continue;
        }
        const uint64_t d = rel < na + nb ? rel : na + nb;
        uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
        const uint32_t *__restrict__ A = in.A + a;
        const uint32_t *__restrict__ B = in.B + b;
        const uint32_t f = in.sflip;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if ((A[mid] ^ f) <= (B[d - 1 - mid] ^ f)) lo = mid + 1; else hi = mid;
        }
        plan.tile_sa[t] = (uint32_t)lo;
        plan.tile_sb[t] = (uint32_t)(d - lo);
        if (in.contig) {
            // CSR: a tile's A run is [tile_ga[t], tile_ga[t+1]) -- one load for the staging loop
            plan.tile_ga[t] = a + lo;
            plan.tile_gb[t] = b + (d - lo);
            if (t + 1 == plan.ntiles) {
                plan.tile_ga[t + 1] = in.a_end[in.P - 1];
                plan.tile_gb[t + 1] = in.b_end[in.P - 1];
            }
        }
    }
}

// ---- tile scan ---------------------------------------------------------------------------------

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
        plan.gsum[blockIdx.x] = sc;
        plan.gsum2[blockIdx.x] = se;
    }
}

// Per-group block scan of the tile counts plus the (already scanned) group base.  The last
// group writes tile_off[ntiles] and flags a clipped output.  Totals were written by the group
// scan.
// With `direct` (few groups) there were no group-sum kernels: a workgroup adds the tile counts
// of the earlier groups itself (both counts for the last group, which writes the totals).
template <int NT>
__global__ __launch_bounds__(NT) void scan_tiles_kernel(TilePlan plan, TileOut out, int direct)
{
    static_assert(NT == (int)kScanGroup, "one thread per tile of a group");
    __shared__ uint32_t scratch[NT / kWave];
    __shared__ unsigned long long red[NT / kWave][2];
    const uint32_t g = blockIdx.x;
    unsigned long long base;
    if (direct) {
        const uint32_t ngroups = (plan.ntiles + kScanGroup - 1) / kScanGroup;
        const bool last = g + 1 == ngroups;
        const uint32_t end = last ? plan.ntiles : g * kScanGroup;  // the last group sums all tiles
        // four words per load, LD loads in flight per thread
        constexpr int LD = 4;
        unsigned long long b = 0, e = 0;
        for (uint32_t i0 = 4 * threadIdx.x; i0 < end; i0 += 4 * LD * NT) {
            uint4 c[LD], el[LD];
#pragma unroll
            for (int k = 0; k < LD; k++) {
                const uint32_t i = i0 + 4 * NT * k;
                if (i < end) {
                    c[k] = *reinterpret_cast<const uint4 *>(plan.tile_count + i);
                    if (last) el[k] = *reinterpret_cast<const uint4 *>(plan.tile_elems + i);
                }
            }
#pragma unroll
            for (int k = 0; k < LD; k++) {
                const uint32_t i = i0 + 4 * NT * k;
                if (i < end) {
                    const uint32_t cw[4] = {c[k].x, c[k].y, c[k].z, c[k].w};
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (i + q < g * kScanGroup) b += cw[q];
                        if (last && i + q < end) {
                            const uint32_t ew[4] = {el[k].x, el[k].y, el[k].z, el[k].w};
                            e += ew[q];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            b += __shfl_xor(b, d, kWave);
            e += __shfl_xor(e, d, kWave);
        }
        if (lane_id() == 0) {
            red[threadIdx.x / kWave][0] = b;
            red[threadIdx.x / kWave][1] = e;
        }
        __syncthreads();
        base = 0;
        unsigned long long el = 0;
        for (int k = 0; k < NT / kWave; k++) {
            base += red[k][0];
            el += red[k][1];
        }
        if (last && threadIdx.x == 0 && out.totals) out.totals[1] = el;  // totals[0] below
    } else {
        base = plan.gsum[g];
    }
    const uint32_t i = g * kScanGroup + threadIdx.x;
    const uint32_t c = (i < plan.ntiles) ? plan.tile_count[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<NT>(c, scratch, &tot);
    if (i < plan.ntiles) plan.tile_off[i] = base + ex;
    if (i + 1 == plan.ntiles) {
        const unsigned long long total = base + ex + c;
        plan.tile_off[plan.ntiles] = total;
        if (direct && out.totals) out.totals[0] = total;
        if (out.out && total > out.cap && out.error) atomicOr(out.error, 1u);
    }
}

// Count-only runs: the per-pair output offsets alone, one thread per pair.  Pair p's results
// start in the tile holding its first unit, at the in-tile offset that tile recorded.
__global__ void pair_offsets_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    const uint64_t gsz = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < in.P; p += gsz) {
        uint64_t t = plan.uoff[p] / plan.U;
        if (t >= plan.ntiles) t = plan.ntiles - 1;  // stale plans (flagged by the tile kernel): stay inside
        out.out_off[p] = plan.tile_off[t] + plan.pair_local[p];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out.out_off[in.P] = plan.tile_off[plan.ntiles];
}

// Write: moves the staged results to their final positions and writes the per-pair output
// offsets.  One workgroup per tile (tiles tile0 + blockIdx.x): the tile's results are the output
// run [tile_off[t], tile_off[t+1]); its lanes copy words x = lane + NT*j of the run from the tile's
// staging slot, with all of a round's loads in flight before its stores, and write the offsets of
// the pairs whose first unit lies in the tile (the tile recorded each such pair's in-tile result
// offset).  A workgroup depends on nothing but its own tile's offsets, so the many small
// workgroups balance tiles of very different result counts; the tile's offsets and pair range
// are read before anything is written (uniform loads through the scalar cache), and the first
// pair offsets load before the copy so their round trip overlaps it.
// FUSED (at most kFuseScanTiles tiles): no tile scan ran before: each workgroup sums the tile
// counts before its tile itself and writes its tile's offset, and the workgroup of the last tile
// writes the totals, the capacity flag and the end offsets.
constexpr uint32_t kFuseScanTiles = 1024;

template <int NT, int Q, bool BSR, bool FUSED>
__global__ __launch_bounds__(NT) void place_tiles_kernel(PairBatch in, TilePlan plan, TileOut out, uint32_t tile0)
{
    __shared__ unsigned long long red[FUSED ? NT / kWave : 1][2];
    const uint32_t tid = threadIdx.x;
    const uint64_t t = (uint64_t)tile0 + blockIdx.x;
    if (t >= plan.ntiles) return;
    const uint32_t sw = plan.stage_words;
    const uint64_t pf = plan.tile_first[t];
    uint64_t pl = plan.tile_last[t];
    if (pl >= in.P) pl = in.P - 1;  // stale maps (flagged by the tile kernel): stay inside the pairs
    uint64_t base;
    uint32_t n;
    if constexpr (FUSED) {
        // results before the tile (and, for the last tile, the elements of every tile)
        const bool last = t + 1 == plan.ntiles;
        unsigned long long b = 0, el = 0;
        // four counts per load, every load in flight together (tile arrays are padded to 256 bytes)
        constexpr int QF = (int)kFuseScanTiles / (4 * NT);
        static_assert(QF * 4 * NT == (int)kFuseScanTiles, "whole 16-byte loads cover the fused tile limit");
        uint4 cv[QF], ev[QF];
#pragma unroll
        for (int q = 0; q < QF; q++) {
            const uint32_t i = 4u * (tid + (uint32_t)q * NT);
            if (i < plan.ntiles) {
                cv[q] = *reinterpret_cast<const uint4 *>(plan.tile_count + i);
                if (last) ev[q] = *reinterpret_cast<const uint4 *>(plan.tile_elems + i);
            }
        }
#pragma unroll
        for (int q = 0; q < QF; q++) {
            const uint32_t i = 4u * (tid + (uint32_t)q * NT);
            if (i < plan.ntiles) {
                const uint32_t cw[4] = {cv[q].x, cv[q].y, cv[q].z, cv[q].w};
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) {
                    if (i + j < t) b += cw[j];
                    if (last && i + j < plan.ntiles) {
                        const uint32_t ew[4] = {ev[q].x, ev[q].y, ev[q].z, ev[q].w};
                        el += ew[j];
                    }
                }
            }
        }
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            b += __shfl_xor(b, d, kWave);
            el += __shfl_xor(el, d, kWave);
        }
        if (lane_id() == 0) {
            red[tid / kWave][0] = b;
            red[tid / kWave][1] = el;
        }
        __syncthreads();
        base = 0;
        el = 0;
        for (int k = 0; k < NT / kWave; k++) {
            base += red[k][0];
            el += red[k][1];
        }
        n = plan.tile_count[t];
        if (tid == 0) {
            plan.tile_off[t] = base;
            if (last) {
                const unsigned long long total = base + n;
                plan.tile_off[plan.ntiles] = total;
                if (out.totals) {
                    out.totals[0] = total;
                    out.totals[1] = el;
                }
                if (out.out && total > out.cap && out.error) atomicOr(out.error, 1u);
                if (out.out_off) out.out_off[in.P] = total;
            }
        }
    } else {
        base = plan.tile_off[t];
        n = (uint32_t)(plan.tile_off[t + 1] - base);
        if (out.out_off && t == 0 && tid == 0) out.out_off[in.P] = plan.tile_off[plan.ntiles];
    }
    // the first pair offsets (a pair whose first unit lies in an earlier tile got its offset there)
    const uint64_t u0 = t * plan.U;
    const uint64_t p0 = pf + tid;
    uint64_t up = 0;
    uint32_t ploc = 0;
    if (out.out_off && p0 <= pl) {
        up = plan.uoff[p0];
        ploc = plan.pair_local[p0];
    }
    if (out.out) {
        const uint64_t room = out.cap > base ? out.cap - base : 0;
        const uint32_t lim = (uint64_t)n < room ? n : (uint32_t)room;
        const uint32_t *src = plan.stage + t * sw;
        uint32_t *dk = out.out + base;
        uint32_t *ds = BSR ? out.out_s + base : nullptr;
        for (uint32_t x0 = 0; x0 < lim; x0 += (uint32_t)NT * Q) {
            uint32_t v[Q], s2[BSR ? Q : 1];
#pragma unroll
            for (int j = 0; j < Q; j++) {
                const uint32_t x = x0 + tid + (uint32_t)j * NT;
                if (x < lim) {
                    v[j] = src[x];
                    if constexpr (BSR) s2[j] = src[sw / 2 + x];
                }
            }
#pragma unroll
            for (int j = 0; j < Q; j++) {
                const uint32_t x = x0 + tid + (uint32_t)j * NT;
                if (x < lim) {
                    dk[x] = v[j];
                    if constexpr (BSR) ds[x] = s2[j];
                }
            }
        }
    }
    if (out.out_off) {
        if (p0 <= pl && up >= u0) out.out_off[p0] = base + ploc;
        for (uint64_t p = p0 + NT; p <= pl; p += NT) {  // tiles of more than NT pairs
            const uint64_t uq = plan.uoff[p];
            const uint32_t lq = plan.pair_local[p];
            if (uq >= u0) out.out_off[p] = base + lq;
        }
    }
}

// ---- per-tile segment geometry -------------------------------------------------------------------

// One covered pair of a tile: its element windows (pair-relative) and where the tile's units
// reach it.  Computed by thread s for segment s.
struct SegWin {
    uint64_t p;
    uint64_t a, b;     // pool index of the pair's first A / B element
    uint64_t na, nb;   // pair sizes
    uint64_t i0, i1;   // A window [i0, i1)
    uint64_t j0, j1;   // B window [j0, j1) (merge), unused for search
    uint32_t u;        // tile-relative unit where the segment starts
    uint32_t flags;    // bit0: pair starts in this tile, bit1: pair ends in this tile
};

// Where a tile starts and where the next one starts (both inside their first pairs), as the
// segment table needs them.
struct TileEdges {
    uint32_t sa, sb;         // tile_sa[t], tile_sb[t]
    uint32_t next_first;     // tile_first[t+1] (~0 for the last tile)
    uint32_t next_sa, next_sb;
};

__device__ __forceinline__ TileEdges tile_edges(const TilePlan &plan, uint32_t t)
{
    TileEdges e;
    e.sa = plan.tile_sa[t];
    e.sb = plan.tile_sb[t];
    const bool more = t + 1 < plan.ntiles;
    e.next_first = more ? plan.tile_first[t + 1] : ~0u;
    e.next_sa = more ? plan.tile_sa[t + 1] : 0u;
    e.next_sb = more ? plan.tile_sb[t + 1] : 0u;
    return e;
}

// Window of segment s of tile t (pairs p_first .. p_last); the last segment ends where the next
// tile starts if that tile starts in the same pair, else at the pair's end.
__device__ __forceinline__ SegWin seg_window(const PairBatch &in, const TilePlan &plan, uint32_t t, uint32_t p_first,
                                             uint32_t nseg, uint32_t s, uint32_t ulen, const TileEdges &e)
{
    SegWin w;
    const uint64_t p = (uint64_t)p_first + s;
    w.p = p;
    w.a = in.a_beg[p];
    w.b = in.b_beg[p];
    w.na = in.a_end[p] - w.a;
    w.nb = in.b_end[p] - w.b;
    const uint64_t up = plan.uoff[p], upn = plan.uoff[p + 1];
    const uint64_t u0 = (uint64_t)t * plan.U;
    w.i0 = (s == 0) ? e.sa : 0;
    w.j0 = (s == 0) ? e.sb : 0;
    const bool cont = (s == nseg - 1) && e.next_first == (uint32_t)p;
    if (plan.policy == POLICY_SEARCH) {
        const uint64_t np = w.na < w.nb ? w.na : w.nb;
        w.i1 = cont ? e.next_sa : np;
        w.j0 = 0;
        w.j1 = 0;
    } else {
        w.i1 = cont ? e.next_sa : w.na;
        w.j1 = cont ? e.next_sb : w.nb;
    }
    w.u = (up > u0) ? (uint32_t)(up - u0) : 0u;
    w.flags = (up >= u0 ? 1u : 0u) | (upn <= u0 + ulen ? 2u : 0u);
    return w;
}

__device__ __forceinline__ SegWin seg_window(const PairBatch &in, const TilePlan &plan, uint32_t t, uint32_t p_first,
                                             uint32_t nseg, uint32_t s, uint32_t ulen)
{
    return seg_window(in, plan, t, p_first, nseg, s, ulen, tile_edges(plan, t));
}

// The one segment of a direct plan (one pair in one tile): the whole pair, starting and ending
// in the tile at unit 0 (what seg_window gives for it under the tile kind's policy).
template <bool SEARCH>
__device__ __forceinline__ SegWin direct_window(const TilePlan &plan)
{
    SegWin w;
    const uint64_t na = plan.single_na, nb = plan.single_nb;
    w.p = 0;
    w.a = 0;
    w.b = 0;
    w.na = na;
    w.nb = nb;
    w.i0 = 0;
    w.j0 = 0;
    if constexpr (SEARCH) {
        w.i1 = na < nb ? na : nb;
        w.j1 = 0;
    } else {
        w.i1 = na;
        w.j1 = nb;
    }
    w.u = 0;
    w.flags = 3u;
    return w;
}

// Pair p of a direct search batch (nothing planned) as a segment of the tile of units
// [u0, u0 + ulen): the pair owns units [pu, pu + units), its probes first (from the smaller side),
// and meets the tile.  What seg_window gives for it from a plan.
__device__ __forceinline__ SegWin search_direct_window(const PairBatch &in, uint32_t p, uint64_t pu, uint64_t units,
                                                       uint64_t u0, uint32_t ulen)
{
    SegWin w;
    w.p = p;
    w.a = in.a_beg[p];
    w.b = in.b_beg[p];
    w.na = in.a_end[p] - w.a;
    w.nb = in.b_end[p] - w.b;
    const uint64_t np = w.na < w.nb ? w.na : w.nb;
    const uint64_t t1 = u0 + ulen;
    w.i0 = pu < u0 ? ((u0 - pu) < np ? (u0 - pu) : np) : 0;
    uint64_t i1 = t1 > pu ? t1 - pu : 0;
    if (i1 > np) i1 = np;
    w.i1 = i1 < w.i0 ? w.i0 : i1;
    w.j0 = 0;
    w.j1 = 0;
    w.u = pu > u0 ? (uint32_t)(pu - u0) : 0u;
    w.flags = (pu >= u0 ? 1u : 0u) | (pu + units <= t1 ? 2u : 0u);
    return w;
}

// One 16-byte chunk of a pool at element index e (16-byte aligned), guarding the pool end.
__device__ __forceinline__ uint4 load_chunk(const uint32_t *__restrict__ src, uint64_t pool, uint64_t e)
{
    uint4 v;
    if (e + 4 <= pool) {
        v = *reinterpret_cast<const uint4 *>(src + e);
    } else {
        v.x = (e + 0 < pool) ? src[e + 0] : 0u;
        v.y = (e + 1 < pool) ? src[e + 1] : 0u;
        v.z = (e + 2 < pool) ? src[e + 2] : 0u;
        v.w = (e + 3 < pool) ? src[e + 3] : 0u;
    }
    return v;
}

// Stages chunks [0, n) of src + base (base a multiple of 4 elements) into dst[4c .. 4c+4) with
// direct global->LDS loads (16 bytes per lane; a wave's row of 64 chunks lands contiguously at
// the row's LDS address).  The chunk that reaches past the pool (at most the last one) is written
// element-wise, zero-filled.  The caller waits (s_waitcnt) and synchronizes before reading dst.
template <int NT>
__device__ __forceinline__ void stage_run_lds(const uint32_t *__restrict__ src, uint64_t pool, uint64_t base,
                                              uint32_t n, uint32_t *dst)
{
    const uint64_t fp = pool > base ? (pool - base) / 4 : 0;
    const uint32_t full = fp < n ? (uint32_t)fp : n;
    const uint4 *p = reinterpret_cast<const uint4 *>(src + base);
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint32_t r = threadIdx.x - lane; r < full; r += NT) {  // r: the wave's row (uniform)
        if (r + lane < full) __builtin_amdgcn_global_load_lds(p + r + lane, dst + 4 * r, 16, 0, 0);
    }
    if (full < n && threadIdx.x == 0) {
        const uint64_t e = base + 4ull * full;
        for (uint32_t j = 0; j < 4; j++) dst[4 * full + j] = (e + j < pool) ? src[e + j] : 0u;
    }
}

// Gather staging: nwin windows, window w = src_w[g0_w .. g0_w + n_w), each placed in LDS at
// dst + 4*cpfx[w] + (g0_w & 3) as whole aligned 16-byte chunks.  cpfx is the exclusive prefix of
// the windows' chunk counts (cpfx[nwin] = total).  Each thread takes chunks c, c + NT, ... (at most
// QMAX of them), finds each chunk's window by binary search over cpfx, and issues all its loads
// before storing.  Every staged word is XORed with `flip` (PairBatch::sflip for keys, 0 for states).
template <int NT, int QMAX>
__device__ __forceinline__ void stage_windows(const uint32_t *__restrict__ srcA, uint64_t poolA,
                                              const uint32_t *__restrict__ srcB, uint64_t poolB,
                                              const uint32_t *cpfx, const uint64_t *gbase, const uint32_t *wsrc,
                                              uint32_t nwin, uint32_t *dst, uint32_t flip)
{
    const uint32_t total = cpfx[nwin];
    uint4 v[QMAX];
#pragma unroll
    for (int k = 0; k < QMAX; k++) {
        const uint32_t c = threadIdx.x + (uint32_t)k * NT;
        if (c < total) {
            uint32_t lo = 0, hi = nwin;  // largest w with cpfx[w] <= c
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (cpfx[mid] <= c) lo = mid; else hi = mid;
            }
            const bool fromB = wsrc[lo] != 0;
            v[k] = load_chunk(fromB ? srcB : srcA, fromB ? poolB : poolA, gbase[lo] + 4ull * (c - cpfx[lo]));
        }
    }
#pragma unroll
    for (int k = 0; k < QMAX; k++) {
        const uint32_t c = threadIdx.x + (uint32_t)k * NT;
        if (c < total) {
            uint4 x = v[k];
            x.x ^= flip;
            x.y ^= flip;
            x.z ^= flip;
            x.w ^= flip;
            *reinterpret_cast<uint4 *>(dst + 4 * c) = x;
        }
    }
}

} // namespace setops
