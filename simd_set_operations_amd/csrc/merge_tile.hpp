// merge_tile.hpp -- batched sorted-set intersection over a global unit space (gfx950).
//
// Work model.  Pair p of a batch owns units [uoff[p], uoff[p+1]).  For the merge policy a pair
// has |A|+|B| real units (one per merge-path diagonal step) plus c0 padding units; the padding
// bounds how many pairs one tile can touch (<= U/c0 + 2), which sizes the LDS segment table.
// Tile t covers units [t*U, (t+1)*U) of the concatenated unit space, so tiny pairs are packed
// many to a tile (several per lane), medium pairs span lanes of one tile, and huge pairs are
// split across tiles by merge-path -- one kernel, no per-size launches, perfectly even tiles.
//
// Per tile:
//   1. claim the next tile from a device counter (persistent grid; tiles are claimed in order),
//   2. read the tile's precomputed boundaries (merge-path splits found by plan_split_kernel, one
//      thread per tile boundary) and the covered pairs' descriptors,
//   3. stage the tile's A range and B range (each contiguous in the CSR pools) in LDS with
//      16-byte loads; the B range carries one element past the tile for the boundary compare,
//   4. every lane finds its K-unit slice by an LDS merge-path search and walks K merge steps,
//      emitting on an A step whose value equals the current B value (exactly once per match,
//      independent of where tile and lane boundaries fall),
//   5. block scan of per-lane match counts, decoupled look-back for the tile's output position,
//   6. ordered compaction through LDS and coalesced stores; per-pair offsets and counts.
//
// The same kernel handles the BSR representation (key = base, payload = 32-bit state): a base
// match emits (base, state_a & state_b) when the AND is non-zero.
#pragma once

#include "tile_common.hpp"

namespace setops {

constexpr int kClaimQueues = 16;      // tile claim counters (tile t belongs to queue t % 16)
constexpr uint32_t kClaimStride = 64;  // words between counters (separate 256-byte lines)
constexpr uint32_t kClaimSlack = 8 * kClaimQueues;

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
    uint64_t *status;      // [ntiles] look-back words
    uint32_t *queue;       // kClaimQueues tile claim counters, kClaimStride words apart
    uint32_t *pend;        // per-workgroup pending areas (pend_words() each)
    uint32_t spin_limit;   // look-back sleeps before a tile's outputs go to the pending area
    uint64_t utotal;
    uint32_t ntiles;
    uint32_t c0;           // padding units per pair
    uint32_t U;            // units per tile
    uint32_t policy;       // 0 = merge units (|A|+|B|+c0)
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
    for (uint64_t t = gid; t < plan.ntiles; t += gsz) plan.status[t] = 0;
    if (gid < (uint64_t)kClaimQueues) plan.queue[gid * kClaimStride] = 0;
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
    uint64_t wave_w[NT / kWave], wave_e[NT / kWave];
    uint64_t ga[2], gb[2];       // tile start / end in the A and B pools
    uint64_t excl;
    uint32_t tile, p_first, p_last, ok;
    uint32_t pend_tile, pend_tot, pend_nseg;  // the tile waiting in the pending area
};

// Copy `n` elements starting at global element index g0 of `src` (pool length `pool`) into LDS
// at dst[shift ...], where shift = g0 & 3 and loads are 16-byte aligned chunks of the pool.
template <int NT>
__device__ __forceinline__ void stage_range(const uint32_t *__restrict__ src, uint64_t pool, uint64_t g0, uint32_t n,
                                            uint32_t *dst)
{
    const uint64_t base = g0 & ~3ull;
    const uint32_t span = (uint32_t)(g0 - base) + n;
    const uint32_t chunks = (span + 3) / 4;
    for (uint32_t c = threadIdx.x; c < chunks; c += NT) {
        const uint64_t e = base + 4ull * c;
        if (e + 4 <= pool) {
            const uint4 v = *reinterpret_cast<const uint4 *>(src + e);
            *reinterpret_cast<uint4 *>(dst + 4 * c) = v;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) dst[4 * c + q] = (e + q < pool) ? src[e + q] : 0u;
        }
    }
}

// Writes a pending tile's outputs (kept in the workgroup's pending area) to their final place
// once the tile's exclusive prefix is known.  All threads of the workgroup.
template <int NT, bool BSR>
__device__ __forceinline__ void flush_pending(uint32_t t, uint64_t excl, uint32_t tot, uint32_t nseg,
                                              const uint32_t *pend_out, const uint32_t *pend_st,
                                              const uint32_t *pend_rec, const PairBatch &in, const TilePlan &plan,
                                              const TileOut &out)
{
    uint64_t lim = tot;
    if (excl + tot > out.cap) {
        lim = (excl < out.cap) ? out.cap - excl : 0;
        if (threadIdx.x == 0 && out.error) atomicOr(out.error, 1u);
    }
    for (uint32_t x = threadIdx.x; x < lim; x += NT) {
        out.out[excl + x] = pend_out[x];
        if constexpr (BSR) out.out_s[excl + x] = pend_st[x];
    }
    if (out.out_off) {
        for (uint32_t s = threadIdx.x; s < nseg; s += NT) {
            const uint32_t r = pend_rec[2 * s + 1];
            if (r & 0x80000000u) out.out_off[pend_rec[2 * s]] = excl + (r & 0x7FFFFFFFu);
        }
        if (t == plan.ntiles - 1 && threadIdx.x == 0) out.out_off[in.P] = excl + tot;
    }
}

// Words of one workgroup's pending area (outputs of a tile whose output position is not known
// yet): keys, states (BSR), then one record per segment (pair index, local start | valid bit).
template <int U, int MAXSEG>
__host__ __device__ constexpr uint32_t pend_words()
{
    return 2u * (U / 2 + 64) + 2u * MAXSEG;
}

template <int NT, int K, int MAXSEG, bool BSR, bool COUNT_ONLY>
__global__ __launch_bounds__(NT) void merge_tile_kernel(PairBatch in, TilePlan plan, TileOut out)
{
    static_assert(K >= 1 && K <= 32, "lane masks hold K bits");
    constexpr int U = NT * K;
    constexpr uint32_t kPendOut = U / 2 + 64;  // >= max outputs of one tile
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    __shared__ MergeSmem<NT, K, MAXSEG, BSR> sm;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / kWave;
    // result words / elements of this thread over all its tiles; one atomic pair per workgroup
    uint64_t acc_w = 0, acc_e = 0;
    const uint32_t home = blockIdx.x % kClaimQueues;
    // this workgroup's pending area
    uint32_t *pend = plan.pend + (uint64_t)blockIdx.x * pend_words<U, MAXSEG>();
    uint32_t *pend_out = pend, *pend_st = pend + kPendOut, *pend_rec = pend + 2 * kPendOut;
    if (tid == 0) sm.pend_tile = kNone;
    // wave 0 reads the queue heads one tile ahead; the decision uses them at the next claim
    uint32_t heads = 0;
    if (wave == 0) heads = claim_heads<kClaimQueues>(plan.queue, kClaimStride);
    __syncthreads();

    for (;;) {
        // ---- 0. claim ------------------------------------------------------------------------
        // Normal mode: home queue (fast, spread over kClaimQueues counters).  With a tile pending,
        // only the lowest unclaimed tile is taken, so the blocking look-back below is safe.
        if (wave == 0) {
            uint32_t t;
            if (COUNT_ONLY || sm.pend_tile == kNone)
                t = claim_finish<kClaimQueues>(plan.queue, kClaimStride, plan.ntiles, home, kClaimSlack, heads);
            else
                t = claim_lowest<kClaimQueues>(plan.queue, kClaimStride, plan.ntiles);
            if (lane == 0) sm.tile = t;
        }
        __syncthreads();
        const uint32_t t = sm.tile;
        if (t >= plan.ntiles) break;
        const bool pending_mode = !COUNT_ONLY && sm.pend_tile != kNone;

        // ---- 1. tile descriptor (one round trip) ------------------------------------------
        if (tid < 2) {
            sm.ga[tid] = plan.tile_ga[t + tid];
            sm.gb[tid] = plan.tile_gb[t + tid];
        } else if (tid == 2) {
            sm.p_first = plan.tile_first[t];
        } else if (tid == 3) {
            sm.p_last = plan.tile_last[t];
        }
        __syncthreads();
        const uint64_t u0 = (uint64_t)t * U;
        const uint64_t u1 = (u0 + U < plan.utotal) ? u0 + U : plan.utotal;
        uint32_t ulen = (uint32_t)(u1 - u0);
        const uint32_t p_first = sm.p_first;
        uint32_t nseg = sm.p_last - p_first + 1;
        const uint64_t ga0 = sm.ga[0], ga1 = sm.ga[1], gb0 = sm.gb[0], gb1 = sm.gb[1];
        const bool bad = nseg == 0 || nseg > (uint32_t)MAXSEG || ga1 < ga0 || gb1 < gb0 ||
                         (ga1 - ga0) + (gb1 - gb0) > (uint64_t)U;
        uint32_t n_a = 0, n_b = 0;
        if (bad) {
            // A plan that violates the tile bounds: report it and run the tile as empty, so the
            // look-back chain stays intact and nothing is read or written.
            if (tid == 0 && out.error) atomicOr(out.error, 2u);
            nseg = 0;
            ulen = 0;
        } else {
            n_a = (uint32_t)(ga1 - ga0);
            n_b = (uint32_t)(gb1 - gb0) + ((gb1 < in.b_pool) ? 1u : 0u);  // + boundary element
        }
        const uint32_t shift_a = (uint32_t)(ga0 & 3), shift_b = (uint32_t)(gb0 & 3);
        const uint32_t boff = 4u * ((shift_a + n_a + 3) / 4);

        // ---- 2. stage the A and B ranges; segment table (one round trip) --------------------
        if (!bad) {
            stage_range<NT>(in.A, in.a_pool, ga0, n_a, sm.key);
            stage_range<NT>(in.B, in.b_pool, gb0, n_b, sm.key + boff);
            if constexpr (BSR) {
                stage_range<NT>(in.As, in.a_pool, ga0, n_a, sm.st);
                stage_range<NT>(in.Bs, in.b_pool, gb0, n_b, sm.st + boff);
            }
        }
        if (tid < nseg) {
            const uint64_t p = (uint64_t)p_first + tid;
            const uint64_t up = plan.uoff[p], upn = plan.uoff[p + 1];
            const uint64_t a = in.a_off[p], b = in.b_off[p];
            const uint64_t na = in.a_off[p + 1] - a, nb = in.b_off[p + 1] - b;
            const uint64_t i0 = (tid == 0) ? ga0 - a : 0, j0 = (tid == 0) ? gb0 - b : 0;
            const uint64_t i1 = (tid == nseg - 1) ? ga1 - a : na, j1 = (tid == nseg - 1) ? gb1 - b : nb;
            sm.seg_u[tid] = (up > u0) ? (uint32_t)(up - u0) : 0u;
            sm.seg_flags[tid] = (up >= u0 ? 1u : 0u) | (upn <= u1 ? 2u : 0u);
            sm.seg_la[tid] = shift_a + (uint32_t)(a + i0 - ga0);
            sm.seg_lb[tid] = boff + shift_b + (uint32_t)(b + j0 - gb0);
            sm.seg_wa[tid] = (uint32_t)(i1 - i0);
            sm.seg_wb[tid] = (uint32_t)(j1 - j0);
            sm.seg_nb[tid] = (uint32_t)(nb - j0);
        }
        if (tid == 0) sm.seg_u[nseg] = ulen;
        __syncthreads();

        // ---- 3. per-lane merge --------------------------------------------------------------
        // smask: bit k set when step k emitted; amask: bit (pa - pa0) set for each emitted A
        // element, so the values are re-read from LDS at write time instead of held in registers.
        if (wave == 0) heads = claim_heads<kClaimQueues>(plan.queue, kClaimStride);
        const uint32_t ub = tid * K;
        uint32_t smask = 0, amask = 0, pa0 = 0;
        uint32_t svals[BSR ? K : 1];
        uint32_t nelems = 0;
        if (ub < ulen) {
            uint32_t lo = 0, hi = nseg;  // largest s with seg_u[s] <= ub (seg_u[0] == 0)
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sm.seg_u[mid] <= ub) lo = mid; else hi = mid;
            }
            uint32_t s = lo;
            uint32_t send = sm.seg_u[s + 1];
            uint32_t la = sm.seg_la[s], lb = sm.seg_lb[s];
            uint32_t ea = la + sm.seg_wa[s], eb = lb + sm.seg_wb[s], ebp = lb + sm.seg_nb[s];
            const uint32_t e = ub - sm.seg_u[s];
            const uint32_t wa = ea - la, wb = eb - lb;
            uint32_t pa, pb;
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
            pa0 = pa;
            uint32_t a = sm.key[pa], b = sm.key[pb];
            const uint32_t kmax = (ulen - ub < (uint32_t)K) ? ulen - ub : (uint32_t)K;
#pragma unroll
            for (int k = 0; k < K; k++) {
                if ((uint32_t)k < kmax) {
                    if (ub + k >= send) {
                        // next pair: its window starts at the pair's first elements (contiguous in LDS)
                        s++;
                        send = sm.seg_u[s + 1];
                        la = sm.seg_la[s];
                        lb = sm.seg_lb[s];
                        ea = la + sm.seg_wa[s];
                        eb = lb + sm.seg_wb[s];
                        ebp = lb + sm.seg_nb[s];
                        pa = la;
                        pb = lb;
                        a = sm.key[pa];
                        b = sm.key[pb];
                    }
                    const bool aok = pa < ea;
                    const bool takeA = aok && (pb >= eb || a <= b);
                    const bool real = aok || pb < eb;
                    bool hit = takeA && pb < ebp && a == b;
                    if constexpr (BSR) {
                        uint32_t stv = 0;
                        if (hit) stv = sm.st[pa] & sm.st[pb];
                        hit = hit && stv != 0;
                        svals[k] = stv;
                        nelems += __popc(stv);
                    }
                    if (hit) {
                        smask |= 1u << k;
                        amask |= 1u << (pa - pa0);
                    }
                    if (real) {
                        if (takeA) {
                            pa++;
                            a = sm.key[pa];
                        } else {
                            pb++;
                            b = sm.key[pb];
                        }
                    }
                }
            }
        }
        const uint32_t nwords = __popc(smask);
        if constexpr (!BSR) nelems = nwords;
        acc_w += nwords;
        acc_e += nelems;

        // ---- 4. block scan; publish the aggregate before any waiting; resolve ----------------
        uint32_t tot = 0;
        uint32_t pfx = 0;
        if (!COUNT_ONLY || out.counts) {
            pfx = block_exclusive_scan<NT>(nwords, sm.scan_scratch, &tot);
            sm.lane_mask[tid] = smask;
            sm.lane_pfx[tid] = pfx;
        }
        if constexpr (!COUNT_ONLY) {
            if (wave == 0) {
                lookback_publish(plan.status, t, tot);
                uint64_t ex = 0;
                const bool ok = lookback_resolve(plan.status, t, tot, pending_mode ? kSpinForever : plan.spin_limit, &ex);
                if (lane == 0) {
                    sm.excl = ex;
                    sm.ok = ok ? 1u : 0u;
                }
            }
        }
        if (!COUNT_ONLY || out.counts) __syncthreads();
        const bool ok = COUNT_ONLY ? true : (sm.ok != 0);
        const uint64_t excl = COUNT_ONLY ? 0 : sm.excl;

        // ---- 5. outputs: final positions, or this workgroup's pending area ---------------------
        if constexpr (!COUNT_ONLY) {
            uint32_t *dk = ok ? out.out : pend_out;
            uint32_t *ds = ok ? out.out_s : pend_st;
            const uint64_t cap = ok ? out.cap : (uint64_t)kPendOut;
            uint64_t pos = (ok ? excl : 0) + pfx;
            if (pos + nwords > cap && out.error && nwords) atomicOr(out.error, 1u);
            if constexpr (BSR) {
                uint32_t am = amask;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (smask & (1u << k)) {
                        const uint32_t bit = __builtin_ctz(am);
                        am &= am - 1;
                        if (pos < cap) {
                            dk[pos] = sm.key[pa0 + bit];
                            ds[pos] = svals[k];
                        }
                        pos++;
                    }
                }
            } else {
                uint32_t am = amask;
                while (am) {
                    const uint32_t bit = __builtin_ctz(am);
                    am &= am - 1;
                    if (pos < cap) dk[pos] = sm.key[pa0 + bit];
                    pos++;
                }
            }
        }

        // ---- 6. per-pair offsets and counts --------------------------------------------------
        if ((!COUNT_ONLY && out.out_off) || out.counts) {
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
                if (!COUNT_ONLY && out.out_off) {
                    if (ok) {
                        if (f & 1u) out.out_off[p] = excl + c0v;
                    } else {
                        pend_rec[2 * tid] = (uint32_t)p;
                        pend_rec[2 * tid + 1] = (f & 1u) ? (c0v | 0x80000000u) : 0u;
                    }
                }
                if (out.counts) {
                    if ((f & 3u) == 3u) out.counts[p] = c;
                    else if (c) atomicAdd(&out.counts[p], c);
                }
            }
        }
        if constexpr (!COUNT_ONLY) {
            if (ok && t == plan.ntiles - 1 && tid == 0 && out.out_off) out.out_off[in.P] = excl + tot;
            if (!ok && tid == 0) {
                sm.pend_tile = t;
                sm.pend_tot = tot;
                sm.pend_nseg = nseg;
            }
            __syncthreads();  // pending area and pend_* visible to the whole workgroup
            // ---- 7. resolve the older pending tile without blocking ------------------------
            if (sm.pend_tile != kNone && sm.pend_tile != t) {
                if (wave == 0) {
                    uint64_t ex = 0;
                    const bool r = lookback_resolve(plan.status, sm.pend_tile, sm.pend_tot, 0, &ex);
                    if (lane == 0) {
                        sm.excl = ex;
                        sm.ok = r ? 1u : 0u;
                    }
                }
                __syncthreads();
                if (sm.ok) {
                    flush_pending<NT, BSR>(sm.pend_tile, sm.excl, sm.pend_tot, sm.pend_nseg, pend_out, pend_st, pend_rec,
                                           in, plan, out);
                    __syncthreads();
                    if (tid == 0) sm.pend_tile = kNone;
                }
            }
        }
    }

    if constexpr (!COUNT_ONLY) {
        // every tile is claimed now, so waiting for the pending tile's predecessors always ends
        if (sm.pend_tile != kNone) {
            if (wave == 0) {
                uint64_t ex = 0;
                (void)lookback_resolve(plan.status, sm.pend_tile, sm.pend_tot, kSpinForever, &ex);
                if (lane == 0) sm.excl = ex;
            }
            __syncthreads();
            flush_pending<NT, BSR>(sm.pend_tile, sm.excl, sm.pend_tot, sm.pend_nseg, pend_out, pend_st, pend_rec, in,
                                   plan, out);
        }
    }

    if (out.totals) {
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            acc_w += __shfl_xor(acc_w, d, kWave);
            acc_e += __shfl_xor(acc_e, d, kWave);
        }
        __syncthreads();
        if (lane == 0) {
            sm.wave_w[wave] = acc_w;
            sm.wave_e[wave] = acc_e;
        }
        __syncthreads();
        if (tid == 0) {
            uint64_t w = 0, e = 0;
            for (int k = 0; k < NT / kWave; k++) {
                w += sm.wave_w[k];
                e += sm.wave_e[k];
            }
            if (w | e) {
                atomicAdd(&out.totals[0], (unsigned long long)w);
                atomicAdd(&out.totals[1], (unsigned long long)e);
            }
        }
    }
}

} // namespace setops
