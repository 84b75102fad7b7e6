// merge_tile.hpp -- merge-path tile kernel for batched sorted-set intersection (gfx950).
//
// One workgroup per tile of U = NT*K merge-path units (tile_plan.hpp).  Per tile:
//   1. the covered pairs' windows: the segment table (one thread per pair),
//   2. stage the windows in LDS with 16-byte loads -- CSR batches as one A run and one B run,
//      gather batches window by window; every B window carries the element after it (if the pair
//      has one) for the boundary compare,
//   3. every lane finds its K-unit slice by an LDS merge-path search and walks K merge steps,
//      emitting on an A step whose value equals the current B value (exactly once per match,
//      independent of where tile and lane boundaries fall),
//   4. block scan of per-lane match counts; results written in order to the tile's staging slot
//      from LDS (plain merge: first gathered in order into the staged B words, then written in
//      whole rows),
//   5. tile count, per-pair counts and in-tile offsets.
//
// The same kernel handles the BSR representation (key = base, payload = 32-bit state): a base
// match emits (base, state_a & state_b) when the AND is non-zero (merge.rs:62-87).
#pragma once

#include "tile_plan.hpp"

namespace setops {

template <int NT, int K, int MAXSEG, bool BSR, bool GATHER, bool BLOCKP>
struct MergeSmem {
    static constexpr int U = NT * K;
    // CSR: the A run and the B run (+ alignment shifts); gather: every window rounded out to
    // whole 16-byte chunks (<= 7 extra words per A and B window, +1 boundary element)
    static constexpr int KEYS = GATHER ? U + 16 * MAXSEG + 16 : U + 32;
    // + K + 8: the merge steps of a lane may read up to K words past its windows
    alignas(16) uint32_t key_guard[4];  // key[-1] for the split search: read, never used
    alignas(16) uint32_t key[KEYS + K + 8];  // starts 16 bytes after key_guard (checked below)
    alignas(16) uint32_t st[BSR ? KEYS : 4];
    static_assert(KEYS + K + 8 < 65536 && U < 65536, "segment fields are 16-bit LDS indices");
    // merge policy: per-segment result prefixes replace the per-lane masks (LDS is what limits how
    // many workgroups share a CU, so the segment table is 16-bit)
    static constexpr bool LEAN = !BLOCKP;
    uint16_t seg_u[MAXSEG + 1];  // tile-relative unit where segment s starts; [nseg] = tile length
    uint16_t seg_la[MAXSEG];     // LDS index of the segment's first A element in the window
    uint16_t seg_lb[MAXSEG];     // LDS index of the segment's first B element in the window
    uint16_t seg_wa[MAXSEG];     // A elements in the window
    uint16_t seg_wb[MAXSEG];     // B elements in the window (the boundary element excluded)
    uint16_t seg_nb[MAXSEG];     // B elements readable from the window start (wb, +1 with a boundary)
    uint8_t seg_flags[MAXSEG];   // bit0: pair starts in this tile, bit1: pair ends in this tile
    uint32_t seg_pre[LEAN ? MAXSEG + 1 : 1];  // plain merge: the tile's results before segment s
    uint32_t lane_mask[LEAN ? 1 : NT];        // other policies: per-lane step mask of matches
    uint32_t lane_pfx[LEAN ? 1 : NT];         // other policies: per-lane exclusive prefix of matches
    uint32_t lane_nfirst[BLOCKP ? NT : 1];  // block policies: lane's A elements before a pair boundary
    uint32_t scan_scratch[NT / kWave];
    uint32_t wave_el[NT / kWave];
    // gather staging: chunk prefix, aligned global start and source pool of each window
    uint32_t win_c[GATHER ? 2 * MAXSEG + 1 : 1];
    uint64_t win_g[GATHER ? 2 * MAXSEG : 1];
    uint32_t win_src[GATHER ? 2 * MAXSEG : 1];
    // byte offset of key[i] from key_guard is 4 * (i + 4): the split search addresses both from there
    static constexpr uint32_t kGuardWords = 4;
};

// Lane policies: how a lane intersects its slice of the tile.
enum : int {
    LP_MERGE = 0,    // K branch-free merge steps (merge.rs:39-87)
    LP_SHUFFLE = 1,  // 4x4 all-pairs block compares (shuffling.rs:26-168, 174-356)
    LP_QFILTER = 2,  // 4x4 blocks with a low-byte prefilter (qfilter.rs:32-180)
};

// A-count of the merge-path split of diagonal d in the window A = key[la, la+wa), B = key[lb, lb+wb)
// (ties go to A, as in the merge step).
__device__ __forceinline__ uint32_t mp_split(const uint32_t *key, uint32_t la, uint32_t wa, uint32_t lb, uint32_t wb,
                                             uint32_t d)
{
    uint32_t lo = (d > wb) ? d - wb : 0u, hi = (d < wa) ? d : wa;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key[la + mid] <= key[lb + d - 1 - mid]) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// One merge step on the current A value a and B value b (branchless_merge, merge.rs:39-60): the
// step takes A when a <= b (ties go to A) and matches when a == b.  Both masks gain the step's bit
// at the bottom, doubling as they go (step k of K ends at bit K-1-k), and ia counts the A elements
// taken.  Each compare result is the carry-in of an add (v_addc_co_u32 with VCC): one VALU per mask
// bit.  Then the next A element's byte offset oa = a4 + 4 ia and bo = s4 - oa, from which the
// caller forms the next B element's offset (s4 + 4(k+1) - oa): seven VALU per step.  s_nop 1 gives
// the two wait states that a VALU read of VCC as a carry needs after a VALU write of it.  Pure
// register arithmetic (no memory access).
__device__ __forceinline__ void merge_step(uint32_t a, uint32_t b, uint32_t a4, uint32_t s4, uint32_t &tmask,
                                           uint32_t &smask, uint32_t &ia, uint32_t &oa, uint32_t &bo)
{
    uint64_t c0, c1, c2;  // carry-outs (unused)
    asm("v_cmp_le_u32 vcc, %[a], %[b]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[t], %[c0], %[t], %[t], vcc\n\t"
        "v_addc_co_u32_e64 %[i], %[c1], %[i], 0, vcc\n\t"
        "v_cmp_eq_u32 vcc, %[a], %[b]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %[s], %[c2], %[s], %[s], vcc\n\t"
        "v_lshl_add_u32 %[oa], %[i], 2, %[a4]\n\t"
        "v_sub_u32 %[bo], %[s4], %[oa]"
        : [t] "+v"(tmask), [i] "+v"(ia), [s] "+v"(smask), [oa] "=&v"(oa), [bo] "=v"(bo), [c0] "=&s"(c0),
          [c1] "=&s"(c1), [c2] "=&s"(c2)
        : [a] "v"(a), [b] "v"(b), [a4] "v"(a4), [s4] "v"(s4)
        : "vcc");
}

// Bytes of x that are zero -> 0x80 in that byte, else 0 (exact, no carries across bytes).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x)
{
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// Block intersection of two short sorted windows A[0, na), B[0, nb) in LDS: W = 4 blocks, all 16
// pairs compared, then the block with the smaller (or equal) maximum advances by 4 -- the block
// rule of shuffling.rs:42-66, with partial blocks masked instead of a scalar tail.  QFILTER first
// compares the blocks' low bytes (four rotations of one packed word, SWAR) and skips the full
// compares when no byte pair agrees -- qfilter's byte-check prefilter (qfilter.rs:32-89) in 32-bit
// SWAR form.  Returns bit i set when A[i] matched (BSR: when the state AND is non-zero); nelems
// accumulates the popcount of the emitted states (BSR).
template <bool QF, bool BSR>
__device__ __forceinline__ uint32_t block_intersect(const uint32_t *key, const uint32_t *st, uint32_t ia,
                                                    uint32_t na, uint32_t ib, uint32_t nb, uint32_t &nelems)
{
    uint32_t mask = 0, i = 0, j = 0;
    while (i < na && j < nb) {
        uint32_t a[4], b[4];
#pragma unroll
        for (int x = 0; x < 4; x++) {
            a[x] = key[ia + i + x];
            b[x] = key[ib + j + x];
        }
        const uint32_t av = (na - i < 4u) ? na - i : 4u, bv = (nb - j < 4u) ? nb - j : 4u;
        const uint32_t amax = key[ia + i + av - 1], bmax = key[ib + j + bv - 1];
        bool any = true;
        if constexpr (QF) {
            const uint32_t alo = (a[0] & 0xFFu) | (a[1] & 0xFFu) << 8 | (a[2] & 0xFFu) << 16 | a[3] << 24;
            const uint32_t blo = (b[0] & 0xFFu) | (b[1] & 0xFFu) << 8 | (b[2] & 0xFFu) << 16 | b[3] << 24;
            uint32_t cand = 0;
#pragma unroll
            for (int r = 0; r < 4; r++) cand |= zero_bytes(alo ^ ((blo >> (8 * r)) | (blo << ((32 - 8 * r) & 31))));
            any = cand != 0;
        }
        if (any) {
#pragma unroll
            for (int x = 0; x < 4; x++) {
                uint32_t hit = 0, stv = 0;
#pragma unroll
                for (int y = 0; y < 4; y++) {
                    const bool eq = (uint32_t)x < av && (uint32_t)y < bv && a[x] == b[y];
                    if constexpr (BSR) stv |= eq ? (st[ia + i + x] & st[ib + j + y]) : 0u;
                    else hit |= eq ? 1u : 0u;
                }
                if constexpr (BSR) {
                    hit = stv != 0 ? 1u : 0u;
                    nelems += __popc(stv);
                }
                mask |= hit << (i + x);
            }
        }
        i += (amax <= bmax) ? 4u : 0u;
        j += (bmax <= amax) ? 4u : 0u;
    }
    return mask;
}

// Steps 3-6 of a tile whose windows and segment table are in LDS (seg_* arrays, key/st): the
// lanes' merge (or block policy), the tile scan, staged outputs, tile and per-pair counts.
template <int NT, int K, int MAXSEG, bool BSR, bool COUNT_ONLY, bool GATHER, int LP>
__device__ __forceinline__ void merge_tile_compute(MergeSmem<NT, K, MAXSEG, BSR, GATHER, LP != LP_MERGE> &sm,
                                                   const TilePlan &plan, const TileOut &out, uint32_t t,
                                                   uint32_t nseg, uint32_t ulen, uint32_t p_first, uint32_t in_sflip,
                                                   uint32_t nstaged, uint32_t bstart)
{
    constexpr bool BLOCKP = LP != LP_MERGE;
    using Smem = MergeSmem<NT, K, MAXSEG, BSR, GATHER, BLOCKP>;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / kWave;
    {
        // ---- 3. per-lane merge ---------------------------------------------------------------------
        // smask: bit k set when step k emitted; tmask: bit k set when step k consumed an A element.
        // The emitted values are A elements, re-read from LDS at write time: the A offset of the
        // emit at step k is popcount(tmask & ((1 << k) - 1)) from where the lane's merge starts
        // (pa0: its own pair, or the next pair's window when it starts there at step kt).
        const uint32_t ub = tid * K;
        uint32_t smask = 0, tmask = 0, pa0 = 0, pa1 = 0;
        uint32_t nelems = 0;
        uint32_t shift_o = 0, pbl = 0;  // merge policy: masks' shift, loop's first B index
        uint32_t kt = K;  // merge policy: step at which the lane enters the next pair
        uint32_t s_first = 0;  // plain merge: the lane's first segment
        bool lane_on = false;  // plain merge: the lane has units in the tile
        // block policies: smask bit x is the lane's x-th A element (first sub-window, then the
        // next pair's); pa0 / pa1 locate them, pb0/nb0 and pb1/nb1 the B sub-windows (BSR output)
        uint32_t nfirst = 0, pb0 = 0, nb0 = 0, pb1 = 0, nb1 = 0;
        if constexpr (BLOCKP) {
            if (ub < ulen) {
                uint32_t lo = 0, hi = nseg;  // largest s with seg_u[s] <= ub
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (sm.seg_u[mid] <= ub) lo = mid; else hi = mid;
                }
                const uint32_t ue = (ulen - ub < (uint32_t)K) ? ulen : ub + K;
                for (uint32_t part = 0, s = lo; part < 2; part++, s++) {
                    const uint32_t su = sm.seg_u[s];
                    const uint32_t sn = sm.seg_u[s + 1];
                    if (part == 1 && ue <= su) break;
                    const uint32_t la = sm.seg_la[s], lb = sm.seg_lb[s];
                    const uint32_t wa = sm.seg_wa[s], wb = sm.seg_wb[s], wn = wa + wb;
                    uint32_t d0 = (ub > su) ? ub - su : 0u, d1 = ((ue < sn) ? ue : sn) - su;
                    d0 = d0 < wn ? d0 : wn;
                    d1 = d1 < wn ? d1 : wn;
                    const uint32_t i0 = mp_split(sm.key, la, wa, lb, wb, d0);
                    const uint32_t i1 = mp_split(sm.key, la, wa, lb, wb, d1);
                    const uint32_t j0 = d0 - i0, j1 = d1 - i1;
                    const uint32_t nbw = j1 - j0 + ((j1 < sm.seg_nb[s]) ? 1u : 0u);  // + boundary element
                    const uint32_t m = block_intersect<LP == LP_QFILTER, BSR>(sm.key, sm.st, la + i0, i1 - i0,
                                                                              lb + j0, nbw, nelems);
                    if (part == 0) {
                        smask = m;
                        nfirst = i1 - i0;
                        pa0 = la + i0;
                        pb0 = lb + j0;
                        nb0 = nbw;
                    } else {
                        smask |= m << nfirst;
                        pa1 = la + i0;
                        pb1 = lb + j0;
                        nb1 = nbw;
                    }
                    if (ue <= sn) break;
                }
            }
        } else {
        uint32_t kmax = 0;  // steps to run
        uint32_t pa = 0, pb = 0, ea = 0, eb = 0, ebp = 0, s = 0;
        // the lane's split search (q_n = 0: none; it then probes key[-1] against key[0])
        uint32_t q_lo = 0, q_n = 0, q_a4 = 12u, q_s4 = 28u;
        bool q_on = false;
        lane_on = ub < ulen;
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
                pa = la;
                pb = lb + e;
                q_lo = (e > wb) ? e - wb : 0u;
                q_n = ((e < wa) ? e : wa) - q_lo;
                // byte offsets from key_guard (key[i] at 4 * (i + 4)): A probe key[la + y - 1], B probe
                // key[lb + e - y]; their sum is the lane constant q_s4
                q_a4 = 4u * (la + 3u);
                q_s4 = 4u * (la + lb + e + 7u);
                q_on = true;
            }
            kmax = (ulen - ub < (uint32_t)K) ? ulen - ub : (uint32_t)K;
            const uint32_t send = sm.seg_u[s + 1];  // > ub
            if (send - ub < kmax) kt = send - ub;   // every pair owns >= c0 >= K units: at most one
        }
        {
            // The lane's merge-path split: the largest x in [q_lo, q_lo + q_n] with
            // A[la + x'] <= B[lb + e - 1 - x'] for every x' < x (ties go to A), by binary lifting
            // with the wave's common step sequence (steps 2^(L-1) .. 1, L the wave's longest
            // range); a step that would pass the range end stops at it and changes nothing.
            // Probe x = y - 1 reads A at byte 4(la + y - 1) and B at byte 4(lb + e - y) of key, both
            // addressed from key_guard (16 bytes lower, so key[-1] = key_guard[3] is in range).
            const uint32_t L = wave_max_u32(32u - (uint32_t)__clz(q_n));
            static_assert(offsetof(Smem, key) == offsetof(Smem, key_guard) + 4 * Smem::kGuardWords,
                          "key follows key_guard directly");
            const char *kb = reinterpret_cast<const char *>(sm.key_guard);
            // the search runs on A's probe byte offset xo = q_a4 + 4x (five VALU per step: the
            // clamped step, B's offset, the compare and the select)
            const uint32_t hio = q_a4 + 4u * (q_lo + q_n);
            uint32_t xo = q_a4 + 4u * q_lo;
            for (int bit = (int)L - 1; bit >= 0; bit--) {
                const uint32_t yo = min(xo + (4u << bit), hio);
                if (*reinterpret_cast<const uint32_t *>(kb + yo) <= *reinterpret_cast<const uint32_t *>(kb + (q_s4 - yo)))
                    xo = yo;
            }
            const uint32_t x = (xo - q_a4) >> 2;
            if (q_on) {
                pa += x;  // pa / pb hold the window starts la / lb + e here
                pb -= x;
            }
        }
        {
            // One branch-free loop of K merge steps for every wave; a step re-reads both current
            // values (two LDS reads, no selects).  What a checked merge step
            // would test each time is settled per lane before and after the loop:
            //  - a lane whose units run into the next pair (at step kt < K) spends steps [0, kt)
            //    in its pair's padding (every pair ends with c0 >= K padding units): it starts at
            //    the next pair's window instead, keeps K - kt steps, and its masks move up by kt;
            //  - once a side is used up, the steps read past its window and decide arbitrarily,
            //    so a step emits only while both sides still hold real elements: the A element
            //    taken exists (fewer than rA A elements taken before it) and so does the current
            //    B element (fewer than rB taken before it, or the staged element after the
            //    window); steps past the tile's end emit nothing.
            uint32_t shift = 0;
            s_first = s;
            if (kt < (uint32_t)K) {
                shift = kt;
                s++;
                pa = sm.seg_la[s];
                pb = sm.seg_lb[s];
                ea = pa + sm.seg_wa[s];
                eb = pb + sm.seg_wb[s];
                ebp = pb + sm.seg_nb[s];
            }
            const uint32_t kv = kmax - shift;
            pa0 = pa;
            pbl = pb;
            shift_o = shift;
            uint32_t a = sm.key[pa], b = sm.key[pb];
            // A's current element is at byte 4(pa + ia) of key, B's at 4(pb + k - ia): their sum
            // is s4 + 4k.  The masks come out of the loop with step k at bit K-1-k.
            const char *kb = reinterpret_cast<const char *>(sm.key);
            const uint32_t s4 = 4u * (pa + pb), a4 = 4u * pa;
            uint32_t ia = 0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                uint32_t oa, bo;
                merge_step(a, b, a4, s4, tmask, smask, ia, oa, bo);
                a = *reinterpret_cast<const uint32_t *>(kb + oa);
                // B's byte offset 4(pb + k + 1 - taken A) >= 4 pb (bo + 4(k+1), 32-bit arithmetic)
                b = *reinterpret_cast<const uint32_t *>(kb + (bo + 4u * (uint32_t)(k + 1)));
            }
            // step k to bit k
            tmask = __builtin_bitreverse32(tmask) >> (32 - K);
            smask = __builtin_bitreverse32(smask) >> (32 - K);
            smask &= (kv >= 32u) ? ~0u : ((1u << kv) - 1u);
            const uint32_t rA = ea - pa;
            const uint32_t rB = (ebp > eb) ? ~0u : eb - pb;
            const bool edge = (rA < (uint32_t)K || rB < (uint32_t)K) && smask != 0u;
            if (ballot(edge) != 0 && edge) {
                uint32_t m = smask, keep = 0;
                while (m) {
                    const uint32_t k = __builtin_ctz(m);
                    m &= m - 1;
                    const uint32_t ta = __popc(tmask & ((1u << k) - 1u));  // A taken before step k
                    keep |= (ta < rA && k - ta < rB) ? (1u << k) : 0u;
                }
                smask = keep;
            }
            if constexpr (BSR) {
                // a base match emits (base, state_a & state_b) only when the AND is non-zero
                // (merge.rs:62-87): the A element taken at step k and the B element current then
                uint32_t m = smask, keep = 0;
                while (m) {
                    const uint32_t k = __builtin_ctz(m);
                    m &= m - 1;
                    const uint32_t ta = __popc(tmask & ((1u << k) - 1u));
                    const uint32_t stv = sm.st[pa + ta] & sm.st[pb + k - ta];
                    keep |= stv ? (1u << k) : 0u;
                    nelems += __popc(stv);
                }
                smask = keep;
            }
            smask <<= shift;
            tmask <<= shift;
        }
        }
        const uint32_t nwords = __popc(smask);
        if constexpr (!BSR) nelems = nwords;

        // ---- 4. tile scan -----------------------------------------------------------------------
        uint32_t tot = 0;
        const uint32_t pfx = block_exclusive_scan<NT>(nwords, sm.scan_scratch, &tot);
        if constexpr (Smem::LEAN) {
            // the segment that starts in this lane's units starts at its first unit, or at step
            // kt after padding: either way no result of the lane precedes it
            if (lane_on && sm.seg_u[s_first] == ub) sm.seg_pre[s_first] = pfx;
            if (lane_on && kt < (uint32_t)K) sm.seg_pre[s_first + 1] = pfx;
            if (tid == 0) sm.seg_pre[nseg] = tot;
        } else {
            sm.lane_mask[tid] = smask;
            sm.lane_pfx[tid] = pfx;
        }
        if constexpr (BLOCKP) sm.lane_nfirst[tid] = nfirst;
        if constexpr (BSR) {
            uint32_t el = nelems;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) el += __shfl_xor(el, dd, kWave);
            if (lane == 0) sm.wave_el[wave] = el;
        }
        __syncthreads();

        // plain merge: results are gathered in order into the B part of the staged words (key[bstart
        // ..)), which the output phase no longer reads, and copied out in whole rows
        constexpr bool COMPACT = !BLOCKP && !BSR && !COUNT_ONLY;
        // int32 sets: the keys go back to their given bits before they are written out (the merges
        // that compared them are done: the scan above waited for every wave)
        if constexpr (!COUNT_ONLY && !BSR && !COMPACT) {
            if (in_sflip) {
                for (uint32_t i = tid; i < nstaged; i += NT) sm.key[i] ^= in_sflip;
                __syncthreads();
            }
        }

        // ---- 5. staged outputs (in order inside the tile's slot) ----------------------------------
        // Sorted unique windows give at most (K + 1) / 2 results per lane, which the slot holds; a
        // tile planned from sizes that disagree with the device offsets may not, and then writes
        // nothing (error bit 1) rather than past its slot.
        const uint32_t slot_keys = BSR ? plan.stage_words / 2 : plan.stage_words;
        const uint32_t lds_room = COMPACT ? (nstaged > bstart ? nstaged - bstart : 0u) : ~0u;
        if (tot > slot_keys || tot > lds_room) {
            if (tid == 0 && out.error) atomicOr(out.error, 2u);
            tot = 0;
            smask = 0;
        }
        if constexpr (!COUNT_ONLY) {
            uint32_t *dk = plan.stage + (uint64_t)t * plan.stage_words;
            uint32_t *ds = dk + plan.stage_words / 2;
            uint32_t pos = pfx;
            if constexpr (BLOCKP) {
                uint32_t sm_ = smask;
                while (sm_) {
                    const uint32_t x = __builtin_ctz(sm_);
                    sm_ &= sm_ - 1;
                    const uint32_t ai = (x < nfirst) ? pa0 + x : pa1 + (x - nfirst);
                    const uint32_t v = sm.key[ai];
                    dk[pos] = v;
                    if constexpr (BSR) {
                        // the matching B base lies in the same sub-window: find it again
                        const uint32_t jb = (x < nfirst) ? pb0 : pb1, nbw = (x < nfirst) ? nb0 : nb1;
                        uint32_t stv = 0;
                        for (uint32_t y = 0; y < nbw; y++) stv |= (sm.key[jb + y] == v) ? sm.st[jb + y] : 0u;
                        ds[pos] = sm.st[ai] & stv;
                    }
                    pos++;
                }
            } else if constexpr (BSR) {
                // the matches' bases and state ANDs, re-read from LDS (A at the emit step, B the
                // element current then)
                uint32_t sm_ = smask;
                while (sm_) {
                    const uint32_t k = __builtin_ctz(sm_);
                    sm_ &= sm_ - 1;
                    const uint32_t ta = __popc(tmask & ((1u << k) - 1u));
                    const uint32_t ai = pa0 + ta, bi = pbl + (k - shift_o) - ta;
                    dk[pos] = sm.key[ai];
                    ds[pos] = sm.st[ai] & sm.st[bi];
                    pos++;
                }
            } else {
                // the A element taken at step k is the lane's A element number popc(tmask below k),
                // counted from pa0 (a lane that moved to the next pair started there and has no
                // steps below kt).  Each result goes to key[bstart + its tile position] (B words:
                // a tile has no more results than staged B elements, and A words are all that the
                // loop reads), then the tile's results leave in order, a row of NT words per store
                uint32_t w = bstart + pos;
                uint32_t sm_ = smask;
                while (sm_) {
                    const uint32_t k = __builtin_ctz(sm_);
                    sm_ &= sm_ - 1;
                    sm.key[w++] = sm.key[pa0 + __popc(tmask & ((1u << k) - 1u))];
                }
                __syncthreads();
                for (uint32_t i = tid; i < tot; i += NT) dk[i] = sm.key[bstart + i] ^ in_sflip;
            }
        }

        // ---- 6. tile counts, per-pair counts and in-tile offsets ----------------------------------
        if (tid == 0) {
            uint32_t el = tot;
            if constexpr (BSR) {
                el = 0;
                for (int k = 0; k < NT / kWave; k++) el += sm.wave_el[k];
            }
            plan.tile_count[t] = tot;
            plan.tile_elems[t] = el;
            if (plan.direct && out.totals) {  // one tile: its counts are the totals
                out.totals[0] = tot;
                out.totals[1] = el;
            }
        }
        if (tid < nseg) {
            // results of the tile's units before x: whole lanes, then the lane's results before x
            // (merge: steps before x; block policies: the lane's A elements before a pair boundary)
            auto before = [&](uint32_t x) -> uint32_t {
                if (x >= ulen) return tot;
                const uint32_t l = x / K, k = x % K;
                uint32_t nb_ = k;
                if constexpr (BLOCKP) nb_ = (k == 0) ? 0u : sm.lane_nfirst[l];
                return sm.lane_pfx[l] + __popc(sm.lane_mask[l] & ((1u << nb_) - 1u));
            };
            uint32_t c0v, c;
            if constexpr (Smem::LEAN) {
                c0v = sm.seg_pre[tid];
                c = sm.seg_pre[tid + 1] - c0v;
            } else {
                c0v = before(sm.seg_u[tid]);
                c = before(sm.seg_u[tid + 1]) - c0v;
            }
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

// Count + stage: one workgroup per tile (tiles tile0 + blockIdx.x), no communication between
// workgroups.  Each tile's results go to its own staging slot (or nowhere in count-only mode);
// its counts go to tile_count / tile_elems; per-pair counts and in-tile offsets are recorded for
// the placement kernel.
template <int NT, int K, int MAXSEG, bool BSR, bool COUNT_ONLY, bool GATHER, int LP = LP_MERGE, bool DIRECT = false>
__global__ __launch_bounds__(NT) void merge_tile_kernel(PairBatch in, TilePlan plan, TileOut out, uint32_t tile0)
{
    static_assert(K >= 1 && K <= 32, "lane masks hold K bits");
    static_assert(!DIRECT || (!GATHER && !BSR), "a direct tile is one plain CSR pair");
    constexpr int U = NT * K;
    constexpr bool BLOCKP = LP != LP_MERGE;
    using Smem = MergeSmem<NT, K, MAXSEG, BSR, GATHER, BLOCKP>;
    __shared__ Smem sm;
    const uint32_t tid = threadIdx.x;

    {
        // one tile per workgroup (a loop over tiles here costs ~20% on the merge kernel)
        const uint32_t t = tile0 + blockIdx.x;
        if (t >= plan.ntiles) return;
        // ---- 1. descriptor and segment windows -------------------------------------------------
        // CSR: the tile's A and B runs load with its pair range (one round trip before the
        // staging is issued, not two)
        // DIRECT (one pair in one tile, CSR form, the launch for plan.direct): the tile is the whole
        // pair and nothing planned it, so its runs, its one segment and the error word come from the
        // plan's fields here.  A template flag: a runtime branch here costs the batch tiles 3-5%.
        constexpr bool direct = DIRECT;
        uint64_t ga0 = 0, ga1 = 0, gb0 = 0, gb1 = 0;
        if constexpr (!GATHER) {
            if constexpr (direct) {
                ga1 = plan.single_na;
                gb1 = plan.single_nb;
                if (tid == 0 && out.error) *out.error = 0;  // before thread 0's own flags below
            } else {
                ga0 = plan.tile_ga[t];
                ga1 = plan.tile_ga[t + 1];
                gb0 = plan.tile_gb[t];
                gb1 = plan.tile_gb[t + 1];
            }
        }
        const uint32_t p_first = direct ? 0u : plan.tile_first[t], p_last = direct ? 0u : plan.tile_last[t];
        const uint64_t u0 = (uint64_t)t * U;
        uint32_t ulen = (uint32_t)((u0 + U < plan.utotal) ? U : plan.utotal - u0);
        uint32_t nseg = p_last - p_first + 1;
        bool bad = p_last < p_first || nseg > (uint32_t)MAXSEG || p_last >= in.P;
        if (bad) nseg = 0;
        if (t == 0 && tid == 0 && out.error && !direct && plan.uoff[in.P] != plan.utotal) atomicOr(out.error, 4u);
        SegWin w{};
        uint32_t staged_words = 0;  // CSR: words loaded straight into LDS (flipped in place for int32 sets)
        uint32_t bstart = 0;        // key index where the staged B words start

        // ---- 2. stage the windows --------------------------------------------------------------
        if constexpr (!GATHER) {
            // CSR: the tile's A run and B run come straight from the split (issued before the
            // segment table's loads)
            if (!bad) bad = ga1 < ga0 || gb1 < gb0 || (ga1 - ga0) + (gb1 - gb0) > (uint64_t)U;
            const uint32_t n_a = bad ? 0u : (uint32_t)(ga1 - ga0);
            const uint32_t n_b = bad ? 0u : (uint32_t)(gb1 - gb0) + ((gb1 < in.b_pool) ? 1u : 0u);
            const uint32_t shift_a = (uint32_t)(ga0 & 3), shift_b = (uint32_t)(gb0 & 3);
            const uint32_t cA = (shift_a + n_a + 3) / 4, cB = (shift_b + n_b + 3) / 4;
            const uint32_t boff = 4u * cA;
            staged_words = 4u * (cA + cB);
            bstart = boff;
            // A run at key[shift_a ...], B run at key[boff + shift_b ...] (boff = the A run's chunks),
            // loaded straight into LDS; the segment table's loads are in flight meanwhile
            if (!bad) {
                stage_run_lds<NT>(in.A, in.a_pool, ga0 & ~3ull, cA, sm.key);
                stage_run_lds<NT>(in.B, in.b_pool, gb0 & ~3ull, cB, sm.key + boff);
                if constexpr (BSR) {
                    stage_run_lds<NT>(in.As, in.a_pool, ga0 & ~3ull, cA, sm.st);
                    stage_run_lds<NT>(in.Bs, in.b_pool, gb0 & ~3ull, cB, sm.st + boff);
                }
            }
            if (tid < nseg) w = direct ? direct_window<false>(plan) : seg_window(in, plan, t, p_first, nseg, tid, ulen);
            if (tid < nseg) {
                sm.seg_la[tid] = shift_a + (uint32_t)(w.a + w.i0 - ga0);
                sm.seg_lb[tid] = boff + shift_b + (uint32_t)(w.b + w.j0 - gb0);
            }
        } else {
            if (tid < nseg) w = seg_window(in, plan, t, p_first, nseg, tid, ulen);
            uint32_t ca = 0, cb = 0;
            uint64_t ga = 0, gb = 0;
            if (tid < nseg) {
                ga = w.a + w.i0;
                gb = w.b + w.j0;
                const uint32_t nwa = (uint32_t)(w.i1 - w.i0);
                const uint32_t nwb = (uint32_t)(w.j1 - w.j0) + (w.j1 < w.nb ? 1u : 0u);
                ca = nwa ? ((uint32_t)(ga & 3) + nwa + 3) / 4 : 0u;
                cb = nwb ? ((uint32_t)(gb & 3) + nwb + 3) / 4 : 0u;
            }
            // per-window chunk counts fit 16 bits (a tile holds <= U + 8*MAXSEG words)
            uint32_t tot = 0;
            const uint32_t pfx = block_exclusive_scan<NT>(ca | (cb << 16), sm.scan_scratch, &tot);
            const uint32_t totA = tot & 0xFFFFu, totB = tot >> 16;
            if (4u * (totA + totB) + 8u > (uint32_t)Smem::KEYS) bad = true;  // not reachable by the unit bound
            if (tid < nseg) {
                const uint32_t pa_ = pfx & 0xFFFFu, pb_ = totA + (pfx >> 16);
                sm.win_c[tid] = pa_;
                sm.win_c[nseg + tid] = pb_;
                sm.win_g[tid] = ga & ~3ull;
                sm.win_g[nseg + tid] = gb & ~3ull;
                sm.win_src[tid] = 0;
                sm.win_src[nseg + tid] = 1;
                sm.seg_la[tid] = 4u * pa_ + (uint32_t)(ga & 3);
                sm.seg_lb[tid] = 4u * pb_ + (uint32_t)(gb & 3);
            }
            if (tid == 0) sm.win_c[2 * nseg] = totA + totB;
            staged_words = 4u * (totA + totB);
            bstart = 4u * totA;
            __syncthreads();
            constexpr int QWIN = (Smem::KEYS + 4 * NT - 1) / (4 * NT);
            if (!bad) {
                stage_windows<NT, QWIN>(in.A, in.a_pool, in.B, in.b_pool, sm.win_c, sm.win_g, sm.win_src, 2 * nseg,
                                        sm.key, in.sflip);
                if constexpr (BSR)
                    stage_windows<NT, QWIN>(in.As, in.a_pool, in.Bs, in.b_pool, sm.win_c, sm.win_g, sm.win_src,
                                            2 * nseg, sm.st, 0u);
            }
        }
        if (bad) {
            nseg = 0;
            ulen = 0;
            if (tid == 0 && out.error) atomicOr(out.error, 2u);
        }
        if (tid < nseg) {
            const uint32_t wb = (uint32_t)(w.j1 - w.j0);
            sm.seg_u[tid] = w.u;
            sm.seg_flags[tid] = w.flags;
            sm.seg_wa[tid] = (uint32_t)(w.i1 - w.i0);
            sm.seg_wb[tid] = wb;
            sm.seg_nb[tid] = wb + (w.j1 < w.nb ? 1u : 0u);
        }
        if (tid == 0) sm.seg_u[nseg] = ulen;
        if constexpr (!GATHER) __builtin_amdgcn_s_waitcnt(0);  // the direct-to-LDS loads have landed
        __syncthreads();
        if constexpr (!GATHER && !BSR) {
            // int32 sets: flip bit 31 of the staged words, so unsigned compares give signed order
            if (in.sflip) {
                for (uint32_t i = tid; i < staged_words; i += NT) sm.key[i] ^= in.sflip;
                __syncthreads();
            }
        }

        merge_tile_compute<NT, K, MAXSEG, BSR, COUNT_ONLY, GATHER, LP>(sm, plan, out, t, nseg, ulen, p_first,
                                                                       in.sflip, staged_words, bstart);
    }
}

} // namespace setops
