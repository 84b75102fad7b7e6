/*
 * setops_oracle.c -- CPU restatement of the reference's sorted-u32 intersection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product path
 * (simd_set_operations_amd + libsetops_hip.so) never links or calls it.
 *
 * Every function restates one function of UNSW-database/simd_set_operations (paths below are
 * relative to that repository) in plain C, written from the algorithm's description:
 *
 *   oracle_naive_merge            setops/src/intersect/merge.rs:7-33
 *   oracle_branchless_merge       setops/src/intersect/merge.rs:39-60      (parity reference)
 *   oracle_branchless_merge_bsr   setops/src/intersect/merge.rs:62-87
 *   oracle_binary_search          setops/src/intersect/galloping.rs:110-131
 *   oracle_galloping              setops/src/intersect/galloping.rs:7-30
 *   oracle_binary_search_intersect setops/src/intersect/galloping.rs:32-49
 *   oracle_galloping_bsr          setops/src/intersect/galloping.rs:51-76
 *   oracle_shuffling              setops/src/intersect/shuffling.rs:26-168 (W = 4, 8, 16)
 *   oracle_shuffling_bsr          setops/src/intersect/shuffling.rs:174-356
 *   oracle_broadcast              setops/src/intersect/broadcast.rs:21-160 (W = 4, 8, 16)
 *   oracle_broadcast_bsr          setops/src/intersect/broadcast.rs:163-320
 *   oracle_simd_galloping         setops/src/intersect/simd_galloping.rs:19-107, 196-296
 *   oracle_simd_galloping_bsr     setops/src/intersect/simd_galloping.rs:109-193
 *   oracle_qfilter                setops/src/intersect/qfilter.rs:32-89 (+ dictionary :254-357)
 *   oracle_qfilter_bsr            setops/src/intersect/qfilter.rs:92-180
 *   oracle_bsr_from_sorted        setops/src/bsr.rs:140-164
 *   oracle_bsr_to_sorted          setops/src/bsr.rs:77-88
 *   oracle_bsr_count              setops/src/visitor.rs:399-403 (Counter::visit_bsr = popcount)
 *   oracle_svs                    setops/src/intersect/svs.rs:50-73 (svs_generic fold)
 *   oracle_small_adaptive         setops/src/intersect/adaptive.rs:107-147
 *
 * Values are compared as uint32_t.  The reference's harness feeds i32 (benchmark/src/datafile.rs:25);
 * for values below 2^31 -- every workload here -- the two orders coincide.
 *
 * Output convention (visitor.rs:87-91 VecWriter / visitor.rs:24-48 Counter): every function
 * writes the visited values in visit order into `out` and returns how many were visited.
 * `out` may be NULL, in which case only the count is returned (the Counter sink).
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>
#include <pthread.h>

#define EMIT(v)                                   \
    do {                                          \
        if (out) out[count] = (v);                \
        count++;                                  \
    } while (0)

/* merge.rs:7-33: three-way branch merge. */
size_t oracle_naive_merge(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    size_t i = 0, j = 0, count = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) {
            i++;
        } else if (a[i] > b[j]) {
            j++;
        } else {
            EMIT(a[i]);
            i++;
            j++;
        }
    }
    return count;
}

/* merge.rs:39-60: on equality emit and advance both; otherwise advance the smaller side by
 * the value of the comparison (no data-dependent branch on order). */
size_t oracle_branchless_merge(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    size_t i = 0, j = 0, count = 0;
    while (i < na && j < nb) {
        uint32_t x = a[i], y = b[j];
        if (x == y) {
            EMIT(x);
            i++;
            j++;
        } else {
            i += (size_t)(x < y);
            j += (size_t)(y < x);
        }
    }
    return count;
}

/* BSR emit: base + state, with a count of elements (popcount) for the Counter sink. */
typedef struct {
    uint32_t *bases;
    uint32_t *states;
    size_t words;
    uint64_t elements;
} bsr_sink;

static inline void bsr_emit(bsr_sink *s, uint32_t base, uint32_t state)
{
    if (s->bases) s->bases[s->words] = base;
    if (s->states) s->states[s->words] = state;
    s->words++;
    s->elements += (uint64_t)__builtin_popcount(state);
}

/* merge.rs:62-87: merge on bases; on equal base AND the states and emit when non-zero. */
size_t oracle_branchless_merge_bsr(const uint32_t *ba, const uint32_t *sa, size_t na,
                                   const uint32_t *bb, const uint32_t *sb, size_t nb,
                                   uint32_t *out_bases, uint32_t *out_states)
{
    bsr_sink s = {out_bases, out_states, 0, 0};
    size_t i = 0, j = 0;
    while (i < na && j < nb) {
        uint32_t x = ba[i], y = bb[j];
        if (x == y) {
            uint32_t st = sa[i] & sb[j];
            if (st != 0) bsr_emit(&s, x, st);
            i++;
            j++;
        } else {
            i += (size_t)(x < y);
            j += (size_t)(y < x);
        }
    }
    return s.words;
}

/* galloping.rs:110-131: inclusive-bounds binary search over set[lo..=hi]; returns the index of
 * the target when found, otherwise the insertion point `lo`. */
size_t oracle_binary_search(const uint32_t *set, uint32_t target, int64_t lo, int64_t hi)
{
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        uint32_t v = set[mid];
        if (v < target)
            lo = mid + 1;
        else if (v > target)
            hi = mid - 1;
        else
            return (size_t)mid;
    }
    return (size_t)lo;
}

/* galloping.rs:7-30: for each element of `small`, double an offset while large[offset] <= target,
 * binary-search the bracket [offset/2, min(offset, len-1)], emit on a hit, then drop the prefix of
 * `large` before the found position. */
size_t oracle_galloping(const uint32_t *small, size_t ns, const uint32_t *large, size_t nl, uint32_t *out)
{
    size_t count = 0;
    for (size_t k = 0; k < ns; k++) {
        uint32_t target = small[k];
        size_t offset = 1;
        while (offset < nl && large[offset] <= target) offset *= 2;
        int64_t lo = (int64_t)(offset / 2);
        int64_t hi = (int64_t)nl - 1;
        if ((int64_t)offset < hi) hi = (int64_t)offset;
        size_t base = oracle_binary_search(large, target, lo, hi);
        if (base < nl && large[base] == target) EMIT(target);
        large += base;
        nl -= base;
    }
    return count;
}

/* galloping.rs:32-49: plain binary search over the whole remaining `large` per element. */
size_t oracle_binary_search_intersect(const uint32_t *small, size_t ns, const uint32_t *large, size_t nl,
                                      uint32_t *out)
{
    size_t count = 0;
    for (size_t k = 0; k < ns; k++) {
        uint32_t target = small[k];
        size_t base = oracle_binary_search(large, target, 0, (int64_t)nl - 1);
        if (base < nl && large[base] == target) EMIT(target);
        large += base;
        nl -= base;
    }
    return count;
}

/* galloping.rs:51-76: galloping on the bases, AND of states on a base hit. */
size_t oracle_galloping_bsr(const uint32_t *ba, const uint32_t *sa, size_t ns,
                            const uint32_t *bb, const uint32_t *sb, size_t nl,
                            uint32_t *out_bases, uint32_t *out_states)
{
    bsr_sink s = {out_bases, out_states, 0, 0};
    for (size_t k = 0; k < ns; k++) {
        uint32_t target = ba[k];
        size_t offset = 1;
        while (offset < nl && bb[offset] <= target) offset *= 2;
        int64_t lo = (int64_t)(offset / 2);
        int64_t hi = (int64_t)nl - 1;
        if ((int64_t)offset < hi) hi = (int64_t)offset;
        size_t idx = oracle_binary_search(bb, target, lo, hi);
        if (idx < nl && bb[idx] == target) {
            uint32_t st = sa[k] & sb[idx];
            if (st != 0) bsr_emit(&s, target, st);
        }
        bb += idx;
        sb += idx;
        nl -= idx;
    }
    return s.words;
}

/* shuffling.rs:26-168: blocks of W from each side; a lane of the A block is emitted when it
 * equals any lane of the B block (the OR over W rotations); the block whose last element is not
 * larger advances (both on a tie); the remainder goes through branchless_merge. */
size_t oracle_shuffling(unsigned W, const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    size_t st_a = (na / W) * W, st_b = (nb / W) * W;
    size_t i = 0, j = 0, count = 0;
    while (i < st_a && j < st_b) {
        for (unsigned x = 0; x < W; x++) {
            int hit = 0;
            for (unsigned r = 0; r < W; r++) hit |= (a[i + x] == b[j + (x + r) % W]);
            if (hit) EMIT(a[i + x]);
        }
        uint32_t amax = a[i + W - 1], bmax = b[j + W - 1];
        i += W * (size_t)(amax <= bmax);
        j += W * (size_t)(bmax <= amax);
    }
    size_t tail = oracle_branchless_merge(a + i, na - i, b + j, nb - j, out ? out + count : NULL);
    return count + tail;
}

/* shuffling.rs:174-356: the same block walk on bases; a lane's state is the OR over rotations of
 * (base equal) & state_a & rotated state_b; emitted when a base matched and the state is non-zero. */
size_t oracle_shuffling_bsr(unsigned W, const uint32_t *ba, const uint32_t *sa, size_t na,
                            const uint32_t *bb, const uint32_t *sb, size_t nb,
                            uint32_t *out_bases, uint32_t *out_states)
{
    bsr_sink s = {out_bases, out_states, 0, 0};
    size_t st_a = (na / W) * W, st_b = (nb / W) * W;
    size_t i = 0, j = 0;
    while (i < st_a && j < st_b) {
        for (unsigned x = 0; x < W; x++) {
            int base_hit = 0;
            uint32_t st = 0;
            for (unsigned r = 0; r < W; r++) {
                unsigned y = (x + r) % W;
                int eq = ba[i + x] == bb[j + y];
                base_hit |= eq;
                st |= eq ? (sa[i + x] & sb[j + y]) : 0u;
            }
            if (base_hit && st != 0) bsr_emit(&s, ba[i + x], st);
        }
        uint32_t amax = ba[i + W - 1], bmax = bb[j + W - 1];
        i += W * (size_t)(amax <= bmax);
        j += W * (size_t)(bmax <= amax);
    }
    size_t tail = oracle_branchless_merge_bsr(ba + i, sa + i, na - i, bb + j, sb + j, nb - j,
                                              out_bases ? out_bases + s.words : NULL,
                                              out_states ? out_states + s.words : NULL);
    return s.words + tail;
}

/* broadcast.rs:21-160: the shuffling block walk, with each B lane of the block splatted and
 * compared against the whole A block (one compare vector per B lane) instead of rotations. */
size_t oracle_broadcast(unsigned W, const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    size_t st_a = (na / W) * W, st_b = (nb / W) * W;
    size_t i = 0, j = 0, count = 0;
    while (i < st_a && j < st_b) {
        unsigned mask = 0;
        for (unsigned y = 0; y < W; y++)
            for (unsigned x = 0; x < W; x++) mask |= (unsigned)(a[i + x] == b[j + y]) << x;
        for (unsigned x = 0; x < W; x++)
            if (mask >> x & 1u) EMIT(a[i + x]);
        uint32_t amax = a[i + W - 1], bmax = b[j + W - 1];
        i += W * (size_t)(amax <= bmax);
        j += W * (size_t)(bmax <= amax);
    }
    size_t tail = oracle_branchless_merge(a + i, na - i, b + j, nb - j, out ? out + count : NULL);
    return count + tail;
}

/* broadcast.rs:163-320: per splatted B lane, (base equal) & state_a & state_b, OR-ed over the B
 * lanes; an A lane is emitted when some base matched and the OR is non-zero. */
size_t oracle_broadcast_bsr(unsigned W, const uint32_t *ba, const uint32_t *sa, size_t na,
                            const uint32_t *bb, const uint32_t *sb, size_t nb,
                            uint32_t *out_bases, uint32_t *out_states)
{
    bsr_sink s = {out_bases, out_states, 0, 0};
    size_t st_a = (na / W) * W, st_b = (nb / W) * W;
    size_t i = 0, j = 0;
    uint32_t st[16];
    while (i < st_a && j < st_b) {
        unsigned base_mask = 0;
        for (unsigned x = 0; x < W; x++) st[x] = 0;
        for (unsigned y = 0; y < W; y++)
            for (unsigned x = 0; x < W; x++) {
                int eq = ba[i + x] == bb[j + y];
                base_mask |= (unsigned)eq << x;
                st[x] |= eq ? (sa[i + x] & sb[j + y]) : 0u;
            }
        for (unsigned x = 0; x < W; x++)
            if ((base_mask >> x & 1u) && st[x] != 0) bsr_emit(&s, ba[i + x], st[x]);
        uint32_t amax = ba[i + W - 1], bmax = bb[j + W - 1];
        i += W * (size_t)(amax <= bmax);
        j += W * (size_t)(bmax <= amax);
    }
    size_t tail = oracle_branchless_merge_bsr(ba + i, sa + i, na - i, bb + j, sb + j, nb - j,
                                              out_bases ? out_bases + s.words : NULL,
                                              out_states ? out_states + s.words : NULL);
    return s.words + tail;
}

/* simd_galloping.rs:196-243 (gallop_wide + binary_search_wide): blocks of `bound` elements of
 * large[0, nl), nl >= bound; the first full block whose last element is >= target, found by
 * doubling a block offset while the block after it ends below the target and binary-searching
 * [offset/2, min(offset, nl/bound - 1)].  Returns the last full block when none qualifies. */
static size_t sg_gallop_wide(uint32_t target, const uint32_t *large, size_t nl, size_t bound)
{
    size_t upper = 0;
    if (large[bound - 1] < target) {
        upper = 1;
        while ((upper + 1) * bound - 1 < nl && large[(upper + 1) * bound - 1] < target) upper *= 2;
    }
    size_t lo = upper / 2, hi = nl / bound - 1;
    if (upper < hi) hi = upper;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (large[(mid + 1) * bound - 1] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* simd_galloping.rs:49-107: Lemire's SIMD galloping with bound = 32 W.  The smaller set probes;
 * each target gallops over whole bound-sized blocks of the larger set; when no block can hold
 * it the larger set drops its full blocks and, if the probe side is at least a block long, the
 * two sides swap; else the block is narrowed to the quarter that can hold the target
 * (reduce_search_bound, :245-264) and its 8 W elements are compared (block_compare, :266-296).
 * Whatever is left (the larger side shorter than a block) goes through branchless_merge. */
size_t oracle_simd_galloping(unsigned W, const uint32_t *small, size_t ns, const uint32_t *large, size_t nl,
                             uint32_t *out)
{
    size_t count = 0;
    const size_t bound = 32 * (size_t)W;
    if (ns > nl) {
        const uint32_t *t = small; small = large; large = t;
        size_t tn = ns; ns = nl; nl = tn;
    }
    while (ns > 0 && nl >= bound) {
        uint32_t target = small[0];
        size_t blk = sg_gallop_wide(target, large, nl, bound);
        if (large[(blk + 1) * bound - 1] < target) {
            large += (blk + 1) * bound;
            nl -= (blk + 1) * bound;
            if (ns >= bound) {
                const uint32_t *t = small; small = large; large = t;
                size_t tn = ns; ns = nl; nl = tn;
                continue;
            }
            break;
        }
        large += blk * bound;
        nl -= blk * bound;
        size_t q;  /* first register (of W elements) of the quarter that can hold the target */
        if (large[bound / 2 - 1] >= target)
            q = (large[bound / 4 - 1] < target) ? 8 : 0;
        else
            q = (large[bound * 3 / 4 - 1] < target) ? 24 : 16;
        int hit = 0;
        for (size_t e = W * q; e < W * (q + 8); e++) hit |= large[e] == target;
        if (hit) EMIT(target);
        small++;
        ns--;
    }
    size_t tail = oracle_branchless_merge(small, ns, large, nl, out ? out + count : NULL);
    return count + tail;
}

/* simd_galloping.rs:109-193: the same walk on bases with bound = W (one register per block): the
 * first matching base of the found block gives the state, target_state & state, emitted non-zero;
 * branchless_merge_bsr for the rest. */
size_t oracle_simd_galloping_bsr(unsigned W, const uint32_t *ba, const uint32_t *sa, size_t ns,
                                 const uint32_t *bb, const uint32_t *sb, size_t nl,
                                 uint32_t *out_bases, uint32_t *out_states)
{
    bsr_sink s = {out_bases, out_states, 0, 0};
    const size_t bound = W;
    if (ns > nl) {
        const uint32_t *t = ba; ba = bb; bb = t;
        t = sa; sa = sb; sb = t;
        size_t tn = ns; ns = nl; nl = tn;
    }
    while (ns > 0 && nl >= bound) {
        uint32_t target = ba[0];
        size_t blk = sg_gallop_wide(target, bb, nl, bound);
        if (bb[(blk + 1) * bound - 1] < target) {
            bb += (blk + 1) * bound;
            sb += (blk + 1) * bound;
            nl -= (blk + 1) * bound;
            if (ns >= bound) {
                const uint32_t *t = ba; ba = bb; bb = t;
                t = sa; sa = sb; sb = t;
                size_t tn = ns; ns = nl; nl = tn;
                continue;
            }
            break;
        }
        bb += blk * bound;
        sb += blk * bound;
        nl -= blk * bound;
        for (size_t e = 0; e < bound; e++)
            if (bb[e] == target) {
                uint32_t st = sa[0] & sb[e];
                if (st != 0) bsr_emit(&s, target, st);
                break;
            }
        ba++;
        sa++;
        ns--;
    }
    size_t tail = oracle_branchless_merge_bsr(ba, sa, ns, bb, sb, nl,
                                              out_bases ? out_bases + s.words : NULL,
                                              out_states ? out_states + s.words : NULL);
    return s.words + tail;
}

/* qfilter.rs:254-357: classification of the 16-bit low-byte equality mask.  Bit 4*x + y is set
 * when the low byte of a[x] equals the low byte of b[y] (BYTE_CHECK_GROUP_A/B, instructions.rs:100-111).
 * Returns -2 (no lane has a candidate), -1 (some lane has several candidates) or the packed 2-bit
 * offsets of the single candidate of each lane (a lane without a candidate points at itself). */
#define QF_NO_MATCH (-2)
#define QF_MULTI_MATCH (-1)

static int qf_lane_offset(unsigned nibble)
{
    switch (nibble) {
    case 0: return QF_NO_MATCH;
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    default: return QF_MULTI_MATCH;
    }
}

int oracle_qfilter_classify(unsigned mask16)
{
    int off[4], any_multi = 0, all_none = 1, packed = 0;
    for (int x = 0; x < 4; x++) {
        off[x] = qf_lane_offset((mask16 >> (4 * x)) & 0xF);
        any_multi |= off[x] == QF_MULTI_MATCH;
        all_none &= off[x] == QF_NO_MATCH;
    }
    if (any_multi) return QF_MULTI_MATCH;
    if (all_none) return QF_NO_MATCH;
    for (int x = 0; x < 4; x++) packed |= (off[x] == QF_NO_MATCH ? x : off[x]) << (2 * x);
    return packed;
}

static unsigned qf_byte_mask(const uint32_t *va, const uint32_t *vb)
{
    unsigned m = 0;
    for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++)
            m |= (unsigned)((va[x] & 0xFF) == (vb[y] & 0xFF)) << (4 * x + y);
    return m;
}

/* qfilter.rs:32-89: 4-wide blocks; skip the block pair when the low-byte check finds nothing;
 * compare against the single shuffled candidate when every lane has at most one; otherwise the
 * all-pairs rotation compare.  Same advance rule and branchless_merge tail as shuffling. */
size_t oracle_qfilter(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    const size_t W = 4;
    size_t st_a = (na / W) * W, st_b = (nb / W) * W;
    size_t i = 0, j = 0, count = 0;
    while (i < st_a && j < st_b) {
        const uint32_t *va = a + i, *vb = b + j;
        int ms = oracle_qfilter_classify(qf_byte_mask(va, vb));
        if (ms != QF_NO_MATCH) {
            for (unsigned x = 0; x < 4; x++) {
                int hit = 0;
                if (ms >= 0) {
                    hit = va[x] == vb[(ms >> (2 * x)) & 3];
                } else {
                    for (unsigned y = 0; y < 4; y++) hit |= va[x] == vb[y];
                }
                if (hit) EMIT(va[x]);
            }
        }
        uint32_t amax = a[i + W - 1], bmax = b[j + W - 1];
        i += W * (size_t)(amax <= bmax);
        j += W * (size_t)(bmax <= amax);
    }
    size_t tail = oracle_branchless_merge(a + i, na - i, b + j, nb - j, out ? out + count : NULL);
    return count + tail;
}

/* qfilter.rs:92-180: the byte check runs on the bases; states are read only for candidate blocks;
 * single-candidate lanes AND with the shuffled state, multi-candidate blocks OR the masked ANDs
 * over the rotations; a lane is emitted when its base matched and its state is non-zero. */
size_t oracle_qfilter_bsr(const uint32_t *ba, const uint32_t *sa, size_t na,
                          const uint32_t *bb, const uint32_t *sb, size_t nb,
                          uint32_t *out_bases, uint32_t *out_states)
{
    const size_t W = 4;
    bsr_sink s = {out_bases, out_states, 0, 0};
    size_t st_a = (na / W) * W, st_b = (nb / W) * W;
    size_t i = 0, j = 0;
    while (i < st_a && j < st_b) {
        const uint32_t *va = ba + i, *vb = bb + j;
        int ms = oracle_qfilter_classify(qf_byte_mask(va, vb));
        if (ms != QF_NO_MATCH) {
            for (unsigned x = 0; x < 4; x++) {
                uint32_t st = 0;
                int base_hit = 0;
                if (ms >= 0) {
                    unsigned y = (ms >> (2 * x)) & 3;
                    base_hit = va[x] == vb[y];
                    st = sa[i + x] & sb[j + y];
                } else {
                    for (unsigned y = 0; y < 4; y++) {
                        int eq = va[x] == vb[y];
                        base_hit |= eq;
                        st |= eq ? (sa[i + x] & sb[j + y]) : 0u;
                    }
                }
                if (base_hit && st != 0) bsr_emit(&s, va[x], st);
            }
        }
        uint32_t amax = ba[i + W - 1], bmax = bb[j + W - 1];
        i += W * (size_t)(amax <= bmax);
        j += W * (size_t)(bmax <= amax);
    }
    size_t tail = oracle_branchless_merge_bsr(ba + i, sa + i, na - i, bb + j, sb + j, nb - j,
                                              out_bases ? out_bases + s.words : NULL,
                                              out_states ? out_states + s.words : NULL);
    return s.words + tail;
}

/* bsr.rs:140-164: base = v >> 5, state bit = 1 << (v & 31), one word per distinct base. */
size_t oracle_bsr_from_sorted(const uint32_t *v, size_t n, uint32_t *bases, uint32_t *states)
{
    size_t w = 0;
    for (size_t k = 0; k < n; k++) {
        uint32_t base = v[k] >> 5, bit = 1u << (v[k] & 31u);
        if (w > 0 && bases[w - 1] == base) {
            states[w - 1] |= bit;
        } else {
            bases[w] = base;
            states[w] = bit;
            w++;
        }
    }
    return w;
}

/* bsr.rs:77-88: expand each state's set bits in ascending order. */
size_t oracle_bsr_to_sorted(const uint32_t *bases, const uint32_t *states, size_t n, uint32_t *out)
{
    size_t count = 0;
    for (size_t k = 0; k < n; k++) {
        uint32_t high = bases[k] << 5, st = states[k];
        while (st) {
            EMIT(high | (uint32_t)__builtin_ctz(st));
            st &= st - 1;
        }
    }
    return count;
}

/* visitor.rs:399-403: the Counter sink counts popcount(state) per visited BSR word. */
uint64_t oracle_bsr_count(const uint32_t *states, size_t n)
{
    uint64_t c = 0;
    for (size_t k = 0; k < n; k++) c += (uint64_t)__builtin_popcount(states[k]);
    return c;
}

/* 2-set selector used by the k-set fold and the batch runner. */
enum {
    ORACLE_NAIVE_MERGE = 0,
    ORACLE_BRANCHLESS_MERGE = 1,
    ORACLE_GALLOPING = 2,
    ORACLE_BINARY_SEARCH = 3,
    ORACLE_SHUFFLING4 = 4,
    ORACLE_SHUFFLING8 = 5,
    ORACLE_SHUFFLING16 = 6,
    ORACLE_QFILTER = 7,
    ORACLE_BROADCAST4 = 8,
    ORACLE_SIMD_GALLOPING4 = 9,
    ORACLE_SIMD_GALLOPING16 = 10,
    /* x86 SIMD restatements (simd_x86.c), for the CPU baseline: callers check oracle_simd_level() */
    ORACLE_SHUFFLING_AVX512 = 11,
    ORACLE_BROADCAST_AVX512 = 12,
    ORACLE_SHUFFLING_AVX2 = 13,
};

size_t oracle_shuffling_avx512(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out);
size_t oracle_broadcast_avx512(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out);
size_t oracle_shuffling_avx2(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out);

size_t oracle_intersect2(int algo, const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    switch (algo) {
    case ORACLE_NAIVE_MERGE: return oracle_naive_merge(a, na, b, nb, out);
    case ORACLE_BRANCHLESS_MERGE: return oracle_branchless_merge(a, na, b, nb, out);
    case ORACLE_GALLOPING: return oracle_galloping(a, na, b, nb, out);
    case ORACLE_BINARY_SEARCH: return oracle_binary_search_intersect(a, na, b, nb, out);
    case ORACLE_SHUFFLING4: return oracle_shuffling(4, a, na, b, nb, out);
    case ORACLE_SHUFFLING8: return oracle_shuffling(8, a, na, b, nb, out);
    case ORACLE_SHUFFLING16: return oracle_shuffling(16, a, na, b, nb, out);
    case ORACLE_QFILTER: return oracle_qfilter(a, na, b, nb, out);
    case ORACLE_BROADCAST4: return oracle_broadcast(4, a, na, b, nb, out);
    case ORACLE_SIMD_GALLOPING4: return oracle_simd_galloping(4, a, na, b, nb, out);
    case ORACLE_SIMD_GALLOPING16: return oracle_simd_galloping(16, a, na, b, nb, out);
    case ORACLE_SHUFFLING_AVX512: return oracle_shuffling_avx512(a, na, b, nb, out);
    case ORACLE_BROADCAST_AVX512: return oracle_broadcast_avx512(a, na, b, nb, out);
    case ORACLE_SHUFFLING_AVX2: return oracle_shuffling_avx2(a, na, b, nb, out);
    default: return (size_t)-1;
    }
}

/* svs.rs:50-73: intersect sets[0] with sets[1] into `left`; for each further set swap the two
 * buffers and intersect the previous result with it.  No reordering: the caller orders the sets.
 * `left`/`right` need capacity for the largest intermediate (the harness uses the max set length,
 * benchmark/src/timer/harness.rs:157-164).  Returns the final length; the result is in *result. */
size_t oracle_svs(int algo, const uint32_t *const *sets, const size_t *lens, size_t k,
                  uint32_t *left, uint32_t *right, uint32_t **result)
{
    if (k < 2) return (size_t)-1;
    size_t n = oracle_intersect2(algo, sets[0], lens[0], sets[1], lens[1], left);
    for (size_t s = 2; s < k; s++) {
        uint32_t *t = left;
        left = right;
        right = t;
        n = oracle_intersect2(algo, right, n, sets[s], lens[s], left);
    }
    if (result) *result = left;
    return n;
}

/* adaptive.rs:107-147 (Demaine, Lopez-Ortiz, Munro): for each element of sets[0], in order, look it
 * up in sets[1], sets[2], ... in turn -- double an offset from the set's remembered position while
 * set[pos + offset] <= element, binary-search [pos, min(len-1, pos+offset)], remember the found
 * position -- and move to the next element at the first set that lacks it; an element found in
 * every set is emitted.  `pos` needs k entries of scratch. */
size_t oracle_small_adaptive(const uint32_t *const *sets, const size_t *lens, size_t k, size_t *pos, uint32_t *out)
{
    size_t count = 0;
    if (k < 2) return (size_t)-1;
    for (size_t i = 0; i < k; i++) pos[i] = 0;
    for (size_t e = 0; e < lens[0]; e++) {
        const uint32_t element = sets[0][e];
        int in_all = 1;
        for (size_t i = 1; i < k && in_all; i++) {
            const uint32_t *set = sets[i];
            const size_t len = lens[i], base = pos[i];
            size_t offset = 1;
            while (base + offset < len && set[base + offset] <= element) offset *= 2;
            int64_t hi = (int64_t)len - 1;
            if ((int64_t)(base + offset) < hi) hi = (int64_t)(base + offset);
            const size_t nb = oracle_binary_search(set, element, (int64_t)base, hi);
            pos[i] = nb;
            if (nb >= len || set[nb] != element) in_all = 0;
        }
        if (in_all) EMIT(element);
    }
    return count;
}

/* ------------------------------------------------------------------------------------------------
 * Batch runner: one intersection per host thread (the CPU baseline).  Pair p is
 * (A[a_beg[p]..a_end[p]), B[b_beg[p]..b_end[p])) -- CSR batches pass end = beg + 1.  Outputs go to a capacity-strided
 * buffer (pair p at cap_off[p], capacity min(|A|,|B|)), counts[p] receives the length.
 * Pairs are handed out dynamically in chunks of 64 so uneven sizes balance across threads.
 * -------------------------------------------------------------------------------------------- */
typedef struct {
    int algo;
    const uint32_t *A, *B;
    const uint64_t *a_beg, *a_end, *b_beg, *b_end, *cap_off;
    uint32_t *out;
    uint32_t *counts;
    size_t npairs;
    size_t next;
    pthread_mutex_t lock;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *job = (batch_job *)arg;
    for (;;) {
        pthread_mutex_lock(&job->lock);
        size_t p0 = job->next;
        job->next += 64;
        pthread_mutex_unlock(&job->lock);
        if (p0 >= job->npairs) break;
        size_t p1 = p0 + 64 < job->npairs ? p0 + 64 : job->npairs;
        for (size_t p = p0; p < p1; p++) {
            const uint32_t *a = job->A + job->a_beg[p], *b = job->B + job->b_beg[p];
            size_t na = job->a_end[p] - job->a_beg[p], nb = job->b_end[p] - job->b_beg[p];
            uint32_t *o = job->out ? job->out + job->cap_off[p] : NULL;
            job->counts[p] = (uint32_t)oracle_intersect2(job->algo, a, na, b, nb, o);
        }
    }
    return NULL;
}

/* Gather batch: pair p is (A[a_beg[p] .. a_end[p]), B[b_beg[p] .. b_end[p])).  Results of pair p go
 * to out + cap_off[p] (out may be NULL: counts only); pairs are shared out to `nthreads` threads in
 * chunks of 64. */
int oracle_batch_intersect_gather(int algo, int nthreads, const uint32_t *A, const uint64_t *a_beg,
                                  const uint64_t *a_end, const uint32_t *B, const uint64_t *b_beg,
                                  const uint64_t *b_end, size_t npairs, const uint64_t *cap_off, uint32_t *out,
                                  uint32_t *counts)
{
    batch_job job;
    job.algo = algo;
    job.A = A;
    job.B = B;
    job.a_beg = a_beg;
    job.a_end = a_end;
    job.b_beg = b_beg;
    job.b_end = b_end;
    job.cap_off = cap_off;
    job.out = out;
    job.counts = counts;
    job.npairs = npairs;
    job.next = 0;
    pthread_mutex_init(&job.lock, NULL);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) return -1;
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, batch_worker, &job);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&job.lock);
    return 0;
}

/* CSR batch: pair p is (A[a_off[p] .. a_off[p+1]), B[b_off[p] .. b_off[p+1])). */
int oracle_batch_intersect(int algo, int nthreads, const uint32_t *A, const uint64_t *a_off,
                           const uint32_t *B, const uint64_t *b_off, size_t npairs,
                           const uint64_t *cap_off, uint32_t *out, uint32_t *counts)
{
    return oracle_batch_intersect_gather(algo, nthreads, A, a_off, a_off + 1, B, b_off, b_off + 1, npairs, cap_off,
                                         out, counts);
}
