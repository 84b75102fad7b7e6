/*
 * setops_hip.h -- C ABI of the MI355X (gfx950) sorted-u32 intersection library.
 *
 * This is the drop-in boundary for the sorted-array intersection path of
 * UNSW-database/simd_set_operations.  Every entry point names the reference interface it
 * replaces (paths relative to that repository).  Conventions shared by all entry points:
 *
 *   - Sets are sorted, strictly increasing uint32_t arrays, or -- with SETOPS_SIGNED_I32 OR'ed into
 *     the algorithm id (plain entries) -- sorted, strictly increasing int32_t arrays passed through
 *     the same pointers, compared as signed.  The reference harness passes i32
 *     (benchmark/src/timer.rs:19, benchmark/src/datafile.rs:25) and its property tests draw
 *     negative values (setops/tests/testlib/mod.rs:11-30); its BSR form casts them to u32 first
 *     (benchmark/src/timer/harness.rs:121-122), so BSR entries have no int32 mode.
 *   - Results are the intersection in ascending order, identical element for element to the
 *     reference's branchless_merge (setops/src/intersect/merge.rs:39-60) and svs_generic
 *     (setops/src/intersect/svs.rs:50-73).
 *   - Return value: SETOPS_OK (0) or a negative SETOPS_ERR_* code; setops_hip_last_error()
 *     gives a message for the calling thread.  The reference has no error channel
 *     (setops/ffi/qfilter/qfilter.h:8-39 return a bare count); misuse that the reference leaves
 *     undefined (short output buffer, unsorted sizes) is reported here instead.
 *   - Output capacity follows the reference harness: min(|A|,|B|) for a 2-set call
 *     (benchmark/src/timer/harness.rs:89), the smallest set for a k-set call.
 *   - "Host" entry points take host pointers and are synchronous, one GPU launch sequence per
 *     call -- the plain per-pair drop-in.  "Batch" and "_dev" entry points take device pointers,
 *     run on the given HIP stream (NULL = the HIP null stream) and return without synchronising.
 */
#ifndef SETOPS_HIP_H
#define SETOPS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SETOPS_OK 0
#define SETOPS_ERR_INVALID (-1)      /* null pointer, bad size, misaligned pool, bad algorithm use */
#define SETOPS_ERR_CAPACITY (-2)     /* output capacity too small for the result */
#define SETOPS_ERR_UNKNOWN_ALGO (-3) /* algorithm id / name not registered */
#define SETOPS_ERR_HIP (-4)          /* HIP runtime error (message names the call) */
#define SETOPS_ERR_NOMEM (-5)        /* device allocation failed or budget exceeded */
#define SETOPS_ERR_UNSUPPORTED (-6)  /* operation not available for this algorithm */
#define SETOPS_ERR_NO_DEVICE (-7)    /* no usable gfx950 device */

/* ---- library state ----------------------------------------------------------------------- */

/* Selects and initialises the device (idempotent).  Every other call initialises device 0 on
 * first use if this was not called.  No reference counterpart (the reference is CPU-only). */
int setops_hip_init(int device);

/* Message for the last error on the calling thread ("" if none). */
const char *setops_hip_last_error(void);

/* Library version string. */
const char *setops_hip_version(void);

/* ---- algorithm registry (benchmark/src/timer.rs:75-276: name -> function table) ---------- */

#define SETOPS_ALGO_TWOSET 1u  /* plain 2-set entry (timer.rs:75-169 try_parse_twoset) */
#define SETOPS_ALGO_BSR 2u     /* BSR 2-set entry (timer.rs:211-259 try_parse_bsr) */
#define SETOPS_ALGO_KSET 4u    /* usable for k sets through svs_generic (timer.rs:190-198) */

/* OR'ed into the id of a plain (non-BSR) entry: the sets are int32_t, ordered as signed integers.
 * Every entry point that takes an algorithm id accepts it; results are the same int32 values. */
#define SETOPS_SIGNED_I32 0x10000

/* OR'ed into an algorithm id: the tile shape of the call's kernels (0 = the default, also when
 * no shape is given).  The plain branchless_merge entry has shapes 0-3 (256 lanes x 20, 16, 24, 12
 * merge steps per lane), the plain search entries (galloping, binary_search, simd_galloping,
 * small_adaptive) shapes 0-2 (256 lanes x 2, 1, 4 probes per lane); every other entry shape 0
 * only (SETOPS_ERR_UNSUPPORTED otherwise).  Results do not depend on the shape; the autotuner
 * (simd_set_operations_amd/autotune.py) times them.  No reference counterpart. */
#define SETOPS_SHAPE_SHIFT 20
#define SETOPS_SHAPE_MASK (0xF << SETOPS_SHAPE_SHIFT)
#define SETOPS_SHAPE(i) ((i) << SETOPS_SHAPE_SHIFT)

/* Number of registered entries; ids are 0 .. count-1. */
int setops_hip_algo_count(void);
/* Id of a registered name (e.g. "hip_branchless_merge"), or SETOPS_ERR_UNKNOWN_ALGO. */
int setops_hip_algo_lookup(const char *name);
/* Name of an id, or NULL. */
const char *setops_hip_algo_name(int algo);
/* SETOPS_ALGO_* flags of an id, or 0. */
unsigned setops_hip_algo_flags(int algo);

/* ---- plain 2-set, host pointers --------------------------------------------------------
 * Replaces Intersect2C (setops/src/intersect.rs:41) as bound for qfilter_c
 * (setops/src/intersect/qfilter_c.rs:6-22 -> setops/ffi/qfilter/qfilter.cpp:221
 * intersect_qfilter_uint_b4_v2).  Writes A∩B ascending to out[0 .. *out_len).  out may be NULL
 * for a count-only call (the Counter visitor, setops/src/visitor.rs:24-48); otherwise out_cap
 * must be >= the result size (min(na, nb) always suffices). */
int setops_hip_intersect2(int algo, const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out,
                          size_t out_cap, size_t *out_len);

/* ---- 2-set (plain or BSR), device pointers, asynchronous --------------------------------
 * The plain one-launch-per-pair drop-in for device-resident sets: the same contract as
 * setops_hip_intersect2 / setops_hip_intersect2_bsr, but the inputs (16-byte aligned) and outputs
 * are device pointers and the call returns after enqueueing on `stream` (NULL = null stream).
 * d_totals[0] receives the result length (words for BSR) and d_totals[1] the element count.
 * d_a_states / d_b_states / d_out_states are NULL for plain entries.  d_out (and d_out_states)
 * NULL selects count-only.  Calls share one workspace: keep them on one stream. */
int setops_hip_intersect2_dev(int algo, const uint32_t *d_a, const uint32_t *d_a_states, size_t na, const uint32_t *d_b,
                              const uint32_t *d_b_states, size_t nb, uint32_t *d_out, uint32_t *d_out_states,
                              size_t out_cap, unsigned long long *d_totals, void *stream);

/* ---- 2-set, device pointers, no allocation -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  The same arguments,
 * contract and results as setops_hip_intersect2_dev, but the call allocates nothing (no hipMalloc,
 * hipFree or allocating library call): its plan and result staging live in a fixed 20 MiB device
 * array of the library, one per GPU, held until the process ends.  Limits: na + nb <= 2^22 for the
 * merge-path entries (branchless_merge, shuffling, broadcast, qfilter and their BSR twins),
 * min(na, nb) <= 2^20 for the search entries (galloping, binary_search, simd_galloping,
 * small_adaptive and their BSR twins); larger pairs return SETOPS_ERR_INVALID and enqueue nothing.
 * Every call on a GPU shares that GPU's array and the array keeps the last call's contents: calls
 * that overlap on one GPU (on other streams or from other threads) race, so a caller serialises
 * them (one stream, or a wait between calls). */
int setops_hip_intersect2_dev_noalloc(int algo, const uint32_t *d_a, const uint32_t *d_a_states, size_t na,
                                      const uint32_t *d_b, const uint32_t *d_b_states, size_t nb, uint32_t *d_out,
                                      uint32_t *d_out_states, size_t out_cap, unsigned long long *d_totals,
                                      void *stream);

/* ---- BSR 2-set, host pointers ------------------------------------------------------------
 * Replaces Intersect2Bsr (setops/src/bsr.rs:16) / UnsafeIntersectBsr
 * (benchmark/src/timer/harness.rs:17) and intersect_qfilter_bsr_b4_v2 (qfilter.h:24-27).
 * Inputs are BsrVec layouts (bsr.rs:48-75): bases strictly increasing, states non-zero.  Output is
 * the (base, state_a & state_b) words with non-zero state, bases ascending; *out_words is their
 * number and *out_elems (may be NULL) the popcount total (Counter::visit_bsr, visitor.rs:399-403).
 * bases_out/states_out may both be NULL for a count-only call. */
int setops_hip_intersect2_bsr(int algo, const uint32_t *bases_a, const uint32_t *states_a, size_t na,
                              const uint32_t *bases_b, const uint32_t *states_b, size_t nb, uint32_t *bases_out,
                              uint32_t *states_out, size_t out_cap, size_t *out_words, uint64_t *out_elems);

/* ---- k-set, host pointers ------------------------------------------------------------------
 * Replaces svs_generic (setops/src/intersect/svs.rs:50-73) driven by a 2-set entry, as timed by
 * time_svs (benchmark/src/timer/harness.rs:152-172): sets[0] ∩ sets[1], then the running result
 * ∩ sets[2], ... in the order given (the caller orders smallest-first).  k >= 2.
 * With "hip_small_adaptive" the same call replaces small_adaptive (setops/src/intersect/
 * adaptive.rs:107-147, the "small_adaptive" k-set entry of benchmark/src/timer.rs): every element
 * of sets[0] is looked up in sets[1], sets[2], ... in turn, in one launch per 1 + 8 further sets,
 * with no intermediate result between the sets of a launch. */
int setops_hip_svs(int algo, const uint32_t *const *sets, const size_t *lens, size_t k, uint32_t *out,
                   size_t out_cap, size_t *out_len);

/* ---- int32 forms of the host-pointer calls -------------------------------------------------
 * setops_hip_intersect2 and setops_hip_svs with SETOPS_SIGNED_I32, with int32_t pointers: the
 * shapes the reference's harness times, Intersect2C<[i32]> (setops/src/intersect.rs:41,
 * time_twoset_c in benchmark/src/timer/harness.rs:98-112) and the i32 k-set fold of time_svs
 * (harness.rs:152-172, svs_generic of setops/src/intersect/svs.rs:50-73). */
int setops_hip_intersect2_i32(int algo, const int32_t *a, size_t na, const int32_t *b, size_t nb, int32_t *out,
                              size_t out_cap, size_t *out_len);
int setops_hip_svs_i32(int algo, const int32_t *const *sets, const size_t *lens, size_t k, int32_t *out, size_t out_cap,
                       size_t *out_len);

/* ---- device-resident batches ----------------------------------------------------------------
 * A batch is P pairs in CSR form: pair p is (A[a_off[p] .. a_off[p+1]), B[b_off[p] .. b_off[p+1])),
 * with a_off[0] == b_off[0] == 0, a_off[P] == a_len, b_off[P] == b_len.  A, B (and the BSR
 * state arrays) must be 16-byte aligned device pointers.  There is no single reference
 * counterpart: this is the batched form of Intersect2 / Intersect2Bsr that the reference's
 * benchmark loop calls once per datafile pair (benchmark/src/bin/benchmark.rs:180-243).
 *
 * Plain batches pass NULL for a_states/b_states; BSR batches pass the state arrays and a BSR
 * algorithm.  Planning depends only on the sizes, so one handle can be run many times. */
typedef struct setops_hip_batch setops_hip_batch;

int setops_hip_batch_create(setops_hip_batch **handle, int algo, const uint32_t *d_a, const uint32_t *d_a_states,
                            uint64_t a_len, const uint64_t *d_a_off, const uint32_t *d_b, const uint32_t *d_b_states,
                            uint64_t b_len, const uint64_t *d_b_off, uint64_t npairs);

/* Gather form: pair p is (A[a_beg[p] .. a_end[p]), B[b_beg[p] .. b_end[p])) anywhere in the pools
 * (pairs may overlap and share sets: graph edges over one CSR, queries over one corpus).  Each
 * set must hold < 2^32 elements.  Creation counts the batch's work on the device and waits for it
 * (one synchronisation); runs do not synchronise.  Same outputs and algorithms as the CSR form;
 * the reference's harness builds such pairs from datafile sets, benchmark/src/timer/harness.rs:81-172. */
int setops_hip_batch_create_gather(setops_hip_batch **handle, int algo, const uint32_t *d_a,
                                   const uint32_t *d_a_states, uint64_t a_len, const uint64_t *d_a_beg,
                                   const uint64_t *d_a_end, const uint32_t *d_b, const uint32_t *d_b_states,
                                   uint64_t b_len, const uint64_t *d_b_beg, const uint64_t *d_b_end, uint64_t npairs);

/* Builds an index of the batch's B pool (every 32nd element, 1/32 of the pool, kept until the
 * handle is destroyed) and lets the search-policy entries (hip_galloping, hip_binary_search,
 * hip_simd_galloping and their BSR twins) use it for larger sets that do not fit on chip: a probe then reads about one
 * 128-byte line of B instead of one per binary-search level.  For resident sets that many pairs
 * share (a large set against many small ones, a corpus).  The B pool's contents must not change
 * afterwards where pairs read them.  Waits for the index build.  No-op for merge-policy entries.
 * No reference counterpart (the reference searches host memory). */
int setops_hip_batch_index_b(setops_hip_batch *handle);

/* k-set groups in one batch: pair p of a gather batch also needs each of its further sets
 * B[x_beg[j] .. x_end[j]), j in [x_off[p], x_off[p+1]) -- the group (A_p, B_p, further sets) is one
 * k-set query, answered in the batch's one launch sequence with no intermediate results: every
 * element of the pair's intersection is looked up in the further sets in turn, the first that lacks
 * it ending its lookup -- small_adaptive's walk (setops/src/intersect/adaptive.rs:107-147), and
 * the result svs_generic's (setops/src/intersect/svs.rs:50-73), the ascending k-way intersection.
 * d_x_off ([P+1], x_off[0] = 0), d_x_beg and d_x_end are device arrays of pool indexes into B, read
 * at run time (their contents may change between runs; a rebound batch keeps them).  For the search
 * entries on plain sets (hip_small_adaptive, hip_galloping, hip_binary_search, hip_simd_galloping);
 * with an index of B (setops_hip_batch_index_b) a lookup reads the index and one 128-byte line.
 * All three NULL detaches the further sets.  The reference runs k-set queries one at a time
 * (benchmark/src/timer/harness.rs:152-172). */
int setops_hip_batch_set_further(setops_hip_batch *handle, const uint64_t *d_x_off, const uint64_t *d_x_beg,
                                 const uint64_t *d_x_end);

/* Points a gather batch at new pairs (same pools and algorithm), reusing its device workspace
 * (grow-only: no allocation once it is large enough).  h_na / h_nb are the pairs' sizes on the host
 * -- they must equal d_a_end - d_a_beg and d_b_end - d_b_beg -- so nothing is read back from the
 * device.  For request loops that batch different pairs every step (continuous batching over a
 * resident corpus); the reference has no counterpart.  The previous run must be complete or on the
 * same stream as the next one. */
int setops_hip_batch_rebind(setops_hip_batch *handle, const uint64_t *d_a_beg, const uint64_t *d_a_end,
                            const uint64_t *d_b_beg, const uint64_t *d_b_end, uint64_t npairs, const uint64_t *h_na,
                            const uint64_t *h_nb);

/* Grows the handle's device workspace and result staging ahead of time for up to `npairs` pairs
 * and `units` work units (merge entries: |A| + |B| per pair, search entries: min(|A|, |B|)), so
 * that later rebinds and runs within those sizes allocate nothing; growing during a run frees and
 * reallocates device memory, which waits for the device.  For request loops that rebind one
 * handle every step; the reference has no counterpart.  The previous run must be complete or on
 * the handle's last stream. */
int setops_hip_batch_reserve(setops_hip_batch *handle, uint64_t npairs, uint64_t units);

/* One request-loop step on a gather handle, enqueued on `stream` with one call (nothing waits):
 * copies the packed step arrays h_idx -> d_idx (uint64: a_beg[P], a_end[P], b_beg[P], b_end[P],
 * then, with nfurther >= 0, x_off[P+1], x_beg[nfurther], x_end[nfurther]; h_idx in pinned host
 * memory; h_idx NULL: no copy, d_idx already holds them, e.g. host memory mapped for the device),
 * rebinds the handle to them (h_na / h_nb: the pairs' sizes, as for
 * setops_hip_batch_rebind), attaches the further sets (nfurther < 0: none), runs it with outputs
 * (as setops_hip_batch_run), then enqueues the copies of d_out_off[0 .. P] to h_out_off (none if
 * h_out_off is NULL, e.g. for offsets in mapped host memory) and of d_out[0 .. copy_words) to
 * h_out.  For serving loops that answer new queries every step; the
 * reference has no counterpart. */
int setops_hip_batch_step(setops_hip_batch *handle, const uint64_t *h_idx, uint64_t *d_idx, uint64_t npairs,
                          int64_t nfurther, const uint64_t *h_na, const uint64_t *h_nb, uint32_t *d_out,
                          uint64_t out_cap, uint64_t *d_out_off, uint32_t *d_counts, unsigned long long *d_totals,
                          uint64_t *h_out_off, uint32_t *h_out, uint64_t copy_words, void *stream);

/* One step slot of the native request loop: a gather handle over the corpus pool, the words of
 * packed rows a step may hold (idx_words, setops_hip_batch_step's layout), its totals (d_totals,
 * 2 words on the device), and the result words a step may hold (out_words).  A step's rows are read by its kernels, and its offsets and results written by
 * them, in host memory that the loop maps for the device (buffers per slot, allocated when the
 * loop starts): there is no upload and no copy back. */
typedef struct setops_hip_serve_slot {
    setops_hip_batch *handle;
    uint64_t idx_words;
    uint64_t out_words;
    unsigned long long *d_totals;
    uint64_t max_pairs;
    void *stream;
} setops_hip_serve_slot;

/* A stream of k-set queries, in arrival order.  Query q is the pair (A, B) =
 * ([pairs[q], pairs[nq + q]), [pairs[2nq + q], pairs[3nq + q])) of pool offsets with sizes na[q],
 * nb[q], and further sets [xb[j], xe[j]) for j in [xoff[q], xoff[q+1]).  Its result needs at most
 * need_pfx[q+1] - need_pfx[q] words; it holds cost[q] words of the admission budget while in
 * flight; it arrives arrivals[q] seconds after the loop starts (ascending). */
typedef struct setops_hip_serve_queries {
    uint64_t nq;
    const uint64_t *pairs;
    const uint64_t *na;
    const uint64_t *nb;
    const int64_t *xoff;
    const uint64_t *xb;
    const uint64_t *xe;
    const int64_t *need_pfx;
    const int64_t *cost;
    const double *arrivals;
} setops_hip_serve_queries;

/* The request loop in native code: serves the query stream with the given step slots (two or more;
 * a step on one runs while the host admits and launches the next on another), admitting
 * arrived queries in arrival order while their costs fit `budget_words` and a step holds at most
 * max_pairs queries, and a step's packed rows and results fit its slot.  Each query is answered
 * completely in one step (setops_hip_batch_step with its further sets).  Outputs: done_t[q] =
 * seconds from the start when query q's step was complete on the host (NaN if the loop stopped
 * at timeout_s first); with results != NULL, query q's result at results[need_pfx[q] ..] and its
 * size in counts[q]; stats = {steps, queries launched, queries served, wall time in ns}.  A query that fits no
 * slot's buffers is SETOPS_ERR_CAPACITY before anything runs.  For serving loops whose host
 * side should cost microseconds per step; the reference has no counterpart (the nearest caller
 * is its k-set timer, harness.rs:152-172). */
int setops_hip_serve(const setops_hip_serve_slot *slots, int nslots, const setops_hip_serve_queries *queries,
                     int64_t budget_words, uint64_t max_pairs, double timeout_s, double *done_t, uint32_t *results,
                     uint64_t *counts, uint64_t *stats);

/* Runs the batch on `stream` (a hipStream_t, NULL for the HIP null stream).  Outputs:
 *   d_out / d_out_states : compacted results of all pairs in pair order (states: BSR only);
 *                          both NULL selects count-only mode (out_cap ignored)
 *   d_out_off            : [P+1] u64, pair p's results start at d_out_off[p]; d_out_off[P] = total
 *   d_counts             : [P] u32 result count of each pair (words for BSR)
 *   d_totals             : [2] u64 totals (words, elements) -- elements = popcount sum for BSR
 * Any output pointer may be NULL.  If the results exceed out_cap the writes are clipped and
 * setops_hip_batch_check reports SETOPS_ERR_CAPACITY. */
int setops_hip_batch_run(setops_hip_batch *handle, uint32_t *d_out, uint32_t *d_out_states, uint64_t out_cap,
                         uint64_t *d_out_off, uint32_t *d_counts, unsigned long long *d_totals, void *stream);

/* Synchronises the handle's last run and reports SETOPS_ERR_CAPACITY if its output was clipped. */
int setops_hip_batch_check(setops_hip_batch *handle);

/* Enables per-run HIP event timing of the main intersection kernel (0 = off). */
int setops_hip_batch_set_timing(setops_hip_batch *handle, int enabled);

/* Synchronises and returns the summed main-kernel time (ms) and number of timed runs since the
 * last call; resets the accumulator. */
int setops_hip_batch_kernel_time(setops_hip_batch *handle, double *total_ms, uint64_t *runs);

/* Number of tiles, the tile kernel's grid size (one workgroup per tile) and the unit total of the
 * handle's plan (for reporting). */
int setops_hip_batch_info(setops_hip_batch *handle, uint64_t *ntiles, uint64_t *grid, uint64_t *units);

int setops_hip_batch_destroy(setops_hip_batch *handle);

#ifdef __cplusplus
}
#endif

#endif /* SETOPS_HIP_H */
