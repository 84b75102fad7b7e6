/* simd_x86.c -- x86 SIMD restatements of the reference's block-compare intersections, for the CPU
 * baseline (test infrastructure only, like the rest of oracle/).
 *
 *   oracle_shuffling_avx512   shuffling_avx512  (setops/src/intersect/shuffling.rs:116-168)
 *   oracle_broadcast_avx512   broadcast_avx512  (setops/src/intersect/broadcast.rs:109-160)
 *   oracle_shuffling_avx2     shuffling_avx2    (setops/src/intersect/shuffling.rs:68-112)
 *
 * Each walks both sets in blocks of W lanes (W = 16 with AVX-512F, 8 with AVX2): the A block is
 * compared with every rotation of the B block (shuffling) or with every B element broadcast
 * (broadcast), the matching A lanes are written in order (the visit_vectorW of an UnsafeWriter,
 * setops/src/visitor.rs:637-755: a compress-store with AVX-512, a permute by a 256-entry index
 * table with AVX2), and the block whose last element is not larger advances by W
 * (shuffling.rs:152-156).  The remainder goes through branchless_merge (merge.rs:39-60), as in the
 * reference.  Compares are on uint32 values (the reference compares i32; every workload here
 * keeps values below 2^31, where the orders agree).
 *
 * The functions carry target attributes, so the file builds with the default flags; callers pick
 * one with oracle_simd_level() (CPUID).  The batch runner in setops_oracle.c reaches them through
 * algorithm codes ORACLE_SHUFFLING_AVX512, ORACLE_BROADCAST_AVX512 and ORACLE_SHUFFLING_AVX2.
 *
 * oracle_stream_read times a STREAM-like read of a buffer by n threads: the host's DRAM read
 * bandwidth beside the CPU baseline.
 */
#include <immintrin.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

size_t oracle_branchless_merge(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out);

/* 2: AVX-512F available, 1: AVX2, 0: neither. */
int oracle_simd_level(void)
{
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f")) return 2;
    if (__builtin_cpu_supports("avx2")) return 1;
    return 0;
}

/* The block loop's tail: the rest of both sets by branchless_merge, appended at out + count (out
 * NULL: count only). */
static size_t tail_merge(const uint32_t *a, size_t na, size_t ia, const uint32_t *b, size_t nb, size_t ib,
                         uint32_t *out, size_t count)
{
    return count + oracle_branchless_merge(a + ia, na - ia, b + ib, nb - ib, out ? out + count : NULL);
}

__attribute__((target("avx512f"))) size_t oracle_shuffling_avx512(const uint32_t *a, size_t na, const uint32_t *b,
                                                                  size_t nb, uint32_t *out)
{
    const size_t W = 16, st_a = na / W * W, st_b = nb / W * W;
    size_t ia = 0, ib = 0, count = 0;
    while (ia < st_a && ib < st_b) {
        const __m512i va = _mm512_loadu_si512((const void *)(a + ia));
        const __m512i vb = _mm512_loadu_si512((const void *)(b + ib));
        __mmask16 m = _mm512_cmpeq_epi32_mask(va, vb);
        /* rotate_elements_left::<k> of the B block: lane i takes lane (i + k) mod 16 */
#define ROT(k) m |= _mm512_cmpeq_epi32_mask(va, _mm512_alignr_epi32(vb, vb, k))
        ROT(1); ROT(2); ROT(3); ROT(4); ROT(5); ROT(6); ROT(7); ROT(8);
        ROT(9); ROT(10); ROT(11); ROT(12); ROT(13); ROT(14); ROT(15);
#undef ROT
        if (out) _mm512_mask_compressstoreu_epi32((void *)(out + count), m, va);
        count += (size_t)__builtin_popcount((unsigned)m);
        const uint32_t amax = a[ia + W - 1], bmax = b[ib + W - 1];
        ia += W * (size_t)(amax <= bmax);
        ib += W * (size_t)(bmax <= amax);
    }
    return tail_merge(a, na, ia, b, nb, ib, out, count);
}

__attribute__((target("avx512f"))) size_t oracle_broadcast_avx512(const uint32_t *a, size_t na, const uint32_t *b,
                                                                  size_t nb, uint32_t *out)
{
    const size_t W = 16, st_a = na / W * W, st_b = nb / W * W;
    size_t ia = 0, ib = 0, count = 0;
    while (ia < st_a && ib < st_b) {
        const __m512i va = _mm512_loadu_si512((const void *)(a + ia));
        __mmask16 m = 0;
        for (size_t k = 0; k < W; k++) m |= _mm512_cmpeq_epi32_mask(va, _mm512_set1_epi32((int)b[ib + k]));
        if (out) _mm512_mask_compressstoreu_epi32((void *)(out + count), m, va);
        count += (size_t)__builtin_popcount((unsigned)m);
        const uint32_t amax = a[ia + W - 1], bmax = b[ib + W - 1];
        ia += W * (size_t)(amax <= bmax);
        ib += W * (size_t)(bmax <= amax);
    }
    return tail_merge(a, na, ia, b, nb, ib, out, count);
}

/* AVX2 compaction table: for each 8-bit mask, the lanes to keep, packed to the front. */
static uint32_t g_perm8[256][8];
static pthread_once_t g_perm8_once = PTHREAD_ONCE_INIT;

static void perm8_init(void)
{
    for (int m = 0; m < 256; m++) {
        int n = 0;
        for (int i = 0; i < 8; i++)
            if (m & (1 << i)) g_perm8[m][n++] = (uint32_t)i;
        for (; n < 8; n++) g_perm8[m][n] = 0;
    }
}

__attribute__((target("avx2"))) size_t oracle_shuffling_avx2(const uint32_t *a, size_t na, const uint32_t *b, size_t nb,
                                                             uint32_t *out)
{
    pthread_once(&g_perm8_once, perm8_init);
    const size_t W = 8, st_a = na / W * W, st_b = nb / W * W;
    size_t ia = 0, ib = 0, count = 0;
    while (ia < st_a && ib < st_b) {
        const __m256i va = _mm256_loadu_si256((const __m256i *)(a + ia));
        const __m256i vb = _mm256_loadu_si256((const __m256i *)(b + ib));
        __m256i eq = _mm256_cmpeq_epi32(va, vb);
        for (int k = 1; k < 8; k++) {
            const __m256i idx = _mm256_setr_epi32(k & 7, (k + 1) & 7, (k + 2) & 7, (k + 3) & 7, (k + 4) & 7, (k + 5) & 7,
                                                  (k + 6) & 7, (k + 7) & 7);
            eq = _mm256_or_si256(eq, _mm256_cmpeq_epi32(va, _mm256_permutevar8x32_epi32(vb, idx)));
        }
        const unsigned m = (unsigned)_mm256_movemask_ps(_mm256_castsi256_ps(eq));
        const unsigned n = (unsigned)__builtin_popcount(m);
        if (out && n) {
            /* the kept lanes packed to the front, then exactly n of them stored (a full 8-lane
             * store could run into another pair's output in the batch runner) */
            uint32_t tmp[8];
            const __m256i p = _mm256_loadu_si256((const __m256i *)g_perm8[m]);
            _mm256_storeu_si256((__m256i *)tmp, _mm256_permutevar8x32_epi32(va, p));
            for (unsigned i = 0; i < n; i++) out[count + i] = tmp[i];
        }
        count += n;
        const uint32_t amax = a[ia + W - 1], bmax = b[ib + W - 1];
        ia += W * (size_t)(amax <= bmax);
        ib += W * (size_t)(bmax <= amax);
    }
    return tail_merge(a, na, ia, b, nb, ib, out, count);
}

/* ---- host DRAM read bandwidth ------------------------------------------------------------------ */
typedef struct {
    const uint64_t *p;
    size_t n;
    int reps;
    uint64_t sum;
} stream_part;

static void *stream_worker(void *arg)
{
    stream_part *s = (stream_part *)arg;
    uint64_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    for (int r = 0; r < s->reps; r++) {
        size_t i = 0;
        for (; i + 4 <= s->n; i += 4) {
            acc0 += s->p[i];
            acc1 += s->p[i + 1];
            acc2 += s->p[i + 2];
            acc3 += s->p[i + 3];
        }
        for (; i < s->n; i++) acc0 += s->p[i];
    }
    s->sum = acc0 + acc1 + acc2 + acc3;
    return NULL;
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* Reads buf[0 .. n) `reps` times with `nthreads` threads (one contiguous share each); returns the
 * seconds taken and the checksum in *sum (so the reads cannot be dropped). */
double oracle_stream_read(const uint64_t *buf, size_t n, int nthreads, int reps, uint64_t *sum)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    stream_part *parts = (stream_part *)malloc(sizeof(stream_part) * (size_t)nthreads);
    if (!th || !parts) {
        free(th);
        free(parts);
        return -1.0;
    }
    const size_t share = (n + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; t++) {
        const size_t b = (size_t)t * share < n ? (size_t)t * share : n;
        const size_t e = b + share < n ? b + share : n;
        parts[t].p = buf + b;
        parts[t].n = e - b;
        parts[t].reps = reps;
        parts[t].sum = 0;
    }
    const double t0 = now_s();
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, stream_worker, &parts[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    const double el = now_s() - t0;
    uint64_t s = 0;
    for (int t = 0; t < nthreads; t++) s += parts[t].sum;
    if (sum) *sum = s;
    free(th);
    free(parts);
    return el;
}
