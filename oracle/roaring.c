/*
 * roaring.c -- CPU restatement of the croaring comparator of the reference's benchmark.
 *
 * TEST INFRASTRUCTURE ONLY, like setops_oracle.c: tests/ and bench.py's CPU baseline leg load it;
 * the HIP library never does.
 *
 * The reference times CRoaring through its Rust bindings as a comparison baseline
 * (benchmark/src/timer/harness.rs:194-257 time_croaring_2set / time_croaring_svs, registered as
 * "croaring" and "croaring_opt" in benchmark/src/timer.rs:280-293; dependency croaring = "0.9.0",
 * benchmark/Cargo.toml:21).  That library is not part of the reference checkout and is not
 * installed here, so this file restates the published Roaring design (Chambi, Lemire, Kaser,
 * Godin, "Better bitmap performance with Roaring bitmaps", 2016; Lemire et al., "Roaring bitmaps:
 * implementation of an optimized software library", 2018) for the operations the harness uses:
 *
 *   - Bitmap::of(values): values are split by their high 16 bits into containers kept in key
 *     order; a container holds the low 16 bits of its values as a sorted uint16 array while it
 *     has at most 4096 of them, else as a 65536-bit bitmap (1024 uint64 words);
 *   - and_cardinality / and: containers with equal keys are intersected -- array & array by a
 *     merge, or by galloping the smaller through the larger when it is at least 64 times smaller;
 *     array & bitmap by testing each array value's bit; bitmap & bitmap by word ANDs (and
 *     popcounts) -- the result container is an array when it holds at most 4096 values;
 *   - run_optimize ("croaring_opt") converts containers into run containers where that is
 *     smaller.  Run containers are not restated: they change the representation, never a
 *     result, so the comparator's outputs are those of "croaring" either way.
 *
 * The results are the sorted intersection (decoded in ascending order) and its cardinality.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

#define RB_ARRAY_MAX 4096
#define RB_WORDS 1024

typedef struct {
    uint16_t key;
    uint8_t bitmap;  /* 0: array of card uint16, 1: bitmap of RB_WORDS uint64 */
    uint32_t card;
    void *data;
} rb_container;

typedef struct {
    uint32_t n;
    rb_container *c;
} rb_bitmap;

static void rb_free(rb_bitmap *b)
{
    for (uint32_t i = 0; i < b->n; i++) free(b->c[i].data);
    free(b->c);
    b->c = NULL;
    b->n = 0;
}

/* Bitmap::of over sorted, strictly increasing values. */
static int rb_from_sorted(const uint32_t *v, size_t n, rb_bitmap *out)
{
    out->n = 0;
    out->c = NULL;
    if (n == 0) return 0;
    uint32_t keys = 1;
    for (size_t i = 1; i < n; i++) keys += (v[i] >> 16) != (v[i - 1] >> 16);
    out->c = (rb_container *)calloc(keys, sizeof(rb_container));
    if (!out->c) return -1;
    size_t i = 0;
    while (i < n) {
        const uint16_t key = (uint16_t)(v[i] >> 16);
        size_t j = i;
        while (j < n && (v[j] >> 16) == key) j++;
        rb_container *c = &out->c[out->n++];
        c->key = key;
        c->card = (uint32_t)(j - i);
        if (c->card <= RB_ARRAY_MAX) {
            uint16_t *a = (uint16_t *)malloc(sizeof(uint16_t) * c->card);
            if (!a) return -1;
            for (size_t k = i; k < j; k++) a[k - i] = (uint16_t)v[k];
            c->data = a;
        } else {
            uint64_t *w = (uint64_t *)calloc(RB_WORDS, sizeof(uint64_t));
            if (!w) return -1;
            for (size_t k = i; k < j; k++) w[(v[k] & 0xFFFFu) >> 6] |= 1ull << (v[k] & 63u);
            c->bitmap = 1;
            c->data = w;
        }
        i = j;
    }
    return 0;
}

/* Galloping lower bound of x in a[lo, n): first index with a[i] >= x. */
static size_t rb_gallop(const uint16_t *a, size_t lo, size_t n, uint16_t x)
{
    if (lo >= n || a[lo] >= x) return lo;
    size_t step = 1, prev = lo, cur = lo + 1;
    while (cur < n && a[cur] < x) {
        prev = cur;
        step <<= 1;
        cur = lo + step;
    }
    size_t hi = cur < n ? cur : n;
    lo = prev + 1;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* array & array into out (may be NULL); returns the count. */
static uint32_t rb_and_arrays(const uint16_t *a, uint32_t na, const uint16_t *b, uint32_t nb, uint16_t *out)
{
    uint32_t cnt = 0;
    if (na > nb) {
        const uint16_t *t = a; a = b; b = t;
        uint32_t tn = na; na = nb; nb = tn;
    }
    if ((uint64_t)na * 64 <= nb) {  /* skewed: gallop the smaller through the larger */
        size_t pos = 0;
        for (uint32_t i = 0; i < na; i++) {
            pos = rb_gallop(b, pos, nb, a[i]);
            if (pos == nb) break;
            if (b[pos] == a[i]) {
                if (out) out[cnt] = a[i];
                cnt++;
            }
        }
        return cnt;
    }
    uint32_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) i++;
        else if (a[i] > b[j]) j++;
        else {
            if (out) out[cnt] = a[i];
            cnt++;
            i++;
            j++;
        }
    }
    return cnt;
}

static uint32_t rb_and_array_bitmap(const uint16_t *a, uint32_t na, const uint64_t *w, uint16_t *out)
{
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < na; i++)
        if (w[a[i] >> 6] >> (a[i] & 63u) & 1u) {
            if (out) out[cnt] = a[i];
            cnt++;
        }
    return cnt;
}

/* Intersection of two bitmaps.  out_vals (may be NULL) receives the decoded sorted values; the
 * result is also returned as a roaring bitmap in res when res is not NULL (and_inplace). */
static uint64_t rb_and(const rb_bitmap *a, const rb_bitmap *b, uint32_t *out_vals, rb_bitmap *res)
{
    uint64_t total = 0;
    uint32_t i = 0, j = 0;
    uint16_t tmp[RB_ARRAY_MAX];
    uint64_t wtmp[RB_WORDS];
    if (res) {
        res->n = 0;
        uint32_t m = a->n < b->n ? a->n : b->n;
        res->c = m ? (rb_container *)calloc(m, sizeof(rb_container)) : NULL;
    }
    while (i < a->n && j < b->n) {
        const rb_container *x = &a->c[i], *y = &b->c[j];
        if (x->key < y->key) { i++; continue; }
        if (x->key > y->key) { j++; continue; }
        const int need = out_vals != NULL || res != NULL;
        uint32_t card = 0;
        int as_bitmap = 0;
        if (!x->bitmap || !y->bitmap) {
            if (!x->bitmap && !y->bitmap)
                card = rb_and_arrays((const uint16_t *)x->data, x->card, (const uint16_t *)y->data, y->card,
                                     need ? tmp : NULL);
            else if (!x->bitmap)
                card = rb_and_array_bitmap((const uint16_t *)x->data, x->card, (const uint64_t *)y->data,
                                           need ? tmp : NULL);
            else
                card = rb_and_array_bitmap((const uint16_t *)y->data, y->card, (const uint64_t *)x->data,
                                           need ? tmp : NULL);
        } else {
            const uint64_t *wx = (const uint64_t *)x->data, *wy = (const uint64_t *)y->data;
            if (need) {
                for (int k = 0; k < RB_WORDS; k++) {
                    wtmp[k] = wx[k] & wy[k];
                    card += (uint32_t)__builtin_popcountll(wtmp[k]);
                }
                as_bitmap = card > RB_ARRAY_MAX;
                if (!as_bitmap) {  /* small results become an array container */
                    uint32_t c = 0;
                    for (int k = 0; k < RB_WORDS; k++)
                        for (uint64_t t = wtmp[k]; t; t &= t - 1)
                            tmp[c++] = (uint16_t)(64 * k + __builtin_ctzll(t));
                }
            } else {
                for (int k = 0; k < RB_WORDS; k++) card += (uint32_t)__builtin_popcountll(wx[k] & wy[k]);
            }
        }
        if (card && out_vals) {
            uint32_t *o = out_vals + total;
            const uint32_t hi = (uint32_t)x->key << 16;
            if (as_bitmap) {
                uint32_t c = 0;
                for (int k = 0; k < RB_WORDS; k++)
                    for (uint64_t t = wtmp[k]; t; t &= t - 1) o[c++] = hi | (uint32_t)(64 * k + __builtin_ctzll(t));
            } else {
                for (uint32_t k = 0; k < card; k++) o[k] = hi | tmp[k];
            }
        }
        if (card && res && res->c) {
            rb_container *r = &res->c[res->n++];
            r->key = x->key;
            r->card = card;
            r->bitmap = (uint8_t)as_bitmap;
            if (as_bitmap) {
                r->data = malloc(sizeof(wtmp));
                if (r->data) memcpy(r->data, wtmp, sizeof(wtmp));
            } else {
                r->data = malloc(sizeof(uint16_t) * card);
                if (r->data) memcpy(r->data, tmp, sizeof(uint16_t) * card);
            }
        }
        total += card;
        i++;
        j++;
    }
    return total;
}

/* ---- C entry points ---------------------------------------------------------------------------- */

/* Intersection of two sorted sets through roaring bitmaps: writes the values to out (capacity
 * min(na, nb), may be NULL) and returns the count. */
size_t roaring_intersect(const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out)
{
    rb_bitmap x, y;
    if (rb_from_sorted(a, na, &x) || rb_from_sorted(b, nb, &y)) return (size_t)-1;
    uint64_t n = rb_and(&x, &y, out, NULL);
    rb_free(&x);
    rb_free(&y);
    return (size_t)n;
}

/* time_croaring_svs (harness.rs:226-253): the first set's bitmap intersected in place with each
 * further set in turn.  Writes the final values to out (capacity lens[0]) and returns the count. */
size_t roaring_svs(const uint32_t *const *sets, const size_t *lens, size_t k, uint32_t *out)
{
    if (k < 2) return (size_t)-1;
    rb_bitmap victim;
    if (rb_from_sorted(sets[0], lens[0], &victim)) return (size_t)-1;
    for (size_t s = 1; s < k; s++) {
        rb_bitmap other, res;
        if (rb_from_sorted(sets[s], lens[s], &other)) return (size_t)-1;
        rb_and(&victim, &other, NULL, &res);
        rb_free(&other);
        rb_free(&victim);
        victim = res;
    }
    uint64_t n = 0;
    /* decode in key order */
    for (uint32_t i = 0; i < victim.n; i++) {
        const rb_container *c = &victim.c[i];
        const uint32_t hi = (uint32_t)c->key << 16;
        if (c->bitmap) {
            const uint64_t *w = (const uint64_t *)c->data;
            for (int q = 0; q < RB_WORDS; q++)
                for (uint64_t t = w[q]; t; t &= t - 1) {
                    if (out) out[n] = hi | (uint32_t)(64 * q + __builtin_ctzll(t));
                    n++;
                }
        } else {
            const uint16_t *a = (const uint16_t *)c->data;
            for (uint32_t q = 0; q < c->card; q++) {
                if (out) out[n] = hi | a[q];
                n++;
            }
        }
    }
    rb_free(&victim);
    return (size_t)n;
}

/* ---- batch comparator: bitmaps built once (the harness's prepare step), pairs timed ----------- */

typedef struct {
    uint64_t P;
    rb_bitmap *a, *b;
} roaring_batch;

void roaring_batch_free(void *h);

void *roaring_batch_create(const uint32_t *A, const uint64_t *a_off, const uint32_t *B, const uint64_t *b_off,
                           uint64_t P)
{
    roaring_batch *rb = (roaring_batch *)calloc(1, sizeof(roaring_batch));
    if (!rb) return NULL;
    rb->a = (rb_bitmap *)calloc(P ? P : 1, sizeof(rb_bitmap));
    rb->b = (rb_bitmap *)calloc(P ? P : 1, sizeof(rb_bitmap));
    if (!rb->a || !rb->b) {
        roaring_batch_free(rb);
        return NULL;
    }
    for (uint64_t p = 0; p < P; p++) {
        rb->P = p + 1;  /* sets built so far (freed on failure) */
        if (rb_from_sorted(A + a_off[p], a_off[p + 1] - a_off[p], &rb->a[p]) ||
            rb_from_sorted(B + b_off[p], b_off[p + 1] - b_off[p], &rb->b[p])) {
            roaring_batch_free(rb);
            return NULL;
        }
    }
    rb->P = P;
    return rb;
}

void roaring_batch_free(void *h)
{
    roaring_batch *rb = (roaring_batch *)h;
    if (!rb) return;
    for (uint64_t p = 0; p < rb->P; p++) {
        rb_free(&rb->a[p]);
        rb_free(&rb->b[p]);
    }
    free(rb->a);
    free(rb->b);
    free(rb);
}

typedef struct {
    roaring_batch *rb;
    int materialize;
    uint32_t *counts;
    uint64_t next;
    pthread_mutex_t lock;
} rb_job;

static void *rb_worker(void *arg)
{
    rb_job *j = (rb_job *)arg;
    for (;;) {
        pthread_mutex_lock(&j->lock);
        const uint64_t p0 = j->next;
        j->next += 64;
        pthread_mutex_unlock(&j->lock);
        if (p0 >= j->rb->P) break;
        const uint64_t p1 = p0 + 64 < j->rb->P ? p0 + 64 : j->rb->P;
        for (uint64_t p = p0; p < p1; p++) {
            uint64_t c;
            if (j->materialize) {  /* and: a new bitmap holding the intersection */
                rb_bitmap res;
                c = rb_and(&j->rb->a[p], &j->rb->b[p], NULL, &res);
                rb_free(&res);
            } else {  /* and_cardinality */
                c = rb_and(&j->rb->a[p], &j->rb->b[p], NULL, NULL);
            }
            j->counts[p] = (uint32_t)c;
        }
    }
    return NULL;
}

/* Every pair's intersection on nthreads threads (pairs handed out in chunks of 64, as the batch
 * runner of setops_oracle.c does); counts[p] receives |A_p & B_p|.  Returns the wall time in
 * seconds, or a negative value on bad arguments. */
double roaring_batch_run(void *h, int nthreads, int materialize, uint32_t *counts)
{
    roaring_batch *rb = (roaring_batch *)h;
    if (!rb || nthreads < 1 || nthreads > 256) return -1.0;
    pthread_t th[256];
    rb_job job;
    job.rb = rb;
    job.materialize = materialize;
    job.counts = counts;
    job.next = 0;
    pthread_mutex_init(&job.lock, NULL);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, rb_worker, &job);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    pthread_mutex_destroy(&job.lock);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
