// datagen.cpp -- seeded synthetic workloads for the intersection path (host only, g++ -fopenmp).
//
// Random numbers come from SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): state += 0x9E3779B97F4A7C15, then two xor-shift-multiply rounds.
// Every pair / list / query draws from its own stream seeded by mixing (master seed, index), so
// generation is parallel and reproducible bit for bit on any host.
//
// Set semantics follow the reference generator (benchmark/src/generators.rs:30-110): the shared
// elements are drawn first, then the elements only in the smaller set, then those only in the
// larger one, all distinct and uniform over the element space; each set is then sorted.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, n) by 128-bit multiply (Lemire's nearly-divisionless reduction, no rejection:
    // the bias is below 2^-32 for the ranges used here)
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

uint64_t stream_seed(uint64_t master, uint64_t index)
{
    SplitMix64 m(master ^ (index * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
    m.next();
    return m.next();
}

// Draw `n` distinct values from [0, universe) in draw order into `order` (class bitmap in `seen`,
// sized universe/8+1 bytes, must be zero on entry and is zeroed again on exit for reuse).
// Rejection sampling: expected draws n / (1 - n/universe), fine for densities up to ~0.5.
void draw_distinct(SplitMix64 &rng, uint64_t universe, uint64_t n, std::vector<uint8_t> &seen, uint32_t *order)
{
    for (uint64_t k = 0; k < n;) {
        const uint64_t v = rng.below(universe);
        uint8_t &byte = seen[v >> 3];
        const uint8_t bit = (uint8_t)(1u << (v & 7));
        if (byte & bit) continue;
        byte |= bit;
        order[k++] = (uint32_t)v;
    }
    for (uint64_t k = 0; k < n; k++) seen[order[k] >> 3] = 0;
}

// Planted pair: `shared` values in both, `only_a` only in A, `only_b` only in B; sorted.
void planted_pair(SplitMix64 &rng, uint64_t universe, uint64_t shared, uint64_t only_a, uint64_t only_b,
                  std::vector<uint8_t> &seen, std::vector<uint32_t> &scratch, uint32_t *a, uint32_t *b)
{
    const uint64_t n = shared + only_a + only_b;
    scratch.resize(n);
    draw_distinct(rng, universe, n, seen, scratch.data());
    uint64_t ia = 0, ib = 0;
    for (uint64_t k = 0; k < shared; k++) {
        a[ia++] = scratch[k];
        b[ib++] = scratch[k];
    }
    for (uint64_t k = shared; k < shared + only_a; k++) a[ia++] = scratch[k];
    for (uint64_t k = shared + only_a; k < n; k++) b[ib++] = scratch[k];
    std::sort(a, a + ia);
    std::sort(b, b + ib);
}

// Uneven-pair parameters of pair p: larger size L = floor(2^U), U ~ U[lmin, lmax]; smaller size
// S = max(1, floor(L / 2^V)), V ~ U[0, log2(max skew)]; element space = L / density; shared =
// round(sel * S) with sel ~ U[0, 1].
struct UnevenParams {
    uint64_t S, L, universe, shared;
};

UnevenParams uneven_params(SplitMix64 &rng, double lmin, double lmax, double log2_skew, double density)
{
    UnevenParams p;
    const double u = lmin + (lmax - lmin) * rng.unit();
    const double v = log2_skew * rng.unit();
    const double sel = rng.unit();
    p.L = (uint64_t)std::floor(std::exp2(u));
    if (p.L < 1) p.L = 1;
    uint64_t s = (uint64_t)std::floor((double)p.L / std::exp2(v));
    p.S = s < 1 ? 1 : s;
    p.universe = (uint64_t)std::ceil((double)p.L / density);
    if (p.universe < p.L + p.S) p.universe = p.L + p.S;
    p.shared = (uint64_t)std::llround(sel * (double)p.S);
    if (p.shared > p.S) p.shared = p.S;
    return p;
}

} // namespace

extern "C" {

uint64_t setops_gen_splitmix64(uint64_t seed, uint64_t *out, uint64_t n)
{
    SplitMix64 r(seed);
    for (uint64_t k = 0; k < n; k++) out[k] = r.next();
    return r.s;
}

// One planted pair over [0, universe): |A| = shared + only_a, |B| = shared + only_b.
int setops_gen_planted_pair(uint64_t seed, uint64_t universe, uint64_t shared, uint64_t only_a, uint64_t only_b,
                            uint32_t *a, uint32_t *b)
{
    if (shared + only_a + only_b > universe) return -1;
    SplitMix64 rng(stream_seed(seed, 0));
    std::vector<uint8_t> seen(universe / 8 + 1, 0);
    std::vector<uint32_t> scratch;
    planted_pair(rng, universe, shared, only_a, only_b, seen, scratch, a, b);
    return 0;
}

// Pass 1 of the uneven batch: sizes of every pair (A = smaller set, B = larger set).
int setops_gen_uneven_sizes(uint64_t seed, uint64_t npairs, double lmin, double lmax, double log2_skew,
                            double density, uint64_t *a_len, uint64_t *b_len)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)npairs; p++) {
        SplitMix64 rng(stream_seed(seed, (uint64_t)p));
        const UnevenParams q = uneven_params(rng, lmin, lmax, log2_skew, density);
        a_len[p] = q.S;
        b_len[p] = q.L;
    }
    return 0;
}

// Pass 2: fill A and B (CSR offsets from pass 1's sizes).  Also returns per-pair shared counts.
int setops_gen_uneven_fill(uint64_t seed, uint64_t npairs, double lmin, double lmax, double log2_skew,
                           double density, const uint64_t *a_off, const uint64_t *b_off, uint32_t *A, uint32_t *B,
                           uint32_t *shared_out)
{
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        std::vector<uint8_t> seen;
        std::vector<uint32_t> scratch;
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = 0; p < (int64_t)npairs; p++) {
            SplitMix64 rng(stream_seed(seed, (uint64_t)p));
            const UnevenParams q = uneven_params(rng, lmin, lmax, log2_skew, density);
            if (a_off[p + 1] - a_off[p] != q.S || b_off[p + 1] - b_off[p] != q.L) {
                bad |= 1;
                continue;
            }
            if (seen.size() < q.universe / 8 + 1) seen.assign(q.universe / 8 + 1, 0);
            planted_pair(rng, q.universe, q.shared, q.S - q.shared, q.L - q.shared, seen, scratch, A + a_off[p],
                         B + b_off[p]);
            if (shared_out) shared_out[p] = (uint32_t)q.shared;
        }
    }
    return bad ? -1 : 0;
}

// Skewed batch: one large set of `large_n` distinct values from [0, universe), sorted, and
// `nsmall` small sets of `small_n` values each: `shared` drawn (distinct) from the large set, the
// rest from its complement; each small set sorted.  small[i*small_n ...] is set i.
// Fixed-shape batch (the autotuner's grid cells): npairs pairs of `smaller` vs `larger` elements
// over [0, universe) with `shared` planted in both; pair p is A[p*smaller ..] and B[p*larger ..].
int setops_gen_fixed_batch(uint64_t seed, uint64_t npairs, uint64_t larger, uint64_t smaller, uint64_t universe,
                           uint64_t shared, uint32_t *A, uint32_t *B)
{
    if (shared > smaller || smaller > larger || larger + smaller - shared > universe) return -1;
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        std::vector<uint8_t> seen(universe / 8 + 1, 0);
        std::vector<uint32_t> scratch;
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = 0; p < (int64_t)npairs; p++) {
            SplitMix64 rng(stream_seed(seed, (uint64_t)p));
            planted_pair(rng, universe, shared, smaller - shared, larger - shared, seen, scratch, A + p * smaller,
                         B + p * larger);
        }
    }
    return bad ? -1 : 0;
}

int setops_gen_skewed(uint64_t seed, uint64_t universe, uint64_t large_n, uint64_t nsmall, uint64_t small_n,
                      uint64_t shared, uint32_t *large, uint32_t *small)
{
    if (large_n + small_n > universe || shared > small_n || shared > large_n) return -1;
    std::vector<uint8_t> in_large(universe / 8 + 1, 0);
    {
        SplitMix64 rng(stream_seed(seed, 0));
        draw_distinct(rng, universe, large_n, in_large, large);
        for (uint64_t k = 0; k < large_n; k++) in_large[large[k] >> 3] |= (uint8_t)(1u << (large[k] & 7));
        std::sort(large, large + large_n);
    }
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        std::vector<uint8_t> seen_idx(large_n / 8 + 1, 0);
        std::vector<uint32_t> idx(shared);
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)nsmall; i++) {
            SplitMix64 rng(stream_seed(seed, 1 + (uint64_t)i));
            uint32_t *out = small + (uint64_t)i * small_n;
            draw_distinct(rng, large_n, shared, seen_idx, idx.data());
            for (uint64_t k = 0; k < shared; k++) out[k] = large[idx[k]];
            // the rest: outside the large set and distinct (a small set, so a linear check)
            for (uint64_t k = shared; k < small_n;) {
                const uint32_t v = (uint32_t)rng.below(universe);
                if (in_large[v >> 3] & (1u << (v & 7))) continue;
                bool dup = false;
                for (uint64_t j = shared; j < k; j++) dup |= out[j] == v;
                if (dup) continue;
                out[k++] = v;
            }
            std::sort(out, out + small_n);
        }
    }
    return bad ? -1 : 0;
}

// R-MAT graph (Chakrabarti, Zhan & Faloutsos, SDM 2004; Graph500 parameters): edge_factor << scale
// edges, each placing itself by `scale` quadrant choices with probabilities (a, b, c, 1-a-b-c);
// self loops dropped, both directions kept, duplicates removed, neighbour lists sorted.
// rowptr has (1 << scale) + 1 entries; col needs room for 2 * (edge_factor << scale) entries.
// Returns the number of stored (directed) entries, or -1.
int64_t setops_gen_rmat(uint64_t seed, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                        uint64_t *rowptr, uint32_t *col)
{
    if (scale == 0 || scale > 30 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -1;
    const uint64_t n = 1ull << scale, m = (uint64_t)edge_factor << scale;
    std::vector<uint32_t> src(m), dst(m);
    const uint64_t chunk = 1 << 16;
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < (int64_t)((m + chunk - 1) / chunk); ch++) {
        SplitMix64 rng(stream_seed(seed, (uint64_t)ch));
        const uint64_t e0 = (uint64_t)ch * chunk, e1 = std::min<uint64_t>(m, e0 + chunk);
        for (uint64_t e = e0; e < e1; e++) {
            uint32_t u = 0, v = 0;
            for (uint32_t l = 0; l < scale; l++) {
                const double r = rng.unit();
                const uint32_t qu = r >= a + b ? 1u : 0u;                 // lower half (c or d)
                const uint32_t qv = (r >= a && r < a + b) || r >= a + b + c ? 1u : 0u;  // right half (b or d)
                u = (u << 1) | qu;
                v = (v << 1) | qv;
            }
            src[e] = u;
            dst[e] = v;
        }
    }
    // degrees of the symmetrized multigraph (self loops dropped)
    std::vector<uint64_t> deg(n + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)m; e++) {
        if (src[e] == dst[e]) continue;
        __atomic_fetch_add(&deg[src[e]], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[dst[e]], 1, __ATOMIC_RELAXED);
    }
    std::vector<uint64_t> start(n + 1, 0);
    for (uint64_t v = 0; v < n; v++) start[v + 1] = start[v] + deg[v];
    std::vector<uint64_t> cur(start.begin(), start.end() - 1);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)m; e++) {
        const uint32_t u = src[e], v = dst[e];
        if (u == v) continue;
        col[__atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED)] = v;
        col[__atomic_fetch_add(&cur[v], 1, __ATOMIC_RELAXED)] = u;
    }
    // sort + dedup each list in place, then compact
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)n; v++) {
        uint32_t *l = col + start[v];
        const uint64_t d = start[v + 1] - start[v];
        std::sort(l, l + d);
        deg[v] = (uint64_t)(std::unique(l, l + d) - l);
    }
    rowptr[0] = 0;
    for (uint64_t v = 0; v < n; v++) rowptr[v + 1] = rowptr[v] + deg[v];
    for (uint64_t v = 0; v < n; v++)  // sequential: destinations never pass their sources
        if (rowptr[v] != start[v]) std::memmove(col + rowptr[v], col + start[v], deg[v] * sizeof(uint32_t));
    return (int64_t)rowptr[n];
}

} // extern "C"
