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

} // extern "C"
