// property_tests.cpp -- the reference's unit and property tests, restated in C++ over the GPU
// entries through the C++ API (simd_set_operations_amd/cpp/setops.hpp).
//
// Follows setops/tests/unit_tests.rs:6-59 (known answers) and setops/tests/property_tests.rs
// (every 2-set entry equals the ascending intersection, the three properties of
// testlib/properties.rs:2-49, BSR round trips and popcount counts, k-set SvS), checked against
// std::set_intersection.  `--registry` lists the registered entries without touching a device.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <random>
#include <set>

#include "../../simd_set_operations_amd/cpp/setops.hpp"

using setops::Slice;
using Set = std::vector<uint32_t>;

static int g_fail = 0, g_checks = 0;
#define EXPECT(cond, ...)                                                 \
    do {                                                                  \
        g_checks++;                                                       \
        if (!(cond)) {                                                    \
            g_fail++;                                                     \
            std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
            std::fprintf(stderr, __VA_ARGS__);                            \
            std::fprintf(stderr, "\n");                                   \
        }                                                                 \
    } while (0)

static Set textbook(const Set &a, const Set &b)
{
    Set r;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(r));
    return r;
}

// testlib/properties.rs:2-49
static bool prop_intersection_correct(const Set &result, const std::vector<Set> &sets)
{
    for (size_t i = 1; i < result.size(); i++)
        if (!(result[i - 1] < result[i])) return false;  // strictly increasing
    for (uint32_t x : result)
        for (const Set &s : sets)
            if (!std::binary_search(s.begin(), s.end(), x)) return false;  // every item is common
    for (uint32_t x : sets[0]) {
        bool common = true;
        for (size_t k = 1; k < sets.size() && common; k++) common = std::binary_search(sets[k].begin(), sets[k].end(), x);
        if (common && !std::binary_search(result.begin(), result.end(), x)) return false;  // every common item is in
    }
    return true;
}

static Set random_set(std::mt19937_64 &rng, size_t n, uint64_t universe)
{
    std::set<uint32_t> s;
    std::uniform_int_distribution<uint64_t> d(0, universe - 1);
    while (s.size() < std::min<uint64_t>(n, universe)) s.insert((uint32_t)d(rng));
    return Set(s.begin(), s.end());
}

static void unit_tests(const std::vector<setops::Entry> &plain)
{
    const Set a5 = {1, 3, 5, 8, 9, 10, 14, 15, 18, 20}, b5 = {1, 2, 3, 4, 9, 10, 11, 15, 20, 21};
    struct Kat {
        Set a, b, out;
    } kats[] = {
        {{1, 2, 3, 4}, {1, 2, 3, 4, 5}, {1, 2, 3, 4}},
        {{0, 4, 5, 8}, {1, 2, 3, 6}, {}},
        {{1, 4, 5}, {1, 4, 5}, {1, 4, 5}},
        {{10, 42}, {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 22, 25, 28, 39, 42, 43, 47, 49}, {10, 42}},
        {a5, b5, {1, 3, 9, 10, 15, 20}},  // const_intersect(A, B) of unit_tests.rs:27-33
    };
    for (const auto &f : plain)
        for (size_t i = 0; i < sizeof kats / sizeof kats[0]; i++) {
            EXPECT(setops::run_2set(kats[i].a, kats[i].b, f) == kats[i].out, "%s kat %zu", f.name().c_str(), i);
            EXPECT(setops::run_2set_c(kats[i].a, kats[i].b, f) == kats[i].out, "%s kat_c %zu", f.name().c_str(), i);
        }
    // test_simd_galloping (unit_tests.rs:47-59): [1 << 12 + 1] = [8192] against 0..12345
    Set small = {8192}, large(12345);
    for (uint32_t i = 0; i < 12345; i++) large[i] = i;
    const auto expected = setops::run_2set(small, large, setops::Entry::lookup("hip_branchless_merge"));
    EXPECT(setops::run_2set(small, large, setops::Entry::lookup("hip_galloping")) == expected, "galloping 8192");
    EXPECT(expected == Set{8192}, "8192 present");
}

static void property_tests(const std::vector<setops::Entry> &plain, const std::vector<setops::Entry> &bsr)
{
    std::mt19937_64 rng(20240517);
    for (int trial = 0; trial < 30; trial++) {
        const uint64_t universe = 16 + rng() % 200000;
        const Set a = random_set(rng, rng() % 20000, universe);
        const Set b = random_set(rng, rng() % 20000, universe);
        const Set exp = textbook(a, b);
        for (const auto &f : plain) {
            const Set got = setops::run_2set(a, b, f);
            EXPECT(got == exp, "%s trial %d", f.name().c_str(), trial);
            EXPECT(prop_intersection_correct(got, {a, b}), "%s props %d", f.name().c_str(), trial);
            EXPECT(setops::run_2set(b, a, f) == exp, "%s swapped %d", f.name().c_str(), trial);
            setops::Counter c;
            f(a, b, c);
            EXPECT(c.count() == exp.size(), "%s counter %d", f.name().c_str(), trial);
        }
        const setops::BsrVec ba = setops::BsrVec::from_sorted(a), bb = setops::BsrVec::from_sorted(b);
        EXPECT(ba.to_sorted_set() == a, "bsr round trip %d", trial);  // bsr_encode_decode
        for (const auto &f : bsr) {
            const setops::BsrVec r = setops::run_2set_bsr(ba.as_ref(), bb.as_ref(), f);
            EXPECT(r.to_sorted_set() == exp, "%s bsr %d", f.name().c_str(), trial);
            bool ok = true;
            for (size_t i = 0; i < r.len(); i++) ok = ok && r.states[i] != 0 && (i == 0 || r.bases[i - 1] < r.bases[i]);
            EXPECT(ok, "%s bsr form %d", f.name().c_str(), trial);
            setops::Counter c;
            f.bsr(ba.as_ref(), bb.as_ref(), c);
            EXPECT(c.count() == exp.size(), "%s bsr counter %d", f.name().c_str(), trial);
        }
    }
    // k sets, smallest first (svs.rs): run_svs == fold of the textbook intersection == run_kset
    for (int trial = 0; trial < 12; trial++) {
        const size_t k = 2 + rng() % 5;
        const Set common = random_set(rng, rng() % 300, 100000);
        std::vector<Set> sets;
        for (size_t i = 0; i < k; i++) {
            Set s = random_set(rng, rng() % 20000, 100000), u;
            std::set_union(s.begin(), s.end(), common.begin(), common.end(), std::back_inserter(u));
            sets.push_back(u);
        }
        std::sort(sets.begin(), sets.end(), [](const Set &x, const Set &y) { return x.size() < y.size(); });
        Set exp = sets[0];
        for (size_t i = 1; i < k; i++) exp = textbook(exp, sets[i]);
        const std::vector<Slice> sl(sets.begin(), sets.end());
        for (const auto &f : plain) {
            if (!f.is_kset()) continue;
            const Set got = setops::run_svs(sl, f);
            EXPECT(got == exp, "%s svs %d", f.name().c_str(), trial);
            EXPECT(prop_intersection_correct(got, sets), "%s svs props %d", f.name().c_str(), trial);
            EXPECT(setops::run_kset(sl, f) == exp, "%s kset %d", f.name().c_str(), trial);
        }
    }
}

static void error_tests()
{
    const setops::Entry f = setops::Entry::lookup("hip_branchless_merge");
    Set a(100);
    for (uint32_t i = 0; i < 100; i++) a[i] = i;
    Set out(10);
    int code = 0;
    try {
        f.intersect_c(a, a, out.data(), out.size());
    } catch (const setops::Error &e) {
        code = e.code();
    }
    EXPECT(code == SETOPS_ERR_CAPACITY, "capacity error %d", code);
    code = 0;
    try {
        setops::Entry::lookup("branchless_merge_but_wrong");
    } catch (const setops::Error &e) {
        code = e.code();
    }
    EXPECT(code == SETOPS_ERR_UNKNOWN_ALGO, "unknown name %d", code);
    code = 0;
    try {
        setops::run_svs({Slice(a)}, f);
    } catch (const setops::Error &e) {
        code = e.code();
    }
    EXPECT(code == SETOPS_ERR_INVALID, "svs needs two sets %d", code);
    setops::Harness h(std::chrono::milliseconds(20));
    const auto [secs, r] = h.time([&] { return Set(); }, [&](Set &d) { d = setops::run_2set(a, a, f); });
    EXPECT(secs > 0 && r == a, "harness time %g", secs);
}

int main(int argc, char **argv)
{
    std::vector<setops::Entry> plain, bsr;
    for (const std::string &n : setops::registered()) {
        const setops::Entry e = setops::Entry::lookup(n);
        (e.is_bsr() ? bsr : plain).push_back(e);
    }
    if (argc > 1 && std::strcmp(argv[1], "--registry") == 0) {
        for (const auto &e : plain) std::printf("%s plain%s\n", e.name().c_str(), e.is_kset() ? " kset" : "");
        for (const auto &e : bsr) std::printf("%s bsr\n", e.name().c_str());
        return plain.empty() || bsr.empty();
    }
    setops::init(0);
    unit_tests(plain);
    property_tests(plain, bsr);
    error_tests();
    std::printf("%d checks, %d failures (%zu plain entries, %zu BSR entries)\n", g_checks, g_fail, plain.size(),
                bsr.size());
    return g_fail ? 1 : 0;
}
