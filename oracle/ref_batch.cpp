// ref_batch.cpp -- TEST INFRASTRUCTURE: batch driver for the reference's own QFilter.
//
// Linked into _ref/libqfilter_ref.so together with the reference's setops/ffi/qfilter/qfilter.cpp
// (compiled from its source where it lies).  Runs intersect_qfilter_uint_b4_v2 once per pair, one
// pair at a time per host thread, over a padded layout that satisfies the reference's contract:
// every set starts 16-byte aligned and is followed by >= 16 readable bytes, and every output slot
// has 3 spare elements (qfilter.cpp:228-229,257,266,272 load/store whole 16-byte vectors).
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

extern "C" int intersect_qfilter_uint_b4_v2(const int *set_a, int size_a, const int *set_b, int size_b, int *set_c);

extern "C" int ref_batch_qfilter(int nthreads, const int32_t *A, const uint64_t *a_start, const uint32_t *a_len,
                                 const int32_t *B, const uint64_t *b_start, const uint32_t *b_len, uint64_t npairs,
                                 int32_t *out, const uint64_t *out_start, uint32_t *counts)
{
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const uint64_t p0 = next.fetch_add(64);
            if (p0 >= npairs) break;
            const uint64_t p1 = p0 + 64 < npairs ? p0 + 64 : npairs;
            for (uint64_t p = p0; p < p1; p++)
                counts[p] = (uint32_t)intersect_qfilter_uint_b4_v2(A + a_start[p], (int)a_len[p], B + b_start[p],
                                                                   (int)b_len[p], out + out_start[p]);
        }
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(work);
    for (auto &t : th) t.join();
    return 0;
}
