// setops_hip.hip -- C ABI of the MI355X sorted-u32 intersection library (see include/setops_hip.h).
//
// Host side: one context per process (one device), an algorithm registry mirroring the
// reference's name table, batch handles that plan once and run many times, and synchronous
// host-pointer entry points built on the same kernels (the plain per-pair drop-in).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/setops_hip.h"
#include "merge_tile.hpp"

using namespace setops;

struct setops_hip_batch;

namespace {

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(SETOPS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------------------------
// kernel configuration
// ---------------------------------------------------------------------------------------------
constexpr int kNT = 256;
constexpr int kK = 16;
constexpr int kU = kNT * kK;
constexpr int kMaxSeg = 130;
// every pair owns at least c0 units so a tile touches at most kMaxSeg pairs
constexpr uint32_t kC0 = (kU + (kMaxSeg - 2) - 1) / (kMaxSeg - 2);

enum Kind { KIND_MERGE = 0 };

struct AlgoEntry {
    const char *name;
    unsigned flags;
    Kind kind;
    bool bsr;
};

// Registered names follow the reference's registry (benchmark/src/timer.rs:80-167, 212-254)
// with a "hip_" prefix.
const AlgoEntry kAlgos[] = {
    {"hip_branchless_merge", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_MERGE, false},
    {"hip_branchless_merge_bsr", SETOPS_ALGO_BSR, KIND_MERGE, true},
};
constexpr int kNumAlgos = (int)(sizeof(kAlgos) / sizeof(kAlgos[0]));

// ---------------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------------
struct Context {
    bool ready = false;
    int device = -1;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    int occ_merge[4] = {0, 0, 0, 0};  // [bsr][count_only] resident blocks per CU
    std::mutex host_mu;               // serialises the host-pointer entry points
    // grow-only device buffers for the host-pointer entry points
    void *host_buf = nullptr;
    size_t host_buf_bytes = 0;
    setops_hip_batch *single[2] = {nullptr, nullptr};  // cached 1-pair handles [bsr]
};

Context g_ctx;
std::mutex g_init_mu;

template <bool BSR, bool CO>
void *merge_kernel_ptr()
{
    return reinterpret_cast<void *>(&merge_tile_kernel<kNT, kK, kMaxSeg, BSR, CO>);
}

int ensure_init(int device)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_ctx.ready) {
        if (device >= 0 && device != g_ctx.device)
            return fail(SETOPS_ERR_INVALID, "library already initialised on device %d", g_ctx.device);
        return SETOPS_OK;
    }
    if (device < 0) device = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(SETOPS_ERR_NO_DEVICE, "no HIP device visible");
    if (device >= n) return fail(SETOPS_ERR_NO_DEVICE, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SETOPS_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950 only", device,
                    prop.gcnArchName);
    g_ctx.num_cu = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    int occ = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)merge_kernel_ptr<false, false>(), kNT, 0));
    g_ctx.occ_merge[0] = occ;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)merge_kernel_ptr<false, true>(), kNT, 0));
    g_ctx.occ_merge[1] = occ;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)merge_kernel_ptr<true, false>(), kNT, 0));
    g_ctx.occ_merge[2] = occ;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)merge_kernel_ptr<true, true>(), kNT, 0));
    g_ctx.occ_merge[3] = occ;
    g_ctx.device = device;
    g_ctx.ready = true;
    return SETOPS_OK;
}

#define ENSURE_INIT()                          \
    do {                                       \
        int rc_ = ensure_init(-1);             \
        if (rc_ != SETOPS_OK) return rc_;      \
    } while (0)

const AlgoEntry *algo_entry(int algo)
{
    if (algo < 0 || algo >= kNumAlgos) return nullptr;
    return &kAlgos[algo];
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

} // namespace

// ---------------------------------------------------------------------------------------------
// batch handle
// ---------------------------------------------------------------------------------------------
struct setops_hip_batch {
    int algo = 0;
    bool bsr = false;
    PairBatch in{};
    TilePlan plan{};
    uint64_t P = 0;
    int grid_plan = 1;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    uint64_t cap_pairs = 0;  // workspace capacity (pairs, tiles)
    uint64_t cap_tiles = 0;
    uint32_t *d_error = nullptr;
    hipStream_t last_stream = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;  // pending timed runs
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
};

namespace {

int batch_alloc(setops_hip_batch *h, uint64_t P, uint64_t ntiles);

// Words of one tile's staging slot: keys then (BSR) states, each >= the most results a tile
// can produce ((U+1)/2 + 1).
constexpr uint32_t kStageHalf = kU / 2 + 64;
constexpr uint32_t kStageWords = 2 * kStageHalf;

// Sizes the handle's workspace for P pairs / utotal units (grow-only) and sets the plan fields.
int batch_reserve(setops_hip_batch *h, uint64_t P, uint64_t utotal)
{
    const uint64_t ntiles = std::max<uint64_t>(1, (utotal + kU - 1) / kU);
    if (ntiles >= 0x7FFFFFFFull) return fail(SETOPS_ERR_INVALID, "batch too large: %llu tiles", (unsigned long long)ntiles);
    if (P > 0xFFFFFFFFull) return fail(SETOPS_ERR_INVALID, "batch too large: %llu pairs", (unsigned long long)P);
    if (!h->ws || P > h->cap_pairs || ntiles > h->cap_tiles) {
        if (h->ws) {
            if (h->last_stream) HIP_TRY(hipStreamSynchronize(h->last_stream));
            HIP_TRY(hipFree(h->ws));
            h->ws = nullptr;
        }
        const uint64_t cp = std::max<uint64_t>(P, h->cap_pairs), ct = std::max<uint64_t>(ntiles, h->cap_tiles);
        int rc = batch_alloc(h, cp, ct);
        if (rc != SETOPS_OK) return rc;
        h->cap_pairs = cp;
        h->cap_tiles = ct;
    }
    h->plan.utotal = utotal;
    h->plan.ntiles = (uint32_t)ntiles;
    h->plan.c0 = kC0;
    h->plan.U = kU;
    h->plan.stage_words = kStageWords;
    const uint64_t gp = std::min<uint64_t>((P + 1 + 255) / 256, 8192);
    h->grid_plan = (int)std::max<uint64_t>(gp, 1);
    return SETOPS_OK;
}

int batch_alloc(setops_hip_batch *h, uint64_t P, uint64_t ntiles)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t o_uoff = take((P + 1) * 8);
    const size_t o_tf = take(ntiles * 4);
    const size_t o_tl = take(ntiles * 4);
    const size_t o_ga = take((ntiles + 1) * 8);
    const size_t o_gb = take((ntiles + 1) * 8);
    const size_t o_tc = take(ntiles * 4);
    const size_t o_te = take(ntiles * 4);
    const size_t o_to = take((ntiles + 1) * 8);
    const size_t ngroups = (ntiles + kScanGroup - 1) / kScanGroup;
    const size_t o_gc = take(ngroups * 8);
    const size_t o_ge = take(ngroups * 8);
    const size_t o_pl = take(P * 4 + 4);
    const size_t o_misc = take(256);  // error word, single-pair offsets
    const size_t o_stage = take(ntiles * (size_t)kStageWords * 4);
    void *ws = nullptr;
    if (hipMalloc(&ws, off) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SETOPS_ERR_NOMEM, "hipMalloc of %zu workspace bytes failed", off);
    }
    char *b = static_cast<char *>(ws);
    h->ws = ws;
    h->ws_bytes = off;
    h->plan.uoff = reinterpret_cast<uint64_t *>(b + o_uoff);
    h->plan.tile_first = reinterpret_cast<uint32_t *>(b + o_tf);
    h->plan.tile_last = reinterpret_cast<uint32_t *>(b + o_tl);
    h->plan.tile_ga = reinterpret_cast<uint64_t *>(b + o_ga);
    h->plan.tile_gb = reinterpret_cast<uint64_t *>(b + o_gb);
    h->plan.tile_count = reinterpret_cast<uint32_t *>(b + o_tc);
    h->plan.tile_elems = reinterpret_cast<uint32_t *>(b + o_te);
    h->plan.tile_off = reinterpret_cast<uint64_t *>(b + o_to);
    h->plan.group_cnt = reinterpret_cast<unsigned long long *>(b + o_gc);
    h->plan.group_el = reinterpret_cast<unsigned long long *>(b + o_ge);
    h->plan.pair_local = reinterpret_cast<uint32_t *>(b + o_pl);
    h->d_error = reinterpret_cast<uint32_t *>(b + o_misc);
    h->plan.single_off = reinterpret_cast<uint64_t *>(b + o_misc + 16);
    h->plan.stage = reinterpret_cast<uint32_t *>(b + o_stage);
    return SETOPS_OK;
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <bool BSR, bool CO>
void launch_merge(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    hipLaunchKernelGGL((merge_tile_kernel<kNT, kK, kMaxSeg, BSR, CO>), dim3(h->plan.ntiles), dim3(kNT), 0, stream, h->in,
                       h->plan, out);
}

int batch_launch(setops_hip_batch *h, uint32_t *d_out, uint32_t *d_out_s, uint64_t cap, uint64_t *d_out_off,
                 uint32_t *d_counts, unsigned long long *d_totals, hipStream_t stream)
{
    const bool count_only = (d_out == nullptr && d_out_s == nullptr);
    if (!count_only) {
        if (d_out == nullptr) return fail(SETOPS_ERR_INVALID, "d_out is NULL but d_out_states is not");
        if (h->bsr && d_out_s == nullptr) return fail(SETOPS_ERR_INVALID, "BSR batch needs d_out_states");
    }
    TileOut out{};
    out.out = d_out;
    out.out_s = d_out_s;
    out.cap = count_only ? 0 : cap;
    out.out_off = d_out_off;
    out.counts = d_counts;
    out.totals = d_totals;
    out.error = h->d_error;
    // plan: sizes only
    hipLaunchKernelGGL(plan_csr_kernel, dim3(h->grid_plan), dim3(256), 0, stream, h->in, h->plan, out);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(plan_split_kernel, dim3((h->plan.ntiles + 1 + 255) / 256), dim3(256), 0, stream, h->in, h->plan);
    HIP_TRY(hipGetLastError());
    // count (and stage)
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (h->timing) {
        if (!h->spare.empty()) {
            ev = h->spare.back();
            h->spare.pop_back();
        } else {
            HIP_TRY(hipEventCreate(&ev.first));
            HIP_TRY(hipEventCreate(&ev.second));
        }
        HIP_TRY(hipEventRecord(ev.first, stream));
    }
    if (h->bsr) {
        if (count_only) launch_merge<true, true>(h, out, stream);
        else launch_merge<true, false>(h, out, stream);
    } else {
        if (count_only) launch_merge<false, true>(h, out, stream);
        else launch_merge<false, false>(h, out, stream);
    }
    HIP_TRY(hipGetLastError());
    if (h->timing) {
        HIP_TRY(hipEventRecord(ev.second, stream));
        h->events.push_back(ev);
    }
    // scan
    const uint32_t ngroups = (h->plan.ntiles + kScanGroup - 1) / kScanGroup;
    hipLaunchKernelGGL((reduce_tiles_kernel<kScanGroup>), dim3(ngroups), dim3(kScanGroup), 0, stream, h->plan);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((scan_tiles_kernel<kScanGroup>), dim3((h->plan.ntiles + kScanGroup - 1) / kScanGroup),
                       dim3(kScanGroup), 0, stream, h->plan, out);
    HIP_TRY(hipGetLastError());
    // write
    if (!count_only || d_out_off) {
        const uint32_t gp = (h->plan.ntiles + 3) / 4;  // 4 tiles (waves) per workgroup
        if (h->bsr)
            hipLaunchKernelGGL((place_tiles_kernel<256, true>), dim3(gp), dim3(256), 0, stream, h->in, h->plan, out);
        else
            hipLaunchKernelGGL((place_tiles_kernel<256, false>), dim3(gp), dim3(256), 0, stream, h->in, h->plan, out);
        HIP_TRY(hipGetLastError());
    }
    h->last_stream = stream;
    return SETOPS_OK;
}

int batch_create(setops_hip_batch **handle, int algo, const uint32_t *d_a, const uint32_t *d_as, uint64_t a_len,
                 const uint64_t *d_a_off, const uint32_t *d_b, const uint32_t *d_bs, uint64_t b_len,
                 const uint64_t *d_b_off, uint64_t npairs)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    *handle = nullptr;
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (npairs == 0) return fail(SETOPS_ERR_INVALID, "batch has no pairs");
    if (!d_a_off || !d_b_off) return fail(SETOPS_ERR_INVALID, "offset arrays must not be NULL");
    if ((a_len && !d_a) || (b_len && !d_b)) return fail(SETOPS_ERR_INVALID, "set pools must not be NULL");
    if (!aligned16(d_a) || !aligned16(d_b)) return fail(SETOPS_ERR_INVALID, "set pools must be 16-byte aligned");
    if (e->bsr) {
        if ((a_len && !d_as) || (b_len && !d_bs)) return fail(SETOPS_ERR_INVALID, "BSR algorithm needs state arrays");
        if (!aligned16(d_as) || !aligned16(d_bs)) return fail(SETOPS_ERR_INVALID, "state arrays must be 16-byte aligned");
    } else if (d_as || d_bs) {
        return fail(SETOPS_ERR_INVALID, "state arrays given to a plain algorithm");
    }
    auto *h = new setops_hip_batch();
    h->algo = algo;
    h->bsr = e->bsr;
    h->P = npairs;
    h->in.A = d_a;
    h->in.As = d_as;
    h->in.B = d_b;
    h->in.Bs = d_bs;
    h->in.a_off = d_a_off;
    h->in.b_off = d_b_off;
    h->in.a_pool = a_len;
    h->in.b_pool = b_len;
    h->in.P = npairs;
    const uint64_t utotal = a_len + b_len + (uint64_t)kC0 * npairs;
    h->plan.single = 0;
    int rc = batch_reserve(h, npairs, utotal);
    if (rc != SETOPS_OK) {
        delete h;
        return rc;
    }
    *handle = h;
    return SETOPS_OK;
}

void batch_free(setops_hip_batch *h)
{
    if (!h) return;
    if (h->last_stream) (void)hipStreamSynchronize(h->last_stream);
    for (auto &e : h->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : h->spare) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (h->ws) (void)hipFree(h->ws);
    delete h;
}

// Grow-only device scratch for the host-pointer entry points (guarded by host_mu).
int host_scratch(size_t bytes, void **p)
{
    if (bytes > g_ctx.host_buf_bytes) {
        if (g_ctx.host_buf) {
            HIP_TRY(hipStreamSynchronize(g_ctx.stream));
            HIP_TRY(hipFree(g_ctx.host_buf));
            g_ctx.host_buf = nullptr;
            g_ctx.host_buf_bytes = 0;
        }
        size_t sz = std::max<size_t>(bytes, 1 << 20);
        if (hipMalloc(&g_ctx.host_buf, sz) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SETOPS_ERR_NOMEM, "hipMalloc of %zu bytes failed", sz);
        }
        g_ctx.host_buf_bytes = sz;
    }
    *p = g_ctx.host_buf;
    return SETOPS_OK;
}

// Context-cached handle for single pairs (plain / BSR).  Its offsets live in the workspace and
// are written by the plan kernel, so a single-pair run needs no host-to-device copy of sizes.
int single_launch(int algo, const uint32_t *dA, const uint32_t *dAs, size_t na, const uint32_t *dB,
                  const uint32_t *dBs, size_t nb, uint32_t *dO, uint32_t *dOs, size_t cap, unsigned long long *dTot,
                  hipStream_t s)
{
    const AlgoEntry *e = algo_entry(algo);
    setops_hip_batch *&h = g_ctx.single[e->bsr ? 1 : 0];
    if (!h) {
        h = new setops_hip_batch();
        h->bsr = e->bsr;
    }
    if (h->last_stream && h->last_stream != s) HIP_TRY(hipStreamSynchronize(h->last_stream));
    h->algo = algo;
    h->P = 1;
    const uint64_t utotal = (uint64_t)na + nb + kC0;
    int rc = batch_reserve(h, 1, utotal);
    if (rc != SETOPS_OK) return rc;
    h->in.A = dA;
    h->in.As = dAs;
    h->in.B = dB;
    h->in.Bs = dBs;
    h->in.a_off = h->plan.single_off;
    h->in.b_off = h->plan.single_off + 2;
    h->in.a_pool = na;
    h->in.b_pool = nb;
    h->in.P = 1;
    h->plan.single = 1;
    h->plan.single_na = na;
    h->plan.single_nb = nb;
    return batch_launch(h, dO, dOs, cap, nullptr, nullptr, dTot, s);
}

// One pair, host pointers: upload, run, download.  keys/states of each side.
int run_single_pair(int algo, const uint32_t *a, const uint32_t *as, size_t na, const uint32_t *b,
                    const uint32_t *bs, size_t nb, uint32_t *out, uint32_t *out_s, size_t out_cap, size_t *out_len,
                    uint64_t *out_elems)
{
    const AlgoEntry *e = algo_entry(algo);
    const bool count_only = (out == nullptr && out_s == nullptr);
    const size_t bytes_a = align_up(na * 4 + 16, 256), bytes_b = align_up(nb * 4 + 16, 256);
    const size_t cap = std::min(na, nb);
    const size_t bytes_o = align_up(cap * 4 + 16, 256);
    const int mult = e->bsr ? 2 : 1;
    const size_t bytes = mult * (bytes_a + bytes_b + bytes_o) + 256;
    void *buf = nullptr;
    int rc = host_scratch(bytes, &buf);
    if (rc != SETOPS_OK) return rc;
    char *p = static_cast<char *>(buf);
    uint32_t *dA = reinterpret_cast<uint32_t *>(p);
    p += bytes_a;
    uint32_t *dB = reinterpret_cast<uint32_t *>(p);
    p += bytes_b;
    uint32_t *dO = reinterpret_cast<uint32_t *>(p);
    p += bytes_o;
    uint32_t *dAs = nullptr, *dBs = nullptr, *dOs = nullptr;
    if (e->bsr) {
        dAs = reinterpret_cast<uint32_t *>(p);
        p += bytes_a;
        dBs = reinterpret_cast<uint32_t *>(p);
        p += bytes_b;
        dOs = reinterpret_cast<uint32_t *>(p);
        p += bytes_o;
    }
    unsigned long long *dTot = reinterpret_cast<unsigned long long *>(p);
    hipStream_t s = g_ctx.stream;
    if (na) HIP_TRY(hipMemcpyAsync(dA, a, na * 4, hipMemcpyHostToDevice, s));
    if (nb) HIP_TRY(hipMemcpyAsync(dB, b, nb * 4, hipMemcpyHostToDevice, s));
    if (e->bsr) {
        if (na) HIP_TRY(hipMemcpyAsync(dAs, as, na * 4, hipMemcpyHostToDevice, s));
        if (nb) HIP_TRY(hipMemcpyAsync(dBs, bs, nb * 4, hipMemcpyHostToDevice, s));
    }
    rc = single_launch(algo, dA, dAs, na, dB, dBs, nb, count_only ? nullptr : dO, count_only ? nullptr : dOs, cap,
                       dTot, s);
    if (rc != SETOPS_OK) return rc;
    unsigned long long tot[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(tot, dTot, sizeof tot, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const size_t n = (size_t)tot[0];
    if (out_len) *out_len = n;
    if (out_elems) *out_elems = tot[1];
    if (!count_only) {
        if (n > out_cap) return fail(SETOPS_ERR_CAPACITY, "result has %zu elements, output capacity is %zu", n, out_cap);
        if (n) {
            HIP_TRY(hipMemcpy(out, dO, n * 4, hipMemcpyDeviceToHost));
            if (e->bsr) HIP_TRY(hipMemcpy(out_s, dOs, n * 4, hipMemcpyDeviceToHost));
        }
    }
    return SETOPS_OK;
}

} // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

int setops_hip_init(int device) { return ensure_init(device); }

const char *setops_hip_last_error(void) { return g_err.c_str(); }

const char *setops_hip_version(void) { return "setops_hip 0.1 gfx950"; }

int setops_hip_algo_count(void) { return kNumAlgos; }

int setops_hip_algo_lookup(const char *name)
{
    if (!name) return fail(SETOPS_ERR_INVALID, "name is NULL");
    for (int i = 0; i < kNumAlgos; i++)
        if (std::strcmp(kAlgos[i].name, name) == 0) return i;
    return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm '%s'", name);
}

const char *setops_hip_algo_name(int algo)
{
    const AlgoEntry *e = algo_entry(algo);
    return e ? e->name : nullptr;
}

unsigned setops_hip_algo_flags(int algo)
{
    const AlgoEntry *e = algo_entry(algo);
    return e ? e->flags : 0u;
}

int setops_hip_intersect2(int algo, const uint32_t *a, size_t na, const uint32_t *b, size_t nb, uint32_t *out,
                          size_t out_cap, size_t *out_len)
{
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (!(e->flags & SETOPS_ALGO_TWOSET)) return fail(SETOPS_ERR_UNSUPPORTED, "%s is not a plain 2-set entry", e->name);
    if ((na && !a) || (nb && !b)) return fail(SETOPS_ERR_INVALID, "input pointer is NULL");
    if (!out_len) return fail(SETOPS_ERR_INVALID, "out_len is NULL");
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    return run_single_pair(algo, a, nullptr, na, b, nullptr, nb, out, nullptr, out_cap, out_len, nullptr);
}

int setops_hip_intersect2_dev(int algo, const uint32_t *d_a, const uint32_t *d_a_states, size_t na, const uint32_t *d_b,
                              const uint32_t *d_b_states, size_t nb, uint32_t *d_out, uint32_t *d_out_states,
                              size_t out_cap, unsigned long long *d_totals, void *stream)
{
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if ((na && !d_a) || (nb && !d_b)) return fail(SETOPS_ERR_INVALID, "input pointer is NULL");
    if (!aligned16(d_a) || !aligned16(d_b)) return fail(SETOPS_ERR_INVALID, "inputs must be 16-byte aligned");
    if (!d_totals) return fail(SETOPS_ERR_INVALID, "d_totals is NULL");
    if (e->bsr) {
        if ((na && !d_a_states) || (nb && !d_b_states)) return fail(SETOPS_ERR_INVALID, "BSR entry needs state arrays");
        if (!aligned16(d_a_states) || !aligned16(d_b_states))
            return fail(SETOPS_ERR_INVALID, "state arrays must be 16-byte aligned");
        if ((d_out == nullptr) != (d_out_states == nullptr))
            return fail(SETOPS_ERR_INVALID, "d_out and d_out_states must both be set or both be NULL");
    } else if (d_a_states || d_b_states || d_out_states) {
        return fail(SETOPS_ERR_INVALID, "state arrays given to a plain entry");
    }
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL = the HIP null stream
    return single_launch(algo, d_a, d_a_states, na, d_b, d_b_states, nb, d_out, d_out_states, out_cap, d_totals, s);
}

int setops_hip_intersect2_bsr(int algo, const uint32_t *bases_a, const uint32_t *states_a, size_t na,
                              const uint32_t *bases_b, const uint32_t *states_b, size_t nb, uint32_t *bases_out,
                              uint32_t *states_out, size_t out_cap, size_t *out_words, uint64_t *out_elems)
{
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (!(e->flags & SETOPS_ALGO_BSR)) return fail(SETOPS_ERR_UNSUPPORTED, "%s is not a BSR entry", e->name);
    if ((na && (!bases_a || !states_a)) || (nb && (!bases_b || !states_b)))
        return fail(SETOPS_ERR_INVALID, "input pointer is NULL");
    if ((bases_out == nullptr) != (states_out == nullptr))
        return fail(SETOPS_ERR_INVALID, "bases_out and states_out must both be set or both be NULL");
    if (!out_words) return fail(SETOPS_ERR_INVALID, "out_words is NULL");
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    return run_single_pair(algo, bases_a, states_a, na, bases_b, states_b, nb, bases_out, states_out, out_cap,
                           out_words, out_elems);
}

int setops_hip_svs(int algo, const uint32_t *const *sets, const size_t *lens, size_t k, uint32_t *out, size_t out_cap,
                   size_t *out_len)
{
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (!(e->flags & SETOPS_ALGO_KSET)) return fail(SETOPS_ERR_UNSUPPORTED, "%s cannot drive svs_generic", e->name);
    if (k < 2) return fail(SETOPS_ERR_INVALID, "svs needs at least 2 sets (svs.rs:61)");
    if (!sets || !lens || !out_len) return fail(SETOPS_ERR_INVALID, "NULL argument");
    for (size_t s = 0; s < k; s++)
        if (lens[s] && !sets[s]) return fail(SETOPS_ERR_INVALID, "set %zu is NULL", s);
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    // svs.rs:50-73: fold the 2-set entry over the sets in the given order.
    std::vector<uint32_t> left(std::max<size_t>(1, std::min(lens[0], lens[1]))), right;
    size_t n = 0;
    int rc = run_single_pair(algo, sets[0], nullptr, lens[0], sets[1], nullptr, lens[1], left.data(), nullptr,
                             left.size(), &n, nullptr);
    if (rc != SETOPS_OK) return rc;
    for (size_t s = 2; s < k; s++) {
        std::swap(left, right);
        left.assign(std::max<size_t>(1, std::min(n, lens[s])), 0u);
        size_t m = 0;
        rc = run_single_pair(algo, right.data(), nullptr, n, sets[s], nullptr, lens[s], left.data(), nullptr,
                             left.size(), &m, nullptr);
        if (rc != SETOPS_OK) return rc;
        n = m;
    }
    *out_len = n;
    if (out) {
        if (n > out_cap) return fail(SETOPS_ERR_CAPACITY, "result has %zu elements, output capacity is %zu", n, out_cap);
        std::memcpy(out, left.data(), n * 4);
    }
    return SETOPS_OK;
}

int setops_hip_batch_create(setops_hip_batch **handle, int algo, const uint32_t *d_a, const uint32_t *d_a_states,
                            uint64_t a_len, const uint64_t *d_a_off, const uint32_t *d_b, const uint32_t *d_b_states,
                            uint64_t b_len, const uint64_t *d_b_off, uint64_t npairs)
{
    ENSURE_INIT();
    return batch_create(handle, algo, d_a, d_a_states, a_len, d_a_off, d_b, d_b_states, b_len, d_b_off, npairs);
}

int setops_hip_batch_run(setops_hip_batch *handle, uint32_t *d_out, uint32_t *d_out_states, uint64_t out_cap,
                         uint64_t *d_out_off, uint32_t *d_counts, unsigned long long *d_totals, void *stream)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL = the HIP null stream
    return batch_launch(handle, d_out, d_out_states, out_cap, d_out_off, d_counts, d_totals, s);
}

int setops_hip_batch_check(setops_hip_batch *handle)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (!handle->last_stream) return SETOPS_OK;
    uint32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, handle->d_error, 4, hipMemcpyDeviceToHost, handle->last_stream));
    HIP_TRY(hipStreamSynchronize(handle->last_stream));
    if (err & 2u) return fail(SETOPS_ERR_INVALID, "internal error: a tile exceeded its segment bound");
    if (err & 1u) return fail(SETOPS_ERR_CAPACITY, "batch output exceeded its capacity (writes were clipped)");
    return SETOPS_OK;
}

int setops_hip_batch_set_timing(setops_hip_batch *handle, int enabled)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    handle->timing = enabled != 0;
    return SETOPS_OK;
}

int setops_hip_batch_kernel_time(setops_hip_batch *handle, double *total_ms, uint64_t *runs)
{
    if (!handle || !total_ms || !runs) return fail(SETOPS_ERR_INVALID, "NULL argument");
    double tot = 0.0;
    for (auto &e : handle->events) {
        HIP_TRY(hipEventSynchronize(e.second));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
        handle->spare.push_back(e);
    }
    *runs = handle->events.size();
    *total_ms = tot;
    handle->events.clear();
    return SETOPS_OK;
}

int setops_hip_batch_info(setops_hip_batch *handle, uint64_t *ntiles, uint64_t *grid, uint64_t *units)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (ntiles) *ntiles = handle->plan.ntiles;
    if (grid) *grid = (uint64_t)handle->plan.ntiles;
    if (units) *units = handle->plan.utotal;
    return SETOPS_OK;
}

int setops_hip_batch_destroy(setops_hip_batch *handle)
{
    batch_free(handle);
    return SETOPS_OK;
}

} // extern "C"
