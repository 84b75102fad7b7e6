// setops_hip.hip -- C ABI of the MI355X sorted-u32 intersection library (see include/setops_hip.h).
//
// Host side: one context per process (one device), an algorithm registry mirroring the
// reference's name table, batch handles that plan once and run many times, and synchronous
// host-pointer entry points built on the same kernels (the plain per-pair drop-in).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/setops_hip.h"
#include "merge_tile.hpp"
#include "search_tile.hpp"

using namespace setops;

struct setops_hip_batch;

// Device scratch of setops_hip_intersect2_dev_noalloc: the plan and staging of one pair call.  Every
// GPU the library runs on holds its own copy until the process ends.  On the host this name is only
// a stand-in for the symbol: host code uses the device address that hipGetSymbolAddress returned
// for the current device at initialisation (g_scratch), never the name itself.
constexpr size_t kScratchBytes = 20u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

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
// merge policy: kNTM lanes x K merge-path steps per tile; search policy: kNT lanes x KS probes.
// The tile shape is selected per call (SETOPS_SHAPE(i) in the algorithm id): shape 0 is the
// default, the others exist for the plain merge and search entries, for the autotuner to pick.
constexpr int kNT = 256;  // workgroup size of the search tile
constexpr int kNTM = 256;
constexpr int kNumShapes = 4;
constexpr int kMergeK[kNumShapes] = {20, 16, 24, 12};  // merge steps per lane
constexpr int kSearchKS[kNumShapes] = {2, 1, 4, 0};    // probes per lane (0: no such shape)
constexpr int kK = kMergeK[0];
// BSR merge tiles: each word carries a 32-bit state beside its key, so a tile's LDS is twice a
// plain one's; half the steps per lane keep as many tiles per CU (measured on the uneven batch in
// BSR form: merge kernel 0.178 ms at 20 steps, 0.154-0.156 at 10, 0.159 at 12, 0.157 at 8)
constexpr int kBsrK = 10;
constexpr int kKS = kSearchKS[0];
constexpr int kMaxSeg = 130;
// gather merge tiles stage every pair's windows as whole 16-byte chunks, so their LDS holds 16
// words of slack per segment: fewer segments per tile, more tiles per CU (R-MAT scale-22 edges,
// count-only merge: 587 ms at 130 segments, 490 at 66, 431 at 50 and at 34; the BSR merge, whose
// tiles hold states too: 600 ms at 50, 526 at 34)
constexpr int kMaxSegGather = 34;
// every pair owns at least c0 = ceil(U / (maxseg - 2)) units, so a tile touches at most maxseg
// pairs; c0 >= K (KS): a lane crosses at most one pair boundary
constexpr uint32_t pad_for(uint32_t U, int maxseg) { return (U + (maxseg - 2) - 1) / (maxseg - 2); }
static_assert(pad_for(kNTM * 12, kMaxSeg) >= 12 && pad_for(kNTM * 20, kMaxSeg) >= 20 &&
                  pad_for(kNTM * kBsrK, kMaxSeg) >= kBsrK && pad_for(kNT * 1, kMaxSeg) >= 1 &&
                  pad_for(kNTM * 12, kMaxSegGather) >= 12 && pad_for(kNTM * 24, kMaxSegGather) >= 24,
              "a lane crosses at most one pair boundary");
constexpr int kCapB = 3072;  // search policy: larger windows staged up to this many words

constexpr uint32_t kOneScanGroups = 8;      // pair-unit scans of up to this many groups run as one workgroup
constexpr uint32_t kScanDirectGroups = 64;  // tile-count scans of up to this many groups run as one kernel
constexpr uint64_t kMaxGrid = 1u << 22;  // grid-stride beyond this many workgroups
constexpr int kPlaceNT = 64;   // lanes of a placement workgroup (one tile each)
constexpr int kPlaceQ = 8;     // words per lane per round of the placement kernel

enum Kind {
    KIND_MERGE = 0,
    KIND_GALLOP = 1,
    KIND_BINARY = 2,
    KIND_SHUFFLE = 3,
    KIND_QFILTER = 4,
    KIND_ADAPTIVE = 5,
    KIND_BLOCK_GALLOP = 6,
};

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
    {"hip_galloping", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_GALLOP, false},
    {"hip_galloping_bsr", SETOPS_ALGO_BSR, KIND_GALLOP, true},
    {"hip_binary_search", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_BINARY, false},
    // simd_galloping (galloping_sse/avx2/avx512, simd_galloping.rs:19-193): gallop over blocks,
    // then compare the one block that can hold the probe in full
    {"hip_simd_galloping", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_BLOCK_GALLOP, false},
    {"hip_simd_galloping_bsr", SETOPS_ALGO_BSR, KIND_BLOCK_GALLOP, true},
    // small_adaptive (adaptive.rs:107-147): k sets in one pass, each element of the first set
    // galloped through the others in turn; on two sets it is galloping
    {"hip_small_adaptive", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_ADAPTIVE, false},
    {"hip_shuffling", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_SHUFFLE, false},
    {"hip_shuffling_bsr", SETOPS_ALGO_BSR, KIND_SHUFFLE, true},
    // broadcast_sse/avx2/avx512 (broadcast.rs): each B lane splatted against the A block.  In one
    // GPU lane's registers the splats and shuffling's rotations are the same 16 compares of a 4x4
    // block, so the entry runs the block-compare tile.
    {"hip_broadcast", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_SHUFFLE, false},
    {"hip_broadcast_bsr", SETOPS_ALGO_BSR, KIND_SHUFFLE, true},
    {"hip_qfilter", SETOPS_ALGO_TWOSET | SETOPS_ALGO_KSET, KIND_QFILTER, false},
    {"hip_qfilter_bsr", SETOPS_ALGO_BSR, KIND_QFILTER, true},
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
    std::mutex host_mu;  // serialises the host-pointer entry points
    // grow-only device buffers for the host-pointer entry points
    void *host_buf = nullptr;
    size_t host_buf_bytes = 0;
    setops_hip_batch *single[kNumAlgos][kNumShapes] = {};  // cached 1-pair handles per algorithm and shape
};

Context g_ctx;
std::mutex g_init_mu;

// The scratch array's device address per device (set at initialisation; null until then or if the
// lookup failed, with the reason in g_scratch_err).
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::string g_scratch_err;

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
    // the no-allocation entry point's scratch on this device (a failed lookup disables only that
    // entry point, which then reports the reason)
    if (device < kMaxDevices) {
        void *sp = nullptr;
        const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (se == hipSuccess && sp) {
            g_scratch[device] = sp;
        } else {
            (void)hipGetLastError();
            g_scratch_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(se);
        }
    } else {
        g_scratch_err = "device index beyond the scratch table";
    }
    g_ctx.device = device;
    g_ctx.ready = true;
    return SETOPS_OK;
}

#define ENSURE_INIT()                          \
    do {                                       \
        int rc_ = ensure_init(-1);             \
        if (rc_ != SETOPS_OK) return rc_;      \
    } while (0)

// Registered entry of an algorithm id; SETOPS_SIGNED_I32 and SETOPS_SHAPE(i) may be OR'ed into it.
constexpr int kAlgoFlags = SETOPS_SIGNED_I32 | SETOPS_SHAPE_MASK;
const AlgoEntry *algo_entry(int algo)
{
    algo &= ~kAlgoFlags;
    if (algo < 0 || algo >= kNumAlgos) return nullptr;
    return &kAlgos[algo];
}
int algo_base(int algo) { return algo & ~kAlgoFlags; }
int shape_of(int algo) { return (algo & SETOPS_SHAPE_MASK) >> SETOPS_SHAPE_SHIFT; }
// PairBatch::sflip of an algorithm id: int32 sets compare with bit 31 flipped.
uint32_t sflip_of(int algo) { return (algo & SETOPS_SIGNED_I32) ? 0x80000000u : 0u; }
// int32 mode exists for plain entries only: the reference's BSR form is built from u32 values
// (benchmark/src/timer/harness.rs:121-122 casts the i32 sets to u32 before BsrVec::from_sorted).
int shape_count(Kind k, bool bsr);
// The flags OR'ed into an algorithm id must fit its entry: the int32 mode for plain entries, a
// tile shape the entry offers.
int check_flags(int algo, const AlgoEntry *e)
{
    if ((algo & SETOPS_SIGNED_I32) && e->bsr)
        return fail(SETOPS_ERR_UNSUPPORTED, "%s: BSR entries take u32 bases (no int32 mode)", e->name);
    if (shape_of(algo) >= shape_count(e->kind, e->bsr))
        return fail(SETOPS_ERR_UNSUPPORTED, "%s has no tile shape %d (it has %d)", e->name, shape_of(algo),
                    shape_count(e->kind, e->bsr));
    return SETOPS_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

} // namespace

// ---------------------------------------------------------------------------------------------
// batch handle
// ---------------------------------------------------------------------------------------------
struct setops_hip_batch {
    int algo = 0;
    int shape = 0;  // tile shape (index into kMergeK / kSearchKS)
    Kind kind = KIND_MERGE;
    bool bsr = false;
    bool gather = false;
    PairBatch in{};
    TilePlan plan{};
    uint64_t P = 0;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    uint64_t cap_pairs = 0;  // workspace capacity (pairs, tiles)
    uint64_t cap_tiles = 0;
    uint32_t *d_error = nullptr;
    uint32_t *stage = nullptr;  // result staging (separate allocation, see stage_reserve)
    uint32_t *b_idx = nullptr;  // index of the B pool (setops_hip_batch_index_b): every 32nd element
    uint32_t *b_top = nullptr;  // its top level: every kTopStride-th element
    uint32_t *b_mid = nullptr;  // its middle level: every kMidStride-th element
    uint64_t stage_tiles = 0;
    unsigned long long *d_tot = nullptr;  // [2] scratch totals (unit total, or totals when the caller passes none)
    hipStream_t last_stream = nullptr;  // the stream of the last run (NULL: the HIP null stream)
    bool ran = false;                   // a run was enqueued (on last_stream)
    bool timing = false;
    // captured launch sequence (batch_launch)
    bool use_graph = true;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    struct GraphKeyT {
        void *out, *out_s;
        uint64_t cap;
        void *out_off, *counts, *totals;
        PairBatch in;
        TilePlan plan;
    } graph_key{}, seen_key{};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;  // pending timed runs
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
};

namespace {

int batch_alloc(setops_hip_batch *h, uint64_t P, uint64_t ntiles);

// Merge-path kinds (merge, shuffling, qfilter) share the merge tile; the search kinds the search tile.
bool merge_like(Kind k) { return k == KIND_MERGE || k == KIND_SHUFFLE || k == KIND_QFILTER; }
uint32_t policy_of(Kind k) { return merge_like(k) ? POLICY_MERGE : POLICY_SEARCH; }
// Tile shapes a kind offers: the plain merge entry and the plain search entries have several,
// every other entry (block policies, BSR) has the default only.
int shape_count(Kind k, bool bsr)
{
    if (bsr) return 1;
    if (k == KIND_MERGE) return kNumShapes;
    if (!merge_like(k)) return 3;
    return 1;
}
uint32_t units_per_tile(Kind k, bool bsr, int shape)
{
    if (k == KIND_MERGE && bsr) return (uint32_t)(kNTM * kBsrK);
    return merge_like(k) ? (uint32_t)(kNTM * kMergeK[shape]) : (uint32_t)(kNT * kSearchKS[shape]);
}
// Segments a tile may touch: gather merge tiles have the smaller limit.
int max_seg(Kind k, bool gather) { return (gather && merge_like(k)) ? kMaxSegGather : kMaxSeg; }
uint32_t pad_units(Kind k, bool bsr, bool gather, int shape)
{
    return pad_for(units_per_tile(k, bsr, shape), max_seg(k, gather));
}
// Words of one tile's staging slot: keys then (BSR) states, each >= the most results a tile
// can produce (merge: (U+1)/2 + 1; search: U).
uint32_t stage_words(Kind k, bool bsr, int shape)
{
    const uint32_t U = units_per_tile(k, bsr, shape);
    const uint32_t half = (merge_like(k) ? U / 2 : U) + 64;
    return bsr ? 2 * half : half;
}
// The plan's unit prefix has a closed form only for merge-policy CSR batches (and single pairs).
bool closed_form(const setops_hip_batch *h) { return h->plan.single || (!h->gather && merge_like(h->kind)); }

// Padding units per pair of a batch whose smallest pair owns min_own units.  The padding bounds
// the pairs one tile touches (<= U / c0 + 2) and keeps every pair at one unit or more.  Search
// policy: when every pair already owns pad_units of its own, those bounds hold without padding,
// and the batch has no padding tiles (the skewed workload's 4,096 pairs of 256 probes make
// exactly 2,048 tiles of 512).  The merge tile needs its padding (a lane's steps past the end of
// its pair run in the padding), so the merge policy always pads.
uint32_t pad_for_batch(Kind k, bool bsr, bool gather, int shape, uint64_t min_own)
{
    const uint32_t pad = pad_units(k, bsr, gather, shape);
    return (!merge_like(k) && min_own >= pad) ? 0u : pad;
}

// Sizes the handle's workspace for P pairs / utotal units (grow-only) and sets the plan fields
// (c0: padding units per pair, counted in utotal).
int batch_reserve(setops_hip_batch *h, uint64_t P, uint64_t utotal, uint32_t c0)
{
    const uint64_t U = units_per_tile(h->kind, h->bsr, h->shape);
    const uint64_t ntiles = std::max<uint64_t>(1, (utotal + U - 1) / U);
    if (ntiles >= 0x7FFFFFFFull) return fail(SETOPS_ERR_INVALID, "batch too large: %llu tiles", (unsigned long long)ntiles);
    if (P >= 0xFFFFFFFFull) return fail(SETOPS_ERR_INVALID, "batch too large: %llu pairs", (unsigned long long)P);
    if (!h->ws || P > h->cap_pairs || ntiles > h->cap_tiles) {
        if (h->ws) {
            if (h->ran) HIP_TRY(hipStreamSynchronize(h->last_stream));
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
    h->plan.c0 = c0;
    h->plan.U = (uint32_t)U;
    h->plan.policy = policy_of(h->kind);
    return SETOPS_OK;
}

// Lays out the workspace of P pairs and ntiles tiles at `b` and points the plan into it (b == NULL:
// sizes only).  Returns its bytes.
size_t ws_layout(setops_hip_batch *h, uint64_t P, uint64_t ntiles, char *b)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const uint32_t sw = stage_words(h->kind, h->bsr, h->shape);
    const size_t o_uoff = take((P + 1) * 8);
    const size_t o_tf = take(ntiles * 4);
    const size_t o_tl = take(ntiles * 4);
    const size_t o_sa = take(ntiles * 4);
    const size_t o_sb = take(ntiles * 4);
    const size_t o_ga = take((ntiles + 1) * 8);
    const size_t o_gb = take((ntiles + 1) * 8);
    const size_t o_tc = take(ntiles * 4);
    const size_t o_te = take(ntiles * 4);
    const size_t o_to = take((ntiles + 1) * 8);
    const size_t ngroups = (std::max<uint64_t>(ntiles, P) + kScanGroup - 1) / kScanGroup;
    const size_t o_gc = take(ngroups * 8);
    const size_t o_ge = take(ngroups * 8);
    const size_t o_pl = take(P * 4 + 4);
    const size_t o_misc = take(256);  // error word, scratch totals, single-pair offsets
    if (!b) return off;
    h->plan.uoff = reinterpret_cast<uint64_t *>(b + o_uoff);
    h->plan.tile_first = reinterpret_cast<uint32_t *>(b + o_tf);
    h->plan.tile_last = reinterpret_cast<uint32_t *>(b + o_tl);
    h->plan.tile_sa = reinterpret_cast<uint32_t *>(b + o_sa);
    h->plan.tile_sb = reinterpret_cast<uint32_t *>(b + o_sb);
    h->plan.tile_ga = reinterpret_cast<uint64_t *>(b + o_ga);
    h->plan.tile_gb = reinterpret_cast<uint64_t *>(b + o_gb);
    h->plan.tile_count = reinterpret_cast<uint32_t *>(b + o_tc);
    h->plan.tile_elems = reinterpret_cast<uint32_t *>(b + o_te);
    h->plan.tile_off = reinterpret_cast<uint64_t *>(b + o_to);
    h->plan.gsum = reinterpret_cast<unsigned long long *>(b + o_gc);
    h->plan.gsum2 = reinterpret_cast<unsigned long long *>(b + o_ge);
    h->plan.pair_local = reinterpret_cast<uint32_t *>(b + o_pl);
    h->d_error = reinterpret_cast<uint32_t *>(b + o_misc);
    h->d_tot = reinterpret_cast<unsigned long long *>(b + o_misc + 16);
    h->plan.single_off = reinterpret_cast<uint64_t *>(b + o_misc + 64);
    h->plan.stage_words = sw;
    return off;
}

int batch_alloc(setops_hip_batch *h, uint64_t P, uint64_t ntiles)
{
    const size_t bytes = ws_layout(h, P, ntiles, nullptr);
    void *ws = nullptr;
    if (hipMalloc(&ws, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SETOPS_ERR_NOMEM, "hipMalloc of %zu workspace bytes failed", bytes);
    }
    h->ws = ws;
    h->ws_bytes = bytes;
    ws_layout(h, P, ntiles, static_cast<char *>(ws));
    return SETOPS_OK;
}

// Per-tile result staging for nt tiles, allocated on the first run that writes results
// (count-only runs never need it), grow-only.
int stage_grow(setops_hip_batch *h, uint64_t nt)
{
    if (h->stage && h->stage_tiles >= nt) {
        h->plan.stage = h->stage;
        return SETOPS_OK;
    }
    if (h->stage) {
        if (h->ran) HIP_TRY(hipStreamSynchronize(h->last_stream));
        HIP_TRY(hipFree(h->stage));
        h->stage = nullptr;
        h->stage_tiles = 0;
    }
    const size_t bytes = (size_t)nt * h->plan.stage_words * 4;
    if (hipMalloc(&h->stage, bytes) != hipSuccess) {
        (void)hipGetLastError();
        h->stage = nullptr;
        return fail(SETOPS_ERR_NOMEM, "hipMalloc of %zu staging bytes failed", bytes);
    }
    h->stage_tiles = nt;
    h->plan.stage = h->stage;
    return SETOPS_OK;
}

int stage_reserve(setops_hip_batch *h) { return stage_grow(h, h->plan.ntiles); }

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

uint32_t grid_for(uint64_t n, uint64_t per_block)
{
    const uint64_t g = (n + per_block - 1) / per_block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, kMaxGrid));
}

// Unit prefix by the three-kernel scan (gather batches, search policy).
int launch_units_scan(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    const uint32_t ng = (uint32_t)((h->P + kScanGroup - 1) / kScanGroup);
    // up to kOneScanGroups groups of pairs (serving steps, small gather batches such as the skewed
    // workload's 4,096 pairs): the scan is one workgroup, a group at a time
    const bool one = ng <= kOneScanGroups;
    if (!one) {
        hipLaunchKernelGGL((units_reduce_kernel<kScanGroup>), dim3(ng), dim3(kScanGroup), 0, stream, h->in,
                           h->plan, 0u);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((groups_scan_kernel<kScanGroup>), dim3(1), dim3(kScanGroup), 0, stream, h->plan.gsum,
                           h->plan.gsum2, (uint64_t)ng, 0, h->d_tot);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL((units_scan_kernel<kScanGroup, kOneScanGroups>), dim3(one ? 1 : ng), dim3(kScanGroup), 0, stream,
                       h->in, h->plan, out, one ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return SETOPS_OK;
}

// Total units of a batch whose prefix has no closed form: group sums and their scan on the
// device, total read back (a synchronisation at handle creation only).
int device_unit_total(setops_hip_batch *h, uint64_t *utotal, uint64_t *small)
{
    const uint32_t ng = (uint32_t)((h->P + kScanGroup - 1) / kScanGroup);
    hipStream_t s = g_ctx.stream;
    // the offsets may have been written by work still pending on another stream (this count runs
    // on the library's own stream, and creation synchronises anyway)
    HIP_TRY(hipDeviceSynchronize());
    hipLaunchKernelGGL((units_reduce_kernel<kScanGroup>), dim3(ng), dim3(kScanGroup), 0, s, h->in, h->plan,
                       h->plan.c0);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((groups_scan_kernel<kScanGroup>), dim3(1), dim3(kScanGroup), 0, s, h->plan.gsum, h->plan.gsum2,
                       (uint64_t)ng, 1, h->d_tot);
    HIP_TRY(hipGetLastError());
    unsigned long long tot[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(tot, h->d_tot, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *utotal = tot[0];
    *small = tot[1];
    return SETOPS_OK;
}

// Tile kernels run one workgroup per tile; batches of more than kMaxGrid tiles take several
// launches (tile offset tile0).  A direct plan (one plain pair in one tile, set only by the
// single-pair entries) launches the kernels' DIRECT form, which takes the pair from the plan's
// fields instead of the tile maps.
template <int K, bool BSR, bool CO, bool GA, int LP>
void launch_merge_t(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    if constexpr (!BSR && !GA) {
        if (h->plan.direct) {
            hipLaunchKernelGGL((merge_tile_kernel<kNTM, K, kMaxSeg, BSR, CO, GA, LP, true>), dim3(1), dim3(kNTM), 0,
                               stream, h->in, h->plan, out, 0u);
            return;
        }
    }
    for (uint64_t t0 = 0; t0 < h->plan.ntiles; t0 += kMaxGrid)
        hipLaunchKernelGGL((merge_tile_kernel<kNTM, K, GA ? kMaxSegGather : kMaxSeg, BSR, CO, GA, LP>),
                           dim3((uint32_t)std::min<uint64_t>(kMaxGrid, h->plan.ntiles - t0)), dim3(kNTM), 0, stream,
                           h->in, h->plan, out, (uint32_t)t0);
}

template <int KS, bool BSR, bool CO, int SMODE>
void launch_search_t(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    if constexpr (!BSR) {
        if (h->plan.direct) {
            hipLaunchKernelGGL((search_tile_kernel<kNT, KS, kMaxSeg, kCapB, BSR, CO, SMODE, true>),
                               dim3((uint32_t)h->plan.ntiles), dim3(kNT), 0, stream, h->in, h->plan, out, 0u);
            return;
        }
    }
    for (uint64_t t0 = 0; t0 < h->plan.ntiles; t0 += kMaxGrid)
        hipLaunchKernelGGL((search_tile_kernel<kNT, KS, kMaxSeg, kCapB, BSR, CO, SMODE>),
                           dim3((uint32_t)std::min<uint64_t>(kMaxGrid, h->plan.ntiles - t0)), dim3(kNT), 0, stream,
                           h->in, h->plan, out, (uint32_t)t0);
}

// The merge entry in the handle's tile shape (plain sets; BSR merge tiles take kBsrK steps).
template <bool BSR, bool CO, bool GA>
void launch_merge_shape(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    if constexpr (BSR) {
        launch_merge_t<kBsrK, BSR, CO, GA, LP_MERGE>(h, out, stream);
    } else {
        switch (h->shape) {
        case 1: launch_merge_t<kMergeK[1], BSR, CO, GA, LP_MERGE>(h, out, stream); break;
        case 2: launch_merge_t<kMergeK[2], BSR, CO, GA, LP_MERGE>(h, out, stream); break;
        case 3: launch_merge_t<kMergeK[3], BSR, CO, GA, LP_MERGE>(h, out, stream); break;
        default: launch_merge_t<kMergeK[0], BSR, CO, GA, LP_MERGE>(h, out, stream); break;
        }
    }
}

template <bool BSR, bool CO, int SMODE>
void launch_search_shape(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    if constexpr (BSR) {
        launch_search_t<kSearchKS[0], BSR, CO, SMODE>(h, out, stream);
    } else {
        switch (h->shape) {
        case 1: launch_search_t<kSearchKS[1], BSR, CO, SMODE>(h, out, stream); break;
        case 2: launch_search_t<kSearchKS[2], BSR, CO, SMODE>(h, out, stream); break;
        default: launch_search_t<kSearchKS[0], BSR, CO, SMODE>(h, out, stream); break;
        }
    }
}

template <bool BSR, bool CO>
void launch_tiles_b(setops_hip_batch *h, const TileOut &out, hipStream_t stream)
{
    switch (h->kind) {
    case KIND_MERGE:
        if (h->gather) launch_merge_shape<BSR, CO, true>(h, out, stream);
        else launch_merge_shape<BSR, CO, false>(h, out, stream);
        break;
    case KIND_SHUFFLE:
        if (h->gather) launch_merge_t<kK, BSR, CO, true, LP_SHUFFLE>(h, out, stream);
        else launch_merge_t<kK, BSR, CO, false, LP_SHUFFLE>(h, out, stream);
        break;
    case KIND_QFILTER:
        if (h->gather) launch_merge_t<kK, BSR, CO, true, LP_QFILTER>(h, out, stream);
        else launch_merge_t<kK, BSR, CO, false, LP_QFILTER>(h, out, stream);
        break;
    case KIND_GALLOP:
    case KIND_ADAPTIVE: launch_search_shape<BSR, CO, SM_GALLOP>(h, out, stream); break;
    case KIND_BINARY: launch_search_shape<BSR, CO, SM_BINARY>(h, out, stream); break;
    case KIND_BLOCK_GALLOP: launch_search_shape<BSR, CO, SM_BLOCK>(h, out, stream); break;
    }
}

// The launch sequence of one run (no allocation, no synchronisation): plan, split, tile kernel,
// scan, placement.  `ev` (may be null) receives events around the tile kernel.
int batch_enqueue(setops_hip_batch *h, const TileOut &out, bool count_only, hipStream_t stream,
                  std::pair<hipEvent_t, hipEvent_t> *ev)
{
    // plan: unit prefix and tile maps, then tile starts (one pair over several merge tiles: both
    // in one launch; one pair in one tile: none, the tile kernel takes the pair as it is)
    const bool single_split = h->plan.single && !h->plan.direct && h->plan.policy == POLICY_MERGE;
    if (h->plan.direct) {
        // no plan launch
    } else if (single_split) {
        hipLaunchKernelGGL(plan_split_single_kernel, dim3(grid_for(h->plan.ntiles, 256)), dim3(256), 0, stream,
                           h->in, h->plan, out);
        HIP_TRY(hipGetLastError());
    } else if (closed_form(h)) {
        hipLaunchKernelGGL(plan_csr_kernel, dim3(grid_for(h->P, 256)), dim3(256), 0, stream, h->in, h->plan, out);
        HIP_TRY(hipGetLastError());
    } else {
        int rc = launch_units_scan(h, out, stream);
        if (rc != SETOPS_OK) return rc;
    }
    // merge policy: the merge-path split of every tile start (the search policy's tile starts were
    // recorded with the tile maps)
    if (!single_split && !h->plan.direct && h->plan.policy == POLICY_MERGE) {
        hipLaunchKernelGGL(plan_split_kernel, dim3(grid_for(h->plan.ntiles, 256)), dim3(256), 0, stream, h->in,
                           h->plan);
        HIP_TRY(hipGetLastError());
    }
    // count (and stage)
    if (ev) HIP_TRY(hipEventRecord(ev->first, stream));
    if (h->bsr) {
        if (count_only) launch_tiles_b<true, true>(h, out, stream);
        else launch_tiles_b<true, false>(h, out, stream);
    } else {
        if (count_only) launch_tiles_b<false, true>(h, out, stream);
        else launch_tiles_b<false, false>(h, out, stream);
    }
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev->second, stream));
    if (h->plan.one_tile) return SETOPS_OK;  // results, offsets and totals are final
    // scan: with few tiles the placement kernel scans the tile counts itself; with few groups one
    // kernel (a workgroup adds the earlier groups' tile counts itself); otherwise group sums,
    // one-block scan of the group sums (writes the totals), per-group scan
    const bool place = !count_only || out.out_off;
    const int fused = (place && h->plan.ntiles <= kFuseScanTiles) ? 1 : 0;
    if (!fused) {
        const uint32_t ngroups = (h->plan.ntiles + kScanGroup - 1) / kScanGroup;
        const int direct = ngroups <= kScanDirectGroups ? 1 : 0;
        TileOut so = out;
        if (!so.totals) so.totals = h->d_tot;
        if (!direct) {
            hipLaunchKernelGGL((reduce_tiles_kernel<kScanGroup>), dim3(ngroups), dim3(kScanGroup), 0, stream,
                               h->plan);
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL((groups_scan_kernel<kScanGroup>), dim3(1), dim3(kScanGroup), 0, stream, h->plan.gsum,
                               h->plan.gsum2, (uint64_t)ngroups, 1, so.totals);
            HIP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL((scan_tiles_kernel<kScanGroup>), dim3(ngroups), dim3(kScanGroup), 0, stream, h->plan, so,
                           direct);
        HIP_TRY(hipGetLastError());
    }
    // count-only runs with many tiles: the per-pair offsets alone (one thread per pair)
    if (place && !fused && count_only) {
        hipLaunchKernelGGL(pair_offsets_kernel, dim3(grid_for(h->P, 256)), dim3(256), 0, stream, h->in, h->plan, out);
        HIP_TRY(hipGetLastError());
        return SETOPS_OK;
    }
    // write: one workgroup per tile
    if (place) {
        for (uint64_t t0 = 0; t0 < h->plan.ntiles; t0 += kMaxGrid) {
            const dim3 gp((uint32_t)std::min<uint64_t>(kMaxGrid, h->plan.ntiles - t0));
            if (h->bsr && fused)
                hipLaunchKernelGGL((place_tiles_kernel<kPlaceNT, kPlaceQ, true, true>), gp, dim3(kPlaceNT), 0, stream,
                                   h->in, h->plan, out, (uint32_t)t0);
            else if (h->bsr)
                hipLaunchKernelGGL((place_tiles_kernel<kPlaceNT, kPlaceQ, true, false>), gp, dim3(kPlaceNT), 0, stream,
                                   h->in, h->plan, out, (uint32_t)t0);
            else if (fused)
                hipLaunchKernelGGL((place_tiles_kernel<kPlaceNT, kPlaceQ, false, true>), gp, dim3(kPlaceNT), 0, stream,
                                   h->in, h->plan, out, (uint32_t)t0);
            else
                hipLaunchKernelGGL((place_tiles_kernel<kPlaceNT, kPlaceQ, false, false>), gp, dim3(kPlaceNT), 0,
                                   stream, h->in, h->plan, out, (uint32_t)t0);
        }
        HIP_TRY(hipGetLastError());
    }
    return SETOPS_OK;
}

void graph_drop(setops_hip_batch *h)
{
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
}

// A run: eager when timed (events around the tile kernel), otherwise the launch sequence is
// captured once into a HIP graph (per handle, per set of output pointers and plan) and replayed
// with one launch.  The sequence records kernels only (no memset or copy nodes).
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
    // a run on another stream than the last one would share the workspace and staging with it
    if (h->ran && h->last_stream != stream) HIP_TRY(hipStreamSynchronize(h->last_stream));
    // a plain search-policy batch of at most kNT pairs runs the search tile's direct form, which
    // finds its pairs itself (no plan launch): in one tile with room for every probe's result, the
    // tile's slot is the output and the tile writes the offsets and totals (no scan, no
    // placement); over a few tiles (a fused placement follows), when no per-pair counts are asked
    // for (they would be added up across tiles from zero, and nothing zeroes them)
    if (!h->plan.single) {
        const bool small = !count_only && !h->bsr && h->plan.policy == POLICY_SEARCH && h->P <= (uint64_t)kNT;
        const bool one = h->plan.ntiles == 1 && h->P <= (uint64_t)kMaxSeg && cap >= h->plan.utotal - h->P * h->plan.c0;
        const bool few = h->plan.ntiles > 1 && h->plan.ntiles <= kFuseScanTiles && d_counts == nullptr;
        h->plan.direct = (small && (one || few)) ? 1u : 0u;
    }
    h->plan.one_tile = (h->plan.direct && h->plan.ntiles == 1) ? 1u : 0u;
    if (!count_only) {
        int rc = stage_reserve(h);
        if (rc != SETOPS_OK) return rc;
        if (h->plan.one_tile) h->plan.stage = d_out;  // the one tile's slot is the output
    }
    h->last_stream = stream;
    h->ran = true;
    if (h->timing || !h->use_graph) {
        std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
        if (h->timing) {
            if (!h->spare.empty()) {
                ev = h->spare.back();
                h->spare.pop_back();
            } else {
                HIP_TRY(hipEventCreate(&ev.first));
                HIP_TRY(hipEventCreate(&ev.second));
            }
        }
        int rc = batch_enqueue(h, out, count_only, stream, h->timing ? &ev : nullptr);
        if (h->timing) h->events.push_back(ev);
        return rc;
    }
    // the captured sequence depends on the outputs and on everything the plan holds
    setops_hip_batch::GraphKeyT key;
    std::memset(static_cast<void *>(&key), 0, sizeof key);
    key.out = d_out;
    key.out_s = d_out_s;
    key.cap = out.cap;
    key.out_off = d_out_off;
    key.counts = d_counts;
    key.totals = d_totals;
    key.in = h->in;
    key.plan = h->plan;
    const bool cached = h->graph_exec && std::memcmp(&key, &h->graph_key, sizeof key) == 0;
    if (!cached && std::memcmp(&key, &h->seen_key, sizeof key) != 0) {
        // first run with these arguments: eager (a run whose arguments change every time, like a
        // rebound serving batch or a single pair, never pays for a capture)
        h->seen_key = key;
        return batch_enqueue(h, out, count_only, stream, nullptr);
    }
    if (!cached) {
        graph_drop(h);
        if (!h->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue(h, out, count_only, h->cap_stream, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->cap_stream, &g);
        if (rc != SETOPS_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(SETOPS_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
        h->graph = g;
        HIP_TRY(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
        h->graph_key = key;
    }
    HIP_TRY(hipGraphLaunch(h->graph_exec, stream));
    return SETOPS_OK;
}

// Shared by the CSR and gather constructors: a_end == NULL means CSR (a_end = a_beg + 1).
int batch_create(setops_hip_batch **handle, int algo, const uint32_t *d_a, const uint32_t *d_as, uint64_t a_len,
                 const uint64_t *d_a_beg, const uint64_t *d_a_end, const uint32_t *d_b, const uint32_t *d_bs,
                 uint64_t b_len, const uint64_t *d_b_beg, const uint64_t *d_b_end, uint64_t npairs)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    *handle = nullptr;
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (int rc = check_flags(algo, e)) return rc;
    if (npairs == 0) return fail(SETOPS_ERR_INVALID, "batch has no pairs");
    const bool gather = d_a_end != nullptr;
    if (!d_a_beg || !d_b_beg || (gather && !d_b_end)) return fail(SETOPS_ERR_INVALID, "offset arrays must not be NULL");
    if ((a_len && !d_a) || (b_len && !d_b)) return fail(SETOPS_ERR_INVALID, "set pools must not be NULL");
    if (!aligned16(d_a) || !aligned16(d_b)) return fail(SETOPS_ERR_INVALID, "set pools must be 16-byte aligned");
    if (e->bsr) {
        if ((a_len && !d_as) || (b_len && !d_bs)) return fail(SETOPS_ERR_INVALID, "BSR algorithm needs state arrays");
        if (!aligned16(d_as) || !aligned16(d_bs)) return fail(SETOPS_ERR_INVALID, "state arrays must be 16-byte aligned");
    } else if (d_as || d_bs) {
        return fail(SETOPS_ERR_INVALID, "state arrays given to a plain algorithm");
    }
    auto *h = new setops_hip_batch();
    h->algo = algo_base(algo);
    h->shape = shape_of(algo);
    h->kind = e->kind;
    h->bsr = e->bsr;
    h->gather = gather;
    h->in.sflip = sflip_of(algo);
    h->P = npairs;
    h->in.A = d_a;
    h->in.As = d_as;
    h->in.B = d_b;
    h->in.Bs = d_bs;
    h->in.a_beg = d_a_beg;
    h->in.a_end = gather ? d_a_end : d_a_beg + 1;
    h->in.b_beg = d_b_beg;
    h->in.b_end = gather ? d_b_end : d_b_beg + 1;
    h->in.a_pool = a_len;
    h->in.b_pool = b_len;
    h->in.P = npairs;
    h->in.contig = gather ? 0u : 1u;
    h->plan.single = 0;
    h->plan.direct = 0;
    h->plan.one_tile = 0;
    int rc;
    const uint32_t pad = pad_units(h->kind, h->bsr, h->gather, h->shape);
    if (!gather && merge_like(e->kind)) {
        rc = batch_reserve(h, npairs, a_len + b_len + (uint64_t)pad * npairs, pad);
    } else {
        // size the scan arrays first, then count units (with full padding) and the pairs that own
        // fewer than `pad` units on the device; without such pairs a search batch needs no padding
        rc = batch_reserve(h, npairs, 1, pad);
        uint64_t utotal = 0, small = 0;
        if (rc == SETOPS_OK) rc = device_unit_total(h, &utotal, &small);
        const uint32_t c0 = pad_for_batch(h->kind, h->bsr, h->gather, h->shape, small ? 0 : pad);
        if (rc == SETOPS_OK) rc = batch_reserve(h, npairs, utotal - (uint64_t)(pad - c0) * npairs, c0);
    }
    if (rc != SETOPS_OK) {
        if (h->ws) (void)hipFree(h->ws);
        delete h;
        return rc;
    }
    *handle = h;
    return SETOPS_OK;
}

void batch_free(setops_hip_batch *h)
{
    if (!h) return;
    if (h->ran) (void)hipStreamSynchronize(h->last_stream);
    for (auto &e : h->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : h->spare) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    graph_drop(h);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    if (h->ws) (void)hipFree(h->ws);
    if (h->stage) (void)hipFree(h->stage);
    if (h->b_idx) (void)hipFree(h->b_idx);
    if (h->b_top) (void)hipFree(h->b_top);
    if (h->b_mid) (void)hipFree(h->b_mid);
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

// Context-cached handle for single pairs, one per algorithm.  Its offsets live in the workspace
// and are written by the plan kernel, so a single-pair run needs no host-to-device copy of sizes.
int single_launch(int algo, const uint32_t *dA, const uint32_t *dAs, size_t na, const uint32_t *dB,
                  const uint32_t *dBs, size_t nb, uint32_t *dO, uint32_t *dOs, size_t cap, unsigned long long *dTot,
                  hipStream_t s, const uint32_t *const *dX = nullptr, const uint64_t *dXlen = nullptr, uint32_t nx = 0)
{
    const AlgoEntry *e = algo_entry(algo);
    setops_hip_batch *&h = g_ctx.single[algo_base(algo)][shape_of(algo)];
    if (!h) {
        h = new setops_hip_batch();
        h->algo = algo_base(algo);
        h->shape = shape_of(algo);
        h->kind = e->kind;
        h->bsr = e->bsr;
    }
    h->in.sflip = sflip_of(algo);
    if (h->ran && h->last_stream != s) HIP_TRY(hipStreamSynchronize(h->last_stream));
    h->P = 1;
    const uint64_t units =
        (merge_like(e->kind) ? (uint64_t)na + nb : (uint64_t)std::min(na, nb)) + pad_units(e->kind, e->bsr, false, h->shape);
    h->plan.single = 1;
    int rc = batch_reserve(h, 1, units, pad_units(e->kind, e->bsr, false, h->shape));
    if (rc != SETOPS_OK) return rc;
    h->in.A = dA;
    h->in.As = dAs;
    h->in.B = dB;
    h->in.Bs = dBs;
    h->in.a_beg = h->plan.single_off;
    h->in.a_end = h->plan.single_off + 1;
    h->in.b_beg = h->plan.single_off + 2;
    h->in.b_end = h->plan.single_off + 3;
    h->in.a_pool = na;
    h->in.b_pool = nb;
    h->in.P = 1;
    h->in.contig = 1;
    h->in.xsets = dX;
    h->in.xlens = dXlen;
    h->in.nx = nx;
    h->plan.single_na = na;
    h->plan.single_nb = nb;
    // one tile, plain sets, and room for any result: the tile writes the output and the totals
    h->plan.direct = (h->plan.ntiles == 1 && !e->bsr && dTot && (dO == nullptr || cap >= std::min(na, nb))) ? 1u : 0u;
    return batch_launch(h, dO, dOs, cap, nullptr, nullptr, dTot, s);
}

// setops_hip_intersect2_dev_noalloc: one pair planned and staged in the device scratch array.  The
// launch sequence is single_launch's, enqueued eagerly (no graph capture, no events), and nothing
// is allocated: the workspace and the staging slots are laid out in the scratch of the current
// device.  Every pipeline kernel writes what a later one reads (the plan kernel clears the totals,
// the error word and the counts), so the scratch needs no clearing between calls.
constexpr uint64_t kNoAllocMaxMerge = 1ull << 22;   // na + nb of a merge-path entry
constexpr uint64_t kNoAllocMaxProbes = 1ull << 20;  // min(na, nb) of a search entry
setops_hip_batch g_noalloc;  // the call's host-side description (reused; owns no device memory)

int noalloc_launch(int algo, const uint32_t *dA, const uint32_t *dAs, size_t na, const uint32_t *dB,
                   const uint32_t *dBs, size_t nb, uint32_t *dO, uint32_t *dOs, size_t cap, unsigned long long *dTot,
                   hipStream_t s)
{
    const AlgoEntry *e = algo_entry(algo);
    const bool ml = merge_like(e->kind);
    const uint64_t work = ml ? (uint64_t)na + nb : (uint64_t)std::min(na, nb);
    const uint64_t limit = ml ? kNoAllocMaxMerge : kNoAllocMaxProbes;
    if (work > limit)
        return fail(SETOPS_ERR_INVALID, "%s: %s = %llu exceeds the no-allocation limit of %llu (a %zu-byte scratch)",
                    e->name, ml ? "na + nb" : "min(na, nb)", (unsigned long long)work, (unsigned long long)limit,
                    kScratchBytes);
    char *base = static_cast<char *>(g_scratch[g_ctx.device]);
    if (!base) return fail(SETOPS_ERR_NO_DEVICE, "no scratch on device %d: %s", g_ctx.device, g_scratch_err.c_str());
    setops_hip_batch *h = &g_noalloc;
    h->algo = algo_base(algo);
    h->shape = shape_of(algo);
    h->kind = e->kind;
    h->bsr = e->bsr;
    h->gather = false;
    h->P = 1;
    h->timing = false;
    h->use_graph = false;
    h->in = PairBatch{};
    h->plan = TilePlan{};
    const uint64_t U = units_per_tile(e->kind, e->bsr, h->shape);
    const uint64_t units = work + pad_units(e->kind, e->bsr, false, h->shape);
    const uint64_t ntiles = std::max<uint64_t>(1, (units + U - 1) / U);
    const size_t ws = align_up(ws_layout(h, 1, ntiles, base), 256);
    const size_t stage_bytes = (size_t)ntiles * h->plan.stage_words * 4;
    if (ws + stage_bytes > kScratchBytes)
        return fail(SETOPS_ERR_INVALID, "internal: %zu scratch bytes needed, %zu available", ws + stage_bytes,
                    kScratchBytes);
    h->plan.stage = reinterpret_cast<uint32_t *>(base + ws);
    h->plan.utotal = units;
    h->plan.ntiles = (uint32_t)ntiles;
    h->plan.c0 = pad_units(e->kind, e->bsr, false, h->shape);
    h->plan.U = (uint32_t)U;
    h->plan.policy = policy_of(e->kind);
    h->plan.single = 1;
    h->plan.single_na = na;
    h->plan.single_nb = nb;
    h->in.A = dA;
    h->in.As = dAs;
    h->in.B = dB;
    h->in.Bs = dBs;
    h->in.a_beg = h->plan.single_off;
    h->in.a_end = h->plan.single_off + 1;
    h->in.b_beg = h->plan.single_off + 2;
    h->in.b_end = h->plan.single_off + 3;
    h->in.a_pool = na;
    h->in.b_pool = nb;
    h->in.P = 1;
    h->in.contig = 1;
    h->in.sflip = sflip_of(algo);
    const bool count_only = (dO == nullptr && dOs == nullptr);
    h->plan.direct = (ntiles == 1 && !e->bsr && (count_only || cap >= std::min(na, nb))) ? 1u : 0u;
    h->plan.one_tile = h->plan.direct;
    if (!count_only && h->plan.direct) h->plan.stage = dO;  // the one tile's slot is the output
    TileOut out{};
    out.out = dO;
    out.out_s = dOs;
    out.cap = count_only ? 0 : cap;
    out.totals = dTot;
    out.error = h->d_error;
    return batch_enqueue(h, out, count_only, s, nullptr);
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


// small_adaptive over k > 2 host sets (adaptive.rs:107-147): every set is uploaded once; one
// launch intersects the first set with the next one and up to kMaxExtraSets further sets (each
// probe found in the second set is looked up in each further set in turn); with more sets the
// result, still on the device, becomes the first set of the next launch.
int run_adaptive(int algo, const uint32_t *const *sets, const size_t *lens, size_t k, uint32_t *out, size_t out_cap,
                 size_t *out_len)
{
    const size_t groups = (k - 1 + kMaxExtraSets) / (kMaxExtraSets + 1);  // launches of 1 + kMaxExtraSets sets
    const size_t table = align_up(groups * kMaxExtraSets * 8, 256);      // pointers, then lengths
    size_t bytes = 256 + 2 * table;
    for (size_t s = 0; s < k; s++) bytes += align_up(lens[s] * 4 + 16, 256);
    const size_t cap = std::min(lens[0], lens[1]);  // bounds every launch's result
    const size_t res_bytes = align_up(std::max<size_t>(cap, 1) * 4 + 16, 256);
    bytes += 2 * res_bytes;
    void *buf = nullptr;
    int rc = host_scratch(bytes, &buf);
    if (rc != SETOPS_OK) return rc;
    char *p = static_cast<char *>(buf);
    hipStream_t st = g_ctx.stream;
    unsigned long long *dTot = reinterpret_cast<unsigned long long *>(p);
    p += 256;
    std::vector<const uint32_t *> dsets(k);
    for (size_t s = 0; s < k; s++) {
        uint32_t *d = reinterpret_cast<uint32_t *>(p);
        if (lens[s]) HIP_TRY(hipMemcpyAsync(d, sets[s], lens[s] * 4, hipMemcpyHostToDevice, st));
        dsets[s] = d;
        p += align_up(lens[s] * 4 + 16, 256);
    }
    uint32_t *res[2] = {reinterpret_cast<uint32_t *>(p), reinterpret_cast<uint32_t *>(p + res_bytes)};
    p += 2 * res_bytes;
    // pointer and length tables of every launch's further sets
    const uint32_t *const *dX = reinterpret_cast<const uint32_t *const *>(p);
    p += table;
    const uint64_t *dL = reinterpret_cast<const uint64_t *>(p);
    p += table;
    // launch g intersects `first` (sets[0], then the previous result) with sets[next] and the
    // further sets sets[next+1 .. next+nx]
    const uint32_t *first = dsets[0];
    size_t nfirst = lens[0];
    size_t next = 1;
    int which = 0;
    for (size_t g = 0; next < k; g++) {
        if (g >= groups) return fail(SETOPS_ERR_INVALID, "internal: launch count");
        const size_t nx = std::min<size_t>(kMaxExtraSets, k - next - 1);
        // the further sets of this launch are sets[next+1 .. next+nx]
        std::vector<const uint32_t *> gx(kMaxExtraSets, nullptr);
        std::vector<uint64_t> gl(kMaxExtraSets, 0);
        for (size_t j = 0; j < nx; j++) {
            gx[j] = dsets[next + 1 + j];
            gl[j] = lens[next + 1 + j];
        }
        const uint32_t *const *dXg = dX + g * kMaxExtraSets;
        const uint64_t *dLg = dL + g * kMaxExtraSets;
        HIP_TRY(hipMemcpyAsync(const_cast<const uint32_t **>(dXg), gx.data(), kMaxExtraSets * sizeof(void *),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(const_cast<uint64_t *>(dLg), gl.data(), kMaxExtraSets * 8, hipMemcpyHostToDevice, st));
        rc = single_launch(algo, first, nullptr, nfirst, dsets[next], nullptr, lens[next], res[which], nullptr,
                           std::max<size_t>(cap, 1), dTot, st, dXg, dLg, (uint32_t)nx);
        if (rc != SETOPS_OK) return rc;
        unsigned long long tot[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(tot, dTot, sizeof tot, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        first = res[which];
        nfirst = (size_t)tot[0];
        which ^= 1;
        next += 1 + nx;
    }
    *out_len = nfirst;
    if (out) {
        if (nfirst > out_cap)
            return fail(SETOPS_ERR_CAPACITY, "result has %zu elements, output capacity is %zu", nfirst, out_cap);
        if (nfirst) HIP_TRY(hipMemcpy(out, first, nfirst * 4, hipMemcpyDeviceToHost));
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
    if (int rc = check_flags(algo, e)) return rc;
    if ((na && !a) || (nb && !b)) return fail(SETOPS_ERR_INVALID, "input pointer is NULL");
    if (!out_len) return fail(SETOPS_ERR_INVALID, "out_len is NULL");
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    return run_single_pair(algo, a, nullptr, na, b, nullptr, nb, out, nullptr, out_cap, out_len, nullptr);
}

namespace {
// Argument checks shared by the device-pointer pair calls.
int check_dev_pair(int algo, const uint32_t *d_a, const uint32_t *d_a_states, size_t na, const uint32_t *d_b,
                   const uint32_t *d_b_states, size_t nb, uint32_t *d_out, uint32_t *d_out_states,
                   unsigned long long *d_totals)
{
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (int rc = check_flags(algo, e)) return rc;
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
    return SETOPS_OK;
}
} // namespace

int setops_hip_intersect2_dev(int algo, const uint32_t *d_a, const uint32_t *d_a_states, size_t na, const uint32_t *d_b,
                              const uint32_t *d_b_states, size_t nb, uint32_t *d_out, uint32_t *d_out_states,
                              size_t out_cap, unsigned long long *d_totals, void *stream)
{
    if (int rc = check_dev_pair(algo, d_a, d_a_states, na, d_b, d_b_states, nb, d_out, d_out_states, d_totals))
        return rc;
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL = the HIP null stream
    return single_launch(algo, d_a, d_a_states, na, d_b, d_b_states, nb, d_out, d_out_states, out_cap, d_totals, s);
}

int setops_hip_intersect2_dev_noalloc(int algo, const uint32_t *d_a, const uint32_t *d_a_states, size_t na,
                                      const uint32_t *d_b, const uint32_t *d_b_states, size_t nb, uint32_t *d_out,
                                      uint32_t *d_out_states, size_t out_cap, unsigned long long *d_totals,
                                      void *stream)
{
    if (int rc = check_dev_pair(algo, d_a, d_a_states, na, d_b, d_b_states, nb, d_out, d_out_states, d_totals))
        return rc;
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    return noalloc_launch(algo, d_a, d_a_states, na, d_b, d_b_states, nb, d_out, d_out_states, out_cap, d_totals,
                          static_cast<hipStream_t>(stream));
}

int setops_hip_intersect2_bsr(int algo, const uint32_t *bases_a, const uint32_t *states_a, size_t na,
                              const uint32_t *bases_b, const uint32_t *states_b, size_t nb, uint32_t *bases_out,
                              uint32_t *states_out, size_t out_cap, size_t *out_words, uint64_t *out_elems)
{
    const AlgoEntry *e = algo_entry(algo);
    if (!e) return fail(SETOPS_ERR_UNKNOWN_ALGO, "unknown algorithm id %d", algo);
    if (!(e->flags & SETOPS_ALGO_BSR)) return fail(SETOPS_ERR_UNSUPPORTED, "%s is not a BSR entry", e->name);
    if (int rc = check_flags(algo, e)) return rc;
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
    if (int rc = check_flags(algo, e)) return rc;
    if (k < 2) return fail(SETOPS_ERR_INVALID, "svs needs at least 2 sets (svs.rs:61)");
    if (!sets || !lens || !out_len) return fail(SETOPS_ERR_INVALID, "NULL argument");
    for (size_t s = 0; s < k; s++)
        if (lens[s] && !sets[s]) return fail(SETOPS_ERR_INVALID, "set %zu is NULL", s);
    ENSURE_INIT();
    std::lock_guard<std::mutex> lk(g_ctx.host_mu);
    if (e->kind == KIND_ADAPTIVE && k > 2) return run_adaptive(algo, sets, lens, k, out, out_cap, out_len);
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

int setops_hip_intersect2_i32(int algo, const int32_t *a, size_t na, const int32_t *b, size_t nb, int32_t *out,
                              size_t out_cap, size_t *out_len)
{
    return setops_hip_intersect2(algo | SETOPS_SIGNED_I32, reinterpret_cast<const uint32_t *>(a), na,
                                 reinterpret_cast<const uint32_t *>(b), nb, reinterpret_cast<uint32_t *>(out), out_cap,
                                 out_len);
}

int setops_hip_svs_i32(int algo, const int32_t *const *sets, const size_t *lens, size_t k, int32_t *out, size_t out_cap,
                       size_t *out_len)
{
    return setops_hip_svs(algo | SETOPS_SIGNED_I32, reinterpret_cast<const uint32_t *const *>(sets), lens, k,
                          reinterpret_cast<uint32_t *>(out), out_cap, out_len);
}

int setops_hip_batch_create(setops_hip_batch **handle, int algo, const uint32_t *d_a, const uint32_t *d_a_states,
                            uint64_t a_len, const uint64_t *d_a_off, const uint32_t *d_b, const uint32_t *d_b_states,
                            uint64_t b_len, const uint64_t *d_b_off, uint64_t npairs)
{
    ENSURE_INIT();
    return batch_create(handle, algo, d_a, d_a_states, a_len, d_a_off, nullptr, d_b, d_b_states, b_len, d_b_off,
                        nullptr, npairs);
}

int setops_hip_batch_create_gather(setops_hip_batch **handle, int algo, const uint32_t *d_a,
                                   const uint32_t *d_a_states, uint64_t a_len, const uint64_t *d_a_beg,
                                   const uint64_t *d_a_end, const uint32_t *d_b, const uint32_t *d_b_states,
                                   uint64_t b_len, const uint64_t *d_b_beg, const uint64_t *d_b_end, uint64_t npairs)
{
    ENSURE_INIT();
    if (!d_a_end || !d_b_end) return fail(SETOPS_ERR_INVALID, "end arrays must not be NULL");
    return batch_create(handle, algo, d_a, d_a_states, a_len, d_a_beg, d_a_end, d_b, d_b_states, b_len, d_b_beg,
                        d_b_end, npairs);
}

int setops_hip_batch_index_b(setops_hip_batch *handle)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (handle->b_idx) return SETOPS_OK;
    if (merge_like(handle->kind)) return SETOPS_OK;  // only the search policy reads an index
    const uint64_t m = (handle->in.b_pool + 31) / 32;
    const uint64_t mt = (handle->in.b_pool + kTopStride - 1) / kTopStride;
    const uint64_t mm = (handle->in.b_pool + kMidStride - 1) / kMidStride;
    if (m == 0) return SETOPS_OK;
    if (hipMalloc(&handle->b_idx, m * 4) != hipSuccess) {
        (void)hipGetLastError();
        handle->b_idx = nullptr;
        return fail(SETOPS_ERR_NOMEM, "hipMalloc of %llu index bytes failed", (unsigned long long)(m * 4));
    }
    if (hipMalloc(&handle->b_top, mt * 4) != hipSuccess || hipMalloc(&handle->b_mid, mm * 4) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(handle->b_idx);
        if (handle->b_top) (void)hipFree(handle->b_top);
        handle->b_idx = nullptr;
        handle->b_top = nullptr;
        handle->b_mid = nullptr;
        return fail(SETOPS_ERR_NOMEM, "hipMalloc of %llu index bytes failed", (unsigned long long)((mt + mm) * 4));
    }
    // the B pool may have been written by work still pending on another stream; the index build
    // waits for it (it synchronises at the end anyway)
    HIP_TRY(hipDeviceSynchronize());
    hipStream_t s = handle->ran ? handle->last_stream : g_ctx.stream;
    const uint32_t f = handle->in.sflip;  // the index holds values in the compare space
    hipLaunchKernelGGL(build_index_kernel, dim3(grid_for(m, 256)), dim3(256), 0, s, handle->in.B, handle->in.b_pool,
                       handle->b_idx, (uint64_t)32, f);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(build_index_kernel, dim3(grid_for(mt, 256)), dim3(256), 0, s, handle->in.B, handle->in.b_pool,
                       handle->b_top, kTopStride, f);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(build_index_kernel, dim3(grid_for(mm, 256)), dim3(256), 0, s, handle->in.B, handle->in.b_pool,
                       handle->b_mid, kMidStride, f);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    handle->in.b_idx = handle->b_idx;
    handle->in.b_top = handle->b_top;
    handle->in.b_mid = handle->b_mid;
    return SETOPS_OK;
}

int setops_hip_batch_set_further(setops_hip_batch *handle, const uint64_t *d_x_off, const uint64_t *d_x_beg,
                                 const uint64_t *d_x_end)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (!d_x_off && !d_x_beg && !d_x_end) {
        handle->in.x_off = nullptr;
        handle->in.x_beg = nullptr;
        handle->in.x_end = nullptr;
        return SETOPS_OK;
    }
    if (!d_x_off || !d_x_beg || !d_x_end) return fail(SETOPS_ERR_INVALID, "further-set arrays must all be set or all NULL");
    if (!handle->gather) return fail(SETOPS_ERR_UNSUPPORTED, "further sets need a gather batch");
    const AlgoEntry *e = algo_entry(handle->algo);
    if (merge_like(handle->kind) || handle->bsr)
        return fail(SETOPS_ERR_UNSUPPORTED,
                    "%s: further sets are looked up by the search entries on plain sets (hip_small_adaptive, "
                    "hip_galloping, hip_binary_search, hip_simd_galloping)", e->name);
    handle->in.x_off = d_x_off;
    handle->in.x_beg = d_x_beg;
    handle->in.x_end = d_x_end;
    return SETOPS_OK;
}

int setops_hip_batch_rebind(setops_hip_batch *handle, const uint64_t *d_a_beg, const uint64_t *d_a_end,
                            const uint64_t *d_b_beg, const uint64_t *d_b_end, uint64_t npairs, const uint64_t *h_na,
                            const uint64_t *h_nb)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (!handle->gather) return fail(SETOPS_ERR_UNSUPPORTED, "only gather batches can be rebound");
    if (npairs == 0) return fail(SETOPS_ERR_INVALID, "batch has no pairs");
    if (!d_a_beg || !d_a_end || !d_b_beg || !d_b_end || !h_na || !h_nb)
        return fail(SETOPS_ERR_INVALID, "NULL argument");
    uint64_t own = 0, min_own = ~0ull;
    for (uint64_t p = 0; p < npairs; p++) {
        if (h_na[p] >= 0xFFFFFFFFull || h_nb[p] >= 0xFFFFFFFFull)
            return fail(SETOPS_ERR_INVALID, "pair %llu: sets must hold < 2^32 elements", (unsigned long long)p);
        const uint64_t u = merge_like(handle->kind) ? h_na[p] + h_nb[p] : std::min(h_na[p], h_nb[p]);
        own += u;
        min_own = std::min(min_own, u);
    }
    const uint32_t c0 = pad_for_batch(handle->kind, handle->bsr, handle->gather, handle->shape, min_own);
    int rc = batch_reserve(handle, npairs, own + (uint64_t)c0 * npairs, c0);
    if (rc != SETOPS_OK) return rc;
    handle->P = npairs;
    handle->in.P = npairs;
    handle->in.a_beg = d_a_beg;
    handle->in.a_end = d_a_end;
    handle->in.b_beg = d_b_beg;
    handle->in.b_end = d_b_end;
    return SETOPS_OK;
}

int setops_hip_batch_reserve(setops_hip_batch *handle, uint64_t npairs, uint64_t units)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (npairs >= 0xFFFFFFFFull) return fail(SETOPS_ERR_INVALID, "too many pairs: %llu", (unsigned long long)npairs);
    setops_hip_batch *h = handle;
    const uint64_t U = units_per_tile(h->kind, h->bsr, h->shape);
    const uint64_t pad = pad_units(h->kind, h->bsr, h->gather, h->shape);
    const uint64_t ntiles = std::max<uint64_t>(1, (units + pad * std::max<uint64_t>(npairs, 1) + U - 1) / U);
    if (ntiles >= 0x7FFFFFFFull) return fail(SETOPS_ERR_INVALID, "too many units: %llu", (unsigned long long)units);
    if (!h->ws || npairs > h->cap_pairs || ntiles > h->cap_tiles) {
        // ws_layout repoints the plan's arrays; every run rewrites them, and a cached launch graph
        // keyed on the old pointers is not replayed
        if (h->ws) {
            if (h->ran) HIP_TRY(hipStreamSynchronize(h->last_stream));
            HIP_TRY(hipFree(h->ws));
            h->ws = nullptr;
        }
        const uint64_t cp = std::max<uint64_t>(npairs, h->cap_pairs), ct = std::max<uint64_t>(ntiles, h->cap_tiles);
        int rc = batch_alloc(h, cp, ct);
        if (rc != SETOPS_OK) return rc;
        h->cap_pairs = cp;
        h->cap_tiles = ct;
    }
    const uint32_t *keep = h->plan.stage;
    int rc = stage_grow(h, ntiles);
    if (rc != SETOPS_OK) return rc;
    if (h->plan.direct) h->plan.stage = const_cast<uint32_t *>(keep);
    return SETOPS_OK;
}

int setops_hip_batch_step(setops_hip_batch *handle, const uint64_t *h_idx, uint64_t *d_idx, uint64_t npairs,
                          int64_t nfurther, const uint64_t *h_na, const uint64_t *h_nb, uint32_t *d_out,
                          uint64_t out_cap, uint64_t *d_out_off, uint32_t *d_counts, unsigned long long *d_totals,
                          uint64_t *h_out_off, uint32_t *h_out, uint64_t copy_words, void *stream)
{
    if (!handle) return fail(SETOPS_ERR_INVALID, "handle is NULL");
    if (!d_idx || !d_out_off || (copy_words && (!h_out || !d_out)))
        return fail(SETOPS_ERR_INVALID, "NULL argument");
    if (copy_words > out_cap) return fail(SETOPS_ERR_INVALID, "copy_words exceeds out_cap");
    if (npairs == 0) return fail(SETOPS_ERR_INVALID, "batch has no pairs");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t P = npairs;
    const uint64_t words = 4 * P + (nfurther >= 0 ? P + 1 + 2 * (uint64_t)nfurther : 0);
    if (h_idx) HIP_TRY(hipMemcpyAsync(d_idx, h_idx, words * 8, hipMemcpyHostToDevice, s));
    int rc = setops_hip_batch_rebind(handle, d_idx, d_idx + P, d_idx + 2 * P, d_idx + 3 * P, P, h_na, h_nb);
    if (rc != SETOPS_OK) return rc;
    if (nfurther >= 0) {
        const uint64_t *xo = d_idx + 4 * P;
        rc = setops_hip_batch_set_further(handle, xo, xo + P + 1, xo + P + 1 + nfurther);
    } else {
        rc = setops_hip_batch_set_further(handle, nullptr, nullptr, nullptr);
    }
    if (rc != SETOPS_OK) return rc;
    rc = batch_launch(handle, d_out, nullptr, out_cap, d_out_off, d_counts, d_totals, s);
    if (rc != SETOPS_OK) return rc;
    if (h_out_off) HIP_TRY(hipMemcpyAsync(h_out_off, d_out_off, (P + 1) * 8, hipMemcpyDeviceToHost, s));
    if (copy_words) HIP_TRY(hipMemcpyAsync(h_out, d_out, copy_words * 4, hipMemcpyDeviceToHost, s));
    return SETOPS_OK;
}

// ---- native request loop ---------------------------------------------------------------------------

namespace {

constexpr uint64_t kStepSlackWords = 64;  // result words a step holds beyond its queries' need

// Packed step rows of query q (setops_hip_batch_step's layout): 4 pair bounds and 1 further-set
// offset per query, 2 words per further set, plus the closing offset of the step.
uint64_t query_rows(const setops_hip_serve_queries *qs, uint64_t q)
{
    return 5 + 2 * (uint64_t)(qs->xoff[q + 1] - qs->xoff[q]);
}

uint64_t query_need(const setops_hip_serve_queries *qs, uint64_t q)
{
    return (uint64_t)(qs->need_pfx[q + 1] - qs->need_pfx[q]);
}

struct ServeState {
    hipEvent_t done = nullptr;
    uint64_t q0 = 0, q1 = 0;  // queries of the step in flight on the slot (none: q0 == q1)
    int64_t cost = 0;          // their admission cost
    // the step's offsets and results, written by the kernels straight into mapped host memory
    // (no copies back): max_pairs + 1 offsets, then out_words result words
    void *host = nullptr;
    uint64_t *h_off = nullptr, *d_off = nullptr;
    uint32_t *h_res = nullptr, *d_res = nullptr;
    // the step's packed rows, read by the kernels from mapped host memory too (no upload)
    void *rows_host = nullptr;
    uint64_t *h_rows = nullptr, *d_rows = nullptr;
};

// Writes a step's packed rows for queries q0 .. q1-1 to h (the slot's mapped rows).
void pack_step(uint64_t *h, const setops_hip_serve_queries *qs, uint64_t q0, uint64_t q1)
{
    const uint64_t P = q1 - q0, nq = qs->nq;
    const int64_t x0 = qs->xoff[q0], X = qs->xoff[q1] - x0;
    for (int r = 0; r < 4; r++) std::memcpy(h + r * P, qs->pairs + r * nq + q0, P * 8);
    for (uint64_t j = 0; j <= P; j++) h[4 * P + j] = (uint64_t)(qs->xoff[q0 + j] - x0);
    std::memcpy(h + 5 * P + 1, qs->xb + x0, (size_t)X * 8);
    std::memcpy(h + 5 * P + 1 + X, qs->xe + x0, (size_t)X * 8);
}

}  // namespace

int setops_hip_serve(const setops_hip_serve_slot *slots, int nslots, const setops_hip_serve_queries *qs,
                     int64_t budget_words, uint64_t max_pairs, double timeout_s, double *done_t, uint32_t *results,
                     uint64_t *counts, uint64_t *stats)
{
    if (!slots || !qs || !done_t || !stats || (results && !counts)) return fail(SETOPS_ERR_INVALID, "NULL argument");
    if (nslots < 2) return fail(SETOPS_ERR_INVALID, "the loop needs at least two slots (got %d)", nslots);
    if (max_pairs == 0) return fail(SETOPS_ERR_INVALID, "max_pairs is 0");
    const uint64_t nq = qs->nq;
    if (nq && (!qs->pairs || !qs->na || !qs->nb || !qs->xoff || !qs->need_pfx || !qs->cost || !qs->arrivals ||
               (qs->xoff[nq] > qs->xoff[0] && (!qs->xb || !qs->xe))))
        return fail(SETOPS_ERR_INVALID, "NULL query array");
    uint64_t idx_min = ~0ull, out_min = ~0ull;
    for (int i = 0; i < nslots; i++) {
        const setops_hip_serve_slot &sl = slots[i];
        if (!sl.handle || !sl.d_totals || sl.max_pairs == 0)
            return fail(SETOPS_ERR_INVALID, "slot %d: NULL buffer or no pair slots", i);
        idx_min = std::min(idx_min, sl.idx_words);
        out_min = std::min(out_min, sl.out_words);
    }
    // every query must fit a step of its own on any slot, and the budget
    for (uint64_t q = 0; q < nq; q++) {
        if (qs->xoff[q + 1] < qs->xoff[q] || qs->need_pfx[q + 1] < qs->need_pfx[q] || qs->cost[q] < 0)
            return fail(SETOPS_ERR_INVALID, "query %llu: offsets decrease or cost is negative", (unsigned long long)q);
        if (q && qs->arrivals[q] < qs->arrivals[q - 1])
            return fail(SETOPS_ERR_INVALID, "arrivals are not ascending at query %llu", (unsigned long long)q);
        if (1 + query_rows(qs, q) > idx_min || query_need(qs, q) + kStepSlackWords > out_min)
            return fail(SETOPS_ERR_CAPACITY, "query %llu does not fit a slot's step buffers", (unsigned long long)q);
        if (qs->cost[q] > budget_words)
            return fail(SETOPS_ERR_CAPACITY, "query %llu needs more than the admission budget", (unsigned long long)q);
    }
    for (uint64_t q = 0; q < nq; q++) done_t[q] = std::numeric_limits<double>::quiet_NaN();

    std::vector<ServeState> st((size_t)nslots);
    // waits for the steps still in flight and frees the events and the mapped buffers
    auto finish = [&](int rc) {
        for (int i = 0; i < nslots; i++) {
            if (st[i].q1 > st[i].q0) (void)hipEventSynchronize(st[i].done);
            if (st[i].done) (void)hipEventDestroy(st[i].done);
            if (st[i].host) (void)hipHostFree(st[i].host);
            if (st[i].rows_host) (void)hipHostFree(st[i].rows_host);
        }
        return rc;
    };
    for (int i = 0; i < nslots; i++) {
        ServeState &s = st[i];
        if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
            s.done = nullptr;
            return finish(fail(SETOPS_ERR_HIP, "hipEventCreateWithFlags failed"));
        }
        const size_t off_bytes = ((size_t)(slots[i].max_pairs + 1) * 8 + 255) & ~(size_t)255;
        const size_t bytes = off_bytes + (size_t)slots[i].out_words * 4;
        if (hipHostMalloc(&s.host, bytes, hipHostMallocMapped) != hipSuccess) {
            (void)hipGetLastError();
            s.host = nullptr;
            return finish(fail(SETOPS_ERR_NOMEM, "hipHostMalloc of %zu mapped bytes failed", bytes));
        }
        void *dev = nullptr;
        if (hipHostGetDevicePointer(&dev, s.host, 0) != hipSuccess || !dev)
            return finish(fail(SETOPS_ERR_HIP, "hipHostGetDevicePointer failed"));
        s.h_off = static_cast<uint64_t *>(s.host);
        s.d_off = static_cast<uint64_t *>(dev);
        s.h_res = reinterpret_cast<uint32_t *>(static_cast<char *>(s.host) + off_bytes);
        s.d_res = reinterpret_cast<uint32_t *>(static_cast<char *>(dev) + off_bytes);
        if (hipHostMalloc(&s.rows_host, (size_t)slots[i].idx_words * 8, hipHostMallocMapped) != hipSuccess) {
            (void)hipGetLastError();
            s.rows_host = nullptr;
            return finish(fail(SETOPS_ERR_NOMEM, "hipHostMalloc of the step rows failed"));
        }
        void *rdev = nullptr;
        if (hipHostGetDevicePointer(&rdev, s.rows_host, 0) != hipSuccess || !rdev)
            return finish(fail(SETOPS_ERR_HIP, "hipHostGetDevicePointer failed"));
        s.h_rows = static_cast<uint64_t *>(s.rows_host);
        s.d_rows = static_cast<uint64_t *>(rdev);
    }

    using clock = std::chrono::steady_clock;
    const clock::time_point t0 = clock::now();
    auto since = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    uint64_t head = 0, nxt = 0, ndone = 0, steps = 0, launched = 0, active = 0;
    int64_t reserved = 0;
    int cur = 0;
    while (ndone < nq) {
        const double now = since();
        if (now > timeout_s) break;
        while (nxt < nq && qs->arrivals[nxt] <= now) nxt++;
        const setops_hip_serve_slot &sl = slots[cur];
        ServeState &ss = st[cur];
        // this slot is free (its last step was finished when it was the oldest): admit the
        // arrived queries in arrival order while they fit the budget, the pair slots and the
        // slot's step buffers (the first that does not fit stops the admission)
        if (head < nxt) {
            uint64_t e = head, rows = 1, need = kStepSlackWords;
            int64_t cost = 0;
            while (e < nxt && active + (e - head) < max_pairs && (e - head) < sl.max_pairs) {
                const int64_t c = qs->cost[e];
                const uint64_t r = query_rows(qs, e), n = query_need(qs, e);
                if (reserved + cost + c > budget_words || rows + r > sl.idx_words || need + n > sl.out_words) break;
                cost += c;
                rows += r;
                need += n;
                e++;
            }
            if (e > head) {
                pack_step(ss.h_rows, qs, head, e);
                const uint64_t P = e - head;
                const int64_t X = qs->xoff[e] - qs->xoff[head];
                int rc = setops_hip_batch_step(sl.handle, nullptr, ss.d_rows, P, X, qs->na + head, qs->nb + head,
                                               ss.d_res, sl.out_words, ss.d_off, nullptr, sl.d_totals,
                                               nullptr, nullptr, 0, sl.stream);
                if (rc != SETOPS_OK) return finish(rc);
                hipStream_t s = static_cast<hipStream_t>(sl.stream);
                if (hipEventRecord(ss.done, s) != hipSuccess) return finish(fail(SETOPS_ERR_HIP, "hipEventRecord failed"));
                ss.q0 = head;
                ss.q1 = e;
                ss.cost = cost;
                reserved += cost;
                active += P;
                head = e;
                steps++;
                launched += P;
            }
        }
        // the oldest step in flight: wait for it and finish its queries
        const int old = (cur + 1) % nslots;
        ServeState &os = st[old];
        if (os.q1 > os.q0) {
            if (hipEventSynchronize(os.done) != hipSuccess)
                return finish(fail(SETOPS_ERR_HIP, "hipEventSynchronize failed"));
            const double t_end = since();
            const setops_hip_serve_slot &ol = slots[old];
            for (uint64_t q = os.q0; q < os.q1; q++) {
                done_t[q] = t_end;
                if (results) {
                    const uint64_t j = q - os.q0, o = os.h_off[j], n = os.h_off[j + 1] - o;
                    if (os.h_off[j + 1] < o || n > query_need(qs, q) || o + n > ol.out_words)
                        return finish(fail(SETOPS_ERR_CAPACITY, "query %llu: result larger than its bound",
                                           (unsigned long long)q));
                    std::memcpy(results + qs->need_pfx[q], os.h_res + o, n * 4);
                    counts[q] = n;
                }
            }
            ndone += os.q1 - os.q0;
            active -= os.q1 - os.q0;
            reserved -= os.cost;
            os.q0 = os.q1 = 0;
        } else if (head == nxt && nxt < nq &&
                   std::all_of(st.begin(), st.end(), [](const ServeState &x) { return x.q1 == x.q0; })) {
            // nothing in flight and nothing arrived: wait for the next arrival (sleep while it is
            // far, spin for the last 100 us)
            const double wait = qs->arrivals[nxt] - since();
            if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait - 1e-4, 1e-3)));
        }
        cur = old;
    }
    const double wall = since();
    stats[0] = steps;
    stats[1] = launched;
    stats[2] = ndone;
    stats[3] = (uint64_t)(wall * 1e9);
    return finish(SETOPS_OK);
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
    if (!handle->ran) return SETOPS_OK;
    uint32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, handle->d_error, 4, hipMemcpyDeviceToHost, handle->last_stream));
    HIP_TRY(hipStreamSynchronize(handle->last_stream));
    if (err & 4u)
        return fail(SETOPS_ERR_INVALID, "the pair sizes on the device differ from the plan (setops_hip_batch_rebind "
                                         "sizes must equal the device offsets); results are incomplete");
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
    if (grid) *grid = handle->plan.ntiles;
    if (units) *units = handle->plan.utotal;
    return SETOPS_OK;
}

int setops_hip_batch_destroy(setops_hip_batch *handle)
{
    batch_free(handle);
    return SETOPS_OK;
}

} // extern "C"
