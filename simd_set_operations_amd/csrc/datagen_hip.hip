// datagen_hip.hip -- synthetic sorted lists generated directly in device memory (gfx950).
//
// Workload generation for the served k-set corpus (2^20 lists, tens of GB): generating it on the
// host and uploading it would take minutes.  Not part of the intersection library's ABI.
//
// List i has length L_i = len[i] over [0, U) and starts at out + beg[i].  Its values are L_i
// order statistics of U(0,1) made distinct: with E_1 .. E_{L+1} i.i.d. Exp(1) and S_k = E_1 + ... + E_k, the ratios
// S_k / S_{L+1} (k = 1..L) are distributed as the sorted uniform sample (the exponential-spacings
// construction, Devroye, "Non-Uniform Random Variate Generation", 1986, V.2); element k is
// floor(S_k / S_{L+1} * (U - L + 1)) + (k - 1), strictly increasing and below U.  E_k comes from
// the k-th output of a SplitMix64 stream seeded per list (counter-based: any thread can produce
// any E_k, and the corpus is the same on every run).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t list_seed(uint64_t master, uint64_t index)
{
    // same stream derivation as the host generator (datagen.cpp stream_seed)
    uint64_t s = master ^ (index * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
    s += 0x9E3779B97F4A7C15ull;  // first output discarded
    s += 0x9E3779B97F4A7C15ull;
    return mix64(s);
}

// E_k, k >= 1: -log of a uniform in (0, 1)
__device__ __forceinline__ double expo(uint64_t seed, uint64_t k)
{
    const uint64_t z = mix64(seed + k * 0x9E3779B97F4A7C15ull);
    const double u = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return -log(u);
}

template <int NT>
__global__ __launch_bounds__(NT) void gen_sorted_lists_kernel(uint64_t master, uint64_t universe, const uint64_t *beg,
                                                              const uint64_t *len, uint64_t nlists, uint32_t *out)
{
    __shared__ double red[NT / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (uint64_t i = blockIdx.x; i < nlists; i += gridDim.x) {
        const uint64_t L = len[i];
        uint32_t *dst = out + beg[i];
        const uint64_t seed = list_seed(master, i);
        // pass 1: S_{L+1}
        double part = 0.0;
        for (uint64_t k = 1 + tid; k <= L + 1; k += NT) part += expo(seed, k);
        for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
        if (lane == 0) red[w] = part;
        __syncthreads();
        double S = 0.0;
        for (int k = 0; k < NT / 64; k++) S += red[k];
        __syncthreads();
        const double scale = (double)(universe - L + 1) / S;
        const uint64_t xmax = universe - L;
        // pass 2: running sums in chunks of NT, each element written once
        double carry = 0.0;
        for (uint64_t k0 = 1; k0 <= L; k0 += NT) {
            const uint64_t k = k0 + tid;
            double x = (k <= L) ? expo(seed, k) : 0.0;
            for (int d = 1; d < 64; d <<= 1) {
                const double y = __shfl_up(x, d, 64);
                if (lane >= (uint32_t)d) x += y;
            }
            if (lane == 63) red[w] = x;
            __syncthreads();
            double pre = carry, all = carry;
            for (int q = 0; q < NT / 64; q++) {
                if (q < (int)w) pre += red[q];
                all += red[q];
            }
            if (k <= L) {
                uint64_t v = (uint64_t)((pre + x) * scale);
                if (v > xmax) v = xmax;
                dst[k - 1] = (uint32_t)(v + (k - 1));
            }
            carry = all;
            __syncthreads();
        }
    }
}

} // namespace

extern "C" {

// Fills out[beg[i] .. beg[i] + len[i]) with list i for every i < nlists (device pointers, on
// `stream`).  Requires len[i] <= universe <= 2^32.  Returns 0 or a negative HIP error.
int setops_gen_sorted_lists_dev(uint64_t seed, uint64_t universe, const uint64_t *d_beg, const uint64_t *d_len,
                                uint64_t nlists, uint32_t *d_out, void *stream)
{
    if (universe == 0 || universe > (1ull << 32) || !d_beg || !d_len || (nlists && !d_out)) return -1;
    const uint32_t grid = (uint32_t)(nlists < (1u << 20) ? (nlists ? nlists : 1) : (1u << 20));
    hipLaunchKernelGGL(gen_sorted_lists_kernel<256>, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), seed,
                       universe, d_beg, d_len, nlists, d_out);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // extern "C"
