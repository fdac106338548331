// Stable LSD radix sort of (u32 key, u32 value) pairs on the GPU, for the photon store (records
// by texel key) and the k-nearest hash grid (records by bucket). Equal keys keep their input
// order, so with values 0..n-1 the output is ordered by (key, input index).
//
// One pass per 8-bit digit of the key's low `key_bits` bits, three launches per pass:
//   k_radix_hist     each 4096-key tile counts its digits (LDS counters) into the digit-major
//                    table hist[digit * tiles + tile];
//   scan             exclusive prefix sum of that table (chunk scans, a scan of the chunk sums,
//                    an add): hist[d * tiles + t] becomes the first output position of tile t's
//                    keys with digit d;
//   k_radix_scatter  each tile ranks its keys: a wavefront takes 64 consecutive keys at a time,
//                    finds the lanes that share its digit with eight ballots, and ranks a key by
//                    the equal-digit lanes below it plus the wavefront's running count for the
//                    digit (one lane per digit updates the count; a wavefront's LDS operations
//                    complete in order). Wavefronts' counts are scanned in tile order, so keys of
//                    one digit leave a tile in input order.
// A pass reads the pairs twice and writes them once (24 B per pair).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "rl_sort.h"

namespace rl {

namespace {

constexpr int kSortThreads = 256;               // four wavefronts
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;                  // keys per thread
constexpr uint32_t kSortTile = kSortThreads * kSortItems;
constexpr int kRadix = 256;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Key index of item j of this thread within its tile: wavefront-major, then item, then lane, so
// a wavefront's loads are contiguous and its items run through the tile in input order.
__device__ __forceinline__ uint32_t tile_index(int j) {
    const uint32_t w = threadIdx.x >> 6;
    return w * (64u * kSortItems) + (uint32_t)j * 64u + lane_id();
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                             int shift, uint32_t tiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kRadix];
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint64_t i = base + tile_index(j);
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of `data` (length m) in chunks of kScanChunk; chunk c's total goes to sums[c].
__global__ __launch_bounds__(kScanThreads) void k_scan_chunks(uint32_t* __restrict__ data, uint64_t m,
                                                              uint32_t* __restrict__ sums) {
    __shared__ uint32_t part[kScanThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        v[j] = base + j < m ? data[base + j] : 0u;
        s += v[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums.
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const uint32_t x = threadIdx.x >= (uint32_t)d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        if (base + j < m) data[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = part[threadIdx.x];
}

// Exclusive scan of the chunk sums in one workgroup (any count, in kScanThreads-wide rounds).
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t count) {
    __shared__ uint32_t part[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < count; r0 += kScanThreads) {
        const uint32_t i = r0 + threadIdx.x;
        const uint32_t v = i < count ? sums[i] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < kScanThreads; d <<= 1) {
            const uint32_t x = threadIdx.x >= (uint32_t)d ? part[threadIdx.x - d] : 0u;
            __syncthreads();
            part[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < count) sums[i] = carry + part[threadIdx.x] - v;
        const uint32_t total = part[kScanThreads - 1];
        __syncthreads();
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* __restrict__ data, uint64_t m,
                                                           const uint32_t* __restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
    const uint32_t add = sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; j++)
        if (base + j < m) data[base + j] += add;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ keys,
                                                                const uint32_t* __restrict__ vals, uint32_t n,
                                                                int shift, uint32_t tiles,
                                                                const uint32_t* __restrict__ offsets,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t wcnt[kSortWaves][kRadix];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    for (int d = threadIdx.x; d < kSortWaves * kRadix; d += kSortThreads) (&wcnt[0][0])[d] = 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    uint32_t k[kSortItems], v[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint64_t i = base + tile_index(j);
        const bool valid = i < n;
        k[j] = valid ? keys[i] : 0u;
        v[j] = valid ? vals[i] : 0u;
        const uint32_t d = (k[j] >> shift) & 0xFFu;
        // Lanes holding a valid key with this lane's digit.
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t ones = __ballot(valid && bit);
            same &= bit ? ones : ~ones;
        }
        const uint32_t before = wcnt[w][d];
        rank[j] = before + (uint32_t)__popcll(same & below);
        // The highest lane of each digit group advances the wavefront's count for that digit.
        if (valid && (same >> lane) == 1ull) wcnt[w][d] = before + (uint32_t)__popcll(same);
    }
    __syncthreads();
    // Per digit: the tile's first output position, then the wavefronts' counts in tile order.
    {
        const uint32_t d = threadIdx.x;
        uint32_t run = offsets[(size_t)d * tiles + blockIdx.x];
        for (int ww = 0; ww < kSortWaves; ww++) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint64_t i = base + tile_index(j);
        if (i < n) {
            const uint32_t pos = wcnt[w][(k[j] >> shift) & 0xFFu] + rank[j];
            keys_out[pos] = k[j];
            vals_out[pos] = v[j];
        }
    }
}

uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

uint64_t hist_len(uint32_t n) { return (uint64_t)kRadix * sort_tiles(n); }

uint32_t scan_chunks(uint64_t m) { return (uint32_t)((m + kScanChunk - 1) / kScanChunk); }

}  // namespace

size_t sort_pairs_temp_bytes(uint32_t n) {
    const uint64_t m = hist_len(n);
    return (size_t)(m + scan_chunks(m)) * sizeof(uint32_t);
}

hipError_t launch_sort_pairs(hipStream_t st, uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                             uint32_t n, int key_bits, void* temp, size_t temp_bytes, uint32_t** keys_out,
                             uint32_t** vals_out) {
    *keys_out = keys;
    *vals_out = vals;
    if (n == 0) return hipSuccess;
    if (key_bits < 1 || key_bits > 32 || temp_bytes < sort_pairs_temp_bytes(n)) return hipErrorInvalidValue;
    const uint32_t tiles = sort_tiles(n);
    const uint64_t m = hist_len(n);
    const uint32_t chunks = scan_chunks(m);
    uint32_t* hist = static_cast<uint32_t*>(temp);
    uint32_t* sums = hist + m;
    uint32_t *ka = keys, *kb = keys_alt, *va = vals, *vb = vals_alt;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, st, ka, n, shift, tiles, hist);
        hipLaunchKernelGGL(k_scan_chunks, dim3(chunks), dim3(kScanThreads), 0, st, hist, m, sums);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, sums, chunks);
        hipLaunchKernelGGL(k_scan_add, dim3(chunks), dim3(kScanThreads), 0, st, hist, m, sums);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, st, ka, va, n, shift, tiles, hist, kb,
                           vb);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint32_t* t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    *keys_out = ka;
    *vals_out = va;
    return hipSuccess;
}

}  // namespace rl
