// Device-wide scan and sort used by the step (see prims.h). Large inputs go
// to hipCUB's scan and radix sort; 64-bit keys up to kSmallSortMax are
// sorted by one workgroup in LDS (a bitonic network), which replaces the
// eight launches of a radix sort over 64 bits with one.
#include "prims.h"
#include <hipcub/hipcub.hpp>

namespace b2g {

namespace {

constexpr int kSmallSortMax = 4096;
constexpr int kSmallSortThreads = 1024;

// Bitonic sort of n <= kSmallSortMax keys, padded to a power of two with the
// largest key (which sorts last and is not written back).
__global__ void __launch_bounds__(kSmallSortThreads) k_sort_small_u64(const unsigned long long* in,
                                                                      unsigned long long* out, int n) {
    __shared__ unsigned long long a[kSmallSortMax];
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += blockDim.x) a[i] = i < n ? in[i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int p = i ^ j;
                if (p > i) {
                    const unsigned long long x = a[i], y = a[p];
                    const bool ascending = (i & k) == 0;
                    if ((x > y) == ascending) {
                        a[i] = y;
                        a[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = a[i];
}

// Stable counting sort of (key, value) pairs with keys < 2^kCountBits, in
// tiles of kCountTile items (one per lane): each tile's per-key counts,
// an exclusive scan over (key, tile) in key-major order, then a scatter in
// which an item's place within its tile is the number of earlier items of
// the tile with its key (lanes with an equal key found by one ballot per key
// bit, then the earlier waves' counts from LDS).
constexpr int kCountBits = 7;
constexpr int kCountKeys = 1 << kCountBits;
constexpr int kCountTile = 256;
constexpr int kCountWaves = kCountTile / 64;

// Lanes of this wavefront holding the same key as this lane, and how many of
// them come before it.
__device__ inline unsigned long long sameKeyLanes(uint32_t key, bool valid) {
    unsigned long long mask = __ballot(valid);
    for (int b = 0; b < kCountBits; ++b) {
        const bool bit = (key >> b) & 1u;
        const unsigned long long ones = __ballot(valid && bit);
        mask &= bit ? ones : ~ones;
    }
    return mask;
}

__device__ inline void tileCounts(const uint32_t* keys, int n, int* waveCount /* [kCountWaves][kCountKeys] */,
                                  uint32_t& key, bool& valid, int& rankInWave) {
    const int i = blockIdx.x * kCountTile + threadIdx.x;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    valid = i < n;
    key = valid ? keys[i] & (kCountKeys - 1) : 0u;
    for (int k = threadIdx.x; k < kCountWaves * kCountKeys; k += kCountTile) waveCount[k] = 0;
    __syncthreads();
    const unsigned long long same = sameKeyLanes(key, valid);
    rankInWave = __popcll(same & ((1ull << lane) - 1ull));
    if (valid && rankInWave == 0) waveCount[wave * kCountKeys + key] = __popcll(same);
    __syncthreads();
}

__global__ void __launch_bounds__(kCountTile) k_count_tiles(const uint32_t* keys, int n, int tiles, int* counts,
                                                            int* totals, int totalsKeys) {
    __shared__ int waveCount[kCountWaves * kCountKeys];
    uint32_t key;
    bool valid;
    int rank;
    tileCounts(keys, n, waveCount, key, valid, rank);
    for (int k = threadIdx.x; k < kCountKeys; k += kCountTile) {
        int c = 0;
        for (int w = 0; w < kCountWaves; ++w) c += waveCount[w * kCountKeys + k];
        counts[k * tiles + blockIdx.x] = c;
        if (totals && c && k < totalsKeys) atomicAdd(&totals[k], c);
    }
}

__global__ void __launch_bounds__(kCountTile) k_count_scatter(const uint32_t* keys, const int* vals, int n, int tiles,
                                                              const int* offsets, uint32_t* keysOut, int* valsOut) {
    __shared__ int waveCount[kCountWaves * kCountKeys];
    uint32_t key;
    bool valid;
    int rank;
    tileCounts(keys, n, waveCount, key, valid, rank);
    if (!valid) return;
    const int wave = threadIdx.x / 64;
    for (int w = 0; w < wave; ++w) rank += waveCount[w * kCountKeys + key];
    const int i = blockIdx.x * kCountTile + threadIdx.x;
    const int dst = offsets[key * tiles + blockIdx.x] + rank;
    keysOut[dst] = keys[i];
    valsOut[dst] = vals[i];
}

}  // namespace

size_t prims_scratch_bytes(int n) {
    size_t a = 0, b = 0, c = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, (const int*)nullptr, (int*)nullptr, n);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, n, 0, 32);
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, c, (const unsigned long long*)nullptr, (unsigned long long*)nullptr, n,
                                      0, 64);
    size_t m = a > b ? a : b;
    m = m > c ? m : c;
    const size_t d = prims_count_sort_bytes(n);
    m = m > d ? m : d;
    return m + 256;
}

hipError_t prims_exclusive_sum(void* scratch, size_t bytes, const int* in, int* out, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return hipcub::DeviceScan::ExclusiveSum(scratch, bytes, in, out, n, s);
}

hipError_t prims_sort_pairs_u32(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                int* vout, int n, int bits, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return hipcub::DeviceRadixSort::SortPairs(scratch, bytes, kin, kout, vin, vout, n, 0, bits, s);
}

size_t prims_count_sort_bytes(int n) {
    const int tiles = (n + kCountTile - 1) / kCountTile;
    size_t scan = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const int*)nullptr, (int*)nullptr, tiles * kCountKeys);
    return 2 * sizeof(int) * (size_t)tiles * kCountKeys + scan + 512;
}

hipError_t prims_count_sort_pairs(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                  int* vout, int n, int* totals, int totalsKeys, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int tiles = (n + kCountTile - 1) / kCountTile;
    const size_t countBytes = (sizeof(int) * (size_t)tiles * kCountKeys + 255) & ~(size_t)255;
    if (bytes < prims_count_sort_bytes(n)) return hipErrorInvalidValue;
    int* counts = (int*)scratch;
    int* offsets = (int*)((char*)scratch + countBytes);
    void* scanTmp = (char*)scratch + 2 * countBytes;
    size_t scanBytes = bytes - 2 * countBytes;
    hipLaunchKernelGGL(k_count_tiles, dim3(tiles), dim3(kCountTile), 0, s, kin, n, tiles, counts, totals, totalsKeys);
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(scanTmp, scanBytes, counts, offsets, tiles * kCountKeys, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_count_scatter, dim3(tiles), dim3(kCountTile), 0, s, kin, vin, n, tiles, offsets, kout, vout);
    return hipGetLastError();
}

hipError_t prims_sort_keys_u64(void* scratch, size_t bytes, const unsigned long long* kin, unsigned long long* kout,
                               int n, hipStream_t s, int endBit) {
    if (n <= 0) return hipSuccess;
    if (n <= kSmallSortMax) {
        hipLaunchKernelGGL(k_sort_small_u64, dim3(1), dim3(kSmallSortThreads), 0, s, kin, kout, n);
        return hipGetLastError();
    }
    return hipcub::DeviceRadixSort::SortKeys(scratch, bytes, kin, kout, n, 0, endBit, s);
}

}  // namespace b2g
