// Device-wide scan and sorts used by the step (see prims.h), written for
// gfx950's 64-lane wavefronts:
//  * exclusive scan: per-tile sums, then every tile scans itself, starting
//    from the sum of the tiles before it (two launches);
//  * stable LSD radix sort of 32- or 64-bit keys (with or without int
//    values), 8-bit digits: per pass, each 256-item tile counts its digits,
//    the counts are scanned digit-major (so equal digits keep tile order),
//    and each item goes to its digit's offset plus its rank among the equal
//    digits before it in the tile (lanes with an equal digit found by one
//    ballot per digit bit, then the earlier wavefronts' counts from LDS);
//  * the same kernels for one pass over 7-bit keys (the colour order);
//  * up to 4096 64-bit keys: one workgroup's merge sort in LDS; up to 2^17,
//    such chunks merged pairwise by merge-path passes.
#include "prims.h"

#include <algorithm>

namespace b2g {

namespace {

// ------------------------------------------------------------------ scan
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kScanWaves = kScanThreads / 64;

__device__ inline int waveInclusiveSum(int x) {
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// Sum over the workgroup (every lane gets it).
__device__ inline int blockSum(int v, int* sums) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) sums[wave] = v;
    __syncthreads();
    int total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) total += sums[w];
    __syncthreads();
    return total;
}

// Exclusive prefix of v over the workgroup's lanes in lane order.
__device__ inline int blockExclusiveSum(int v, int* sums) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = waveInclusiveSum(v);
    if (lane == 63) sums[wave] = x;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += sums[w];
    __syncthreads();
    return before + x - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(const int* in, int n, int* tileSums) {
    __shared__ int sums[kScanWaves];
    const int base = blockIdx.x * kScanTile;
    int v = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const int i = base + j * kScanThreads + (int)threadIdx.x;
        if (i < n) v += in[i];
    }
    v = blockSum(v, sums);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = v;
}

// out may alias in: each lane reads its items before writing them.
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const int* in, int* out, int n, const int* tileSums) {
    __shared__ int sums[kScanWaves];
    int offset = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kScanThreads) offset += tileSums[t];
    offset = blockSum(offset, sums);
    const int base = blockIdx.x * kScanTile + (int)threadIdx.x * kScanItems;
    int v[kScanItems];
    int total = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = base + j < n ? in[base + j] : 0;
        total += v[j];
    }
    int run = offset + blockExclusiveSum(total, sums);
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}

int scanTiles(int n) { return (n + kScanTile - 1) / kScanTile; }

size_t scanBytes(int n) { return ((sizeof(int) * (size_t)scanTiles(n)) + 255) & ~(size_t)255; }

hipError_t exclusiveSum(void* scratch, size_t bytes, const int* in, int* out, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (bytes < scanBytes(n)) return hipErrorInvalidValue;
    const int tiles = scanTiles(n);
    int* tileSums = (int*)scratch;
    if (tiles > 1) hipLaunchKernelGGL(k_scan_sums, dim3(tiles), dim3(kScanThreads), 0, s, in, n, tileSums);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanThreads), 0, s, in, out, n, tileSums);
    return hipGetLastError();
}

// ---------------------------------------------------------- tile sorting
// One item per lane, kSortTile items per tile (four wavefronts).
constexpr int kSortTile = 256;
constexpr int kSortWaves = kSortTile / 64;

template <typename K>
__device__ inline uint32_t digitOf(K key, int shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// Lanes of this wavefront holding the same digit as this lane (digits of
// kBits bits), from one ballot per bit.
template <int kBits>
__device__ inline unsigned long long sameDigitLanes(uint32_t digit, bool valid) {
    unsigned long long mask = __ballot(valid);
    for (int b = 0; b < kBits; ++b) {
        const bool bit = (digit >> b) & 1u;
        const unsigned long long ones = __ballot(valid && bit);
        mask &= bit ? ones : ~ones;
    }
    return mask;
}

// This lane's digit (the key's bits at `shift` under `mask`, mask <
// 2^kBits), validity and rank among the equal digits of its wavefront;
// waveCount[w * kDigits + d] = count of digit d in wavefront w.
template <int kBits, typename K>
__device__ inline void tileDigits(const K* keys, int n, int shift, uint32_t mask, int* waveCount, uint32_t& digit,
                                  bool& valid, int& rank) {
    constexpr int kDigits = 1 << kBits;
    const int i = blockIdx.x * kSortTile + (int)threadIdx.x;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    valid = i < n;
    digit = valid ? digitOf(keys[i], shift, mask) : 0u;
    for (int k = threadIdx.x; k < kSortWaves * kDigits; k += kSortTile) waveCount[k] = 0;
    __syncthreads();
    const unsigned long long same = sameDigitLanes<kBits>(digit, valid);
    rank = __popcll(same & ((1ull << lane) - 1ull));
    if (valid && rank == 0) waveCount[wave * kDigits + digit] = __popcll(same);
    __syncthreads();
}

// counts[d * tiles + tile]: items of the tile with digit d.
template <int kBits, typename K>
__global__ void __launch_bounds__(kSortTile) k_digit_counts(const K* keys, int n, int shift, uint32_t mask, int tiles,
                                                            int* counts) {
    constexpr int kDigits = 1 << kBits;
    __shared__ int waveCount[kSortWaves * kDigits];
    uint32_t digit;
    bool valid;
    int rank;
    tileDigits<kBits>(keys, n, shift, mask, waveCount, digit, valid, rank);
    for (int d = threadIdx.x; d < kDigits; d += kSortTile) {
        int c = 0;
        for (int w = 0; w < kSortWaves; ++w) c += waveCount[w * kDigits + d];
        counts[d * tiles + blockIdx.x] = c;
    }
}

// totals[d] += items with digit d (d < totalsKeys), from the scanned
// digit-major offsets: the next digit's first offset minus this one's.
template <int kBits>
__global__ void k_digit_totals(const int* offsets, int tiles, int n, int* totals, int totalsKeys) {
    constexpr int kDigits = 1 << kBits;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= kDigits || d >= totalsKeys) return;
    const int end = d + 1 < kDigits ? offsets[(d + 1) * tiles] : n;
    totals[d] += end - offsets[d * tiles];
}

template <int kBits, typename K, bool kVals>
__global__ void __launch_bounds__(kSortTile) k_digit_scatter(const K* keys, const int* vals, int n, int shift,
                                                             uint32_t mask, int tiles, const int* offsets, K* keysOut,
                                                             int* valsOut) {
    constexpr int kDigits = 1 << kBits;
    __shared__ int waveCount[kSortWaves * kDigits];
    uint32_t digit;
    bool valid;
    int rank;
    tileDigits<kBits>(keys, n, shift, mask, waveCount, digit, valid, rank);
    if (!valid) return;
    const int wave = threadIdx.x / 64;
    for (int w = 0; w < wave; ++w) rank += waveCount[w * kDigits + digit];
    const int i = blockIdx.x * kSortTile + (int)threadIdx.x;
    const int dst = offsets[digit * tiles + blockIdx.x] + rank;
    keysOut[dst] = keys[i];
    if (kVals) valsOut[dst] = vals[i];
}

int sortTiles(int n) { return (n + kSortTile - 1) / kSortTile; }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch of one pass: digit counts, their offsets, and the scan's tile sums.
size_t passBytes(int n, int bits) {
    const int m = (1 << bits) * sortTiles(n);
    return 2 * align256(sizeof(int) * (size_t)m) + scanBytes(m);
}

// One stable pass by the digit of `bits` <= kBits bits at `shift`.
template <int kBits, typename K, bool kVals>
hipError_t digitPass(char* scratch, const K* kin, const int* vin, K* kout, int* vout, int n, int shift, int bits,
                     int* totals, int totalsKeys, hipStream_t s) {
    const uint32_t mask = (1u << bits) - 1u;
    const int tiles = sortTiles(n);
    const int m = (1 << kBits) * tiles;
    int* counts = (int*)scratch;
    int* offsets = (int*)(scratch + align256(sizeof(int) * (size_t)m));
    char* scanTmp = scratch + 2 * align256(sizeof(int) * (size_t)m);
    hipLaunchKernelGGL((k_digit_counts<kBits, K>), dim3(tiles), dim3(kSortTile), 0, s, kin, n, shift, mask, tiles,
                       counts);
    hipError_t e = exclusiveSum(scanTmp, scanBytes(m), counts, offsets, m, s);
    if (e != hipSuccess) return e;
    if (totals && totalsKeys > 0) {
        constexpr int kDigits = 1 << kBits;
        const int threads = kDigits < 256 ? kDigits : 256;
        hipLaunchKernelGGL(k_digit_totals<kBits>, dim3((kDigits + threads - 1) / threads), dim3(threads), 0, s,
                           offsets, tiles, n, totals, totalsKeys);
    }
    hipLaunchKernelGGL((k_digit_scatter<kBits, K, kVals>), dim3(tiles), dim3(kSortTile), 0, s, kin, vin, n, shift,
                       mask, tiles, offsets, kout, vout);
    return hipGetLastError();
}

// Stable LSD radix sort by bits [0, endBit) in 8-bit digits (the last one
// narrower when endBit is not a multiple of 8, so higher bits are ignored),
// ping-ponging between the output and a scratch copy so the last pass writes
// the output.
template <typename K, bool kVals>
size_t radixBytes(int n) {
    return passBytes(n, 8) + align256(sizeof(K) * (size_t)n) + (kVals ? align256(sizeof(int) * (size_t)n) : 0);
}

template <typename K, bool kVals>
hipError_t radixSort(void* scratch, size_t bytes, const K* kin, K* kout, const int* vin, int* vout, int n, int endBit,
                     hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (bytes < radixBytes<K, kVals>(n)) return hipErrorInvalidValue;
    char* base = (char*)scratch;
    K* ktmp = (K*)(base + passBytes(n, 8));
    int* vtmp = kVals ? (int*)(base + passBytes(n, 8) + align256(sizeof(K) * (size_t)n)) : nullptr;
    endBit = std::min(std::max(endBit, 1), (int)(8 * sizeof(K)));
    const int passes = (endBit + 7) / 8;
    const K* ksrc = kin;
    const int* vsrc = vin;
    for (int p = 0; p < passes; ++p) {
        const bool toOut = ((passes - 1 - p) % 2) == 0;
        K* kdst = toOut ? kout : ktmp;
        int* vdst = toOut ? vout : vtmp;
        const hipError_t e = digitPass<8, K, kVals>(base, ksrc, vsrc, kdst, vdst, n, 8 * p,
                                                            std::min(8, endBit - 8 * p), nullptr, 0, s);
        if (e != hipSuccess) return e;
        ksrc = kdst;
        vsrc = vdst;
    }
    return hipSuccess;
}

// ------------------------------------------------------ small key sorts
constexpr int kSmallSortMax = 4096;
constexpr int kSmallSortThreads = 1024;

// Sort of each kSmallSortMax-key chunk of the input in one workgroup (a
// partial chunk padded with the largest key, which sorts last and is not
// written back): each lane sorts its kChunkItems keys in registers, then
// runs of 4, 8, ... keys are merged pairwise in LDS, each lane finding where
// the merge path crosses its first output by a binary search over the
// diagonal and writing kChunkItems outputs.
constexpr int kChunkItems = kSmallSortMax / kSmallSortThreads;
static_assert(kChunkItems == 4, "four keys per lane");

__device__ inline void sortPair(unsigned long long& a, unsigned long long& b) {
    const unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

__global__ void __launch_bounds__(kSmallSortThreads) k_sort_chunks_u64(const unsigned long long* in,
                                                                       unsigned long long* out, int n) {
    __shared__ unsigned long long buf[2][kSmallSortMax];
    const int base = blockIdx.x * kSmallSortMax;
    const int len = n - base < kSmallSortMax ? n - base : kSmallSortMax;
    const int t = threadIdx.x;
    const int k0 = kChunkItems * t;
    unsigned long long x[kChunkItems];
#pragma unroll
    for (int r = 0; r < kChunkItems; ++r) x[r] = k0 + r < len ? in[base + k0 + r] : ~0ull;
    sortPair(x[0], x[1]);
    sortPair(x[2], x[3]);
    sortPair(x[0], x[2]);
    sortPair(x[1], x[3]);
    sortPair(x[1], x[2]);
#pragma unroll
    for (int r = 0; r < kChunkItems; ++r) buf[0][k0 + r] = x[r];
    __syncthreads();
    int cur = 0;
    for (int w = kChunkItems; w < kSmallSortMax; w *= 2) {
        const unsigned long long* src = buf[cur];
        unsigned long long* dst = buf[cur ^ 1];
        const int pairBase = k0 & ~(2 * w - 1);
        const unsigned long long* A = src + pairBase;
        const unsigned long long* B = A + w;
        const int d = k0 - pairBase;
        int lo = d - w > 0 ? d - w : 0, hi = d < w ? d : w;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (A[mid] <= B[d - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        int i = lo, j = d - lo;
#pragma unroll
        for (int r = 0; r < kChunkItems; ++r) {
            const bool takeA = i < w && (j >= w || A[i] <= B[j]);
            dst[k0 + r] = takeA ? A[i++] : B[j++];
        }
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int r = 0; r < kChunkItems; ++r)
        if (k0 + r < len) out[base + k0 + r] = buf[cur][k0 + r];
}

// One pass of a merge sort: sorted runs of `width` keys merged in pairs.
// Each lane writes kMergeItems consecutive outputs of one pair, starting
// where the merge path crosses its first output (a binary search over the
// diagonal; ties take the left run first).
constexpr int kMergeItems = 8;
constexpr int kMergeThreads = 256;
__global__ void __launch_bounds__(kMergeThreads) k_merge_u64(const unsigned long long* in, unsigned long long* out,
                                                            int n, int width) {
    const long long k0 = ((long long)blockIdx.x * kMergeThreads + threadIdx.x) * kMergeItems;
    if (k0 >= n) return;
    const long long pairBase = k0 / (2LL * width) * (2LL * width);
    const int a = (int)(pairBase + width < n ? width : n - pairBase);
    const int b = (int)(pairBase + width >= n ? 0 : (pairBase + 2LL * width < n ? width : n - pairBase - width));
    const unsigned long long* A = in + pairBase;
    const unsigned long long* B = A + a;
    const int d = (int)(k0 - pairBase);
    int lo = d - b > 0 ? d - b : 0, hi = d < a ? d : a;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    int i = lo, j = d - lo;
    for (int t = 0; t < kMergeItems && d + t < a + b; ++t) {
        const bool takeA = i < a && (j >= b || A[i] <= B[j]);
        out[k0 + t] = takeA ? A[i++] : B[j++];
    }
}

// Merge sort above one chunk: chunks sorted in LDS, then merge passes
// ping-ponging between the output and a scratch copy so the last pass
// writes the output. Fewer launches than the radix sort's four per 8-bit
// digit for the pair keys of a step.
constexpr int kMergeSortMax = 1 << 17;

size_t mergeBytes(int n) { return align256(sizeof(unsigned long long) * (size_t)n); }

hipError_t mergeSort(void* scratch, size_t bytes, const unsigned long long* kin, unsigned long long* kout, int n,
                     hipStream_t s) {
    if (bytes < mergeBytes(n)) return hipErrorInvalidValue;
    unsigned long long* tmp = (unsigned long long*)scratch;
    const int chunks = (n + kSmallSortMax - 1) / kSmallSortMax;
    int passes = 0;
    for (int w = kSmallSortMax; w < n; w *= 2) ++passes;
    unsigned long long* dst = passes % 2 == 0 ? kout : tmp;
    hipLaunchKernelGGL(k_sort_chunks_u64, dim3(chunks), dim3(kSmallSortThreads), 0, s, kin, dst, n);
    const int grid = (int)(((long long)n + (long long)kMergeThreads * kMergeItems - 1) / ((long long)kMergeThreads * kMergeItems));
    for (int w = kSmallSortMax; w < n; w *= 2) {
        unsigned long long* next = dst == kout ? tmp : kout;
        hipLaunchKernelGGL(k_merge_u64, dim3(grid), dim3(kMergeThreads), 0, s, dst, next, n, w);
        dst = next;
    }
    return hipGetLastError();
}

constexpr int kCountBits = 7;

}  // namespace

size_t prims_scratch_bytes(int n) {
    size_t m = scanBytes(n);
    m = std::max(m, radixBytes<uint32_t, true>(n));
    m = std::max(m, radixBytes<unsigned long long, false>(n));
    m = std::max(m, prims_count_sort_bytes(n));
    return m + 256;
}

hipError_t prims_exclusive_sum(void* scratch, size_t bytes, const int* in, int* out, int n, hipStream_t s) {
    return exclusiveSum(scratch, bytes, in, out, n, s);
}

hipError_t prims_sort_pairs_u32(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                int* vout, int n, int bits, hipStream_t s) {
    return radixSort<uint32_t, true>(scratch, bytes, kin, kout, vin, vout, n, bits, s);
}

size_t prims_count_sort_bytes(int n) { return passBytes(n, kCountBits); }

hipError_t prims_count_sort_pairs(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                  int* vout, int n, int* totals, int totalsKeys, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (bytes < prims_count_sort_bytes(n)) return hipErrorInvalidValue;
    return digitPass<kCountBits, uint32_t, true>((char*)scratch, kin, vin, kout, vout, n, 0, kCountBits, totals,
                                                  totalsKeys, s);
}

hipError_t prims_sort_keys_u64(void* scratch, size_t bytes, const unsigned long long* kin, unsigned long long* kout,
                               int n, hipStream_t s, int endBit) {
    if (n <= 0) return hipSuccess;
    if (n <= kSmallSortMax) {
        hipLaunchKernelGGL(k_sort_chunks_u64, dim3(1), dim3(kSmallSortThreads), 0, s, kin, kout, n);
        return hipGetLastError();
    }
    if (n <= kMergeSortMax) return mergeSort(scratch, bytes, kin, kout, n, s);
    return radixSort<unsigned long long, false>(scratch, bytes, kin, kout, (const int*)nullptr, (int*)nullptr, n,
                                               endBit, s);
}

}  // namespace b2g
