// hipCUB-backed device scan and sort (see prims.h).
#include "prims.h"
#include <hipcub/hipcub.hpp>

namespace b2g {

size_t prims_scratch_bytes(int n) {
    size_t a = 0, b = 0, c = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, (const int*)nullptr, (int*)nullptr, n);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, n, 0, 32);
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, c, (const unsigned long long*)nullptr, (unsigned long long*)nullptr, n,
                                      0, 64);
    size_t m = a > b ? a : b;
    m = m > c ? m : c;
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

hipError_t prims_sort_keys_u64(void* scratch, size_t bytes, const unsigned long long* kin, unsigned long long* kout,
                               int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    return hipcub::DeviceRadixSort::SortKeys(scratch, bytes, kin, kout, n, 0, 64, s);
}

}  // namespace b2g
