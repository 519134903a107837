// Device-wide scan and sort used by the step (contact compaction, grid
// build, colour ordering, new-pair ordering), written in prims.hip: a
// two-launch exclusive scan, a stable LSD radix sort in 8-bit digits, one
// pass of it for 7-bit keys, and a one-workgroup bitonic network for small
// 64-bit key sorts. All calls are asynchronous on the given stream and take
// a caller-owned scratch buffer sized by prims_scratch_bytes(); the sorts'
// input and output must not overlap.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace b2g {

// Scratch bytes needed for arrays of up to n items.
size_t prims_scratch_bytes(int n);

// out[i] = sum(in[0..i-1]); in and out may alias.
hipError_t prims_exclusive_sum(void* scratch, size_t bytes, const int* in, int* out, int n, hipStream_t s);

// Stable sort of (key, value) pairs by the low `bits` bits of key (8 bits
// per pass).
hipError_t prims_sort_pairs_u32(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                int* vout, int n, int bits, hipStream_t s);

// Stable sort of (key, value) pairs by the low 7 bits of key (one pass:
// three or four launches); totals[k] (if not null) gains the count of key k
// for k < totalsKeys.
size_t prims_count_sort_bytes(int n);
hipError_t prims_count_sort_pairs(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                  int* vout, int n, int* totals, int totalsKeys, hipStream_t s);

// Ascending sort of 64-bit keys whose bits at and above endBit are equal
// (all keys, for the default 64; a pass per 8 bits). Up to 4096 keys: one
// workgroup in LDS.
hipError_t prims_sort_keys_u64(void* scratch, size_t bytes, const unsigned long long* kin, unsigned long long* kout,
                               int n, hipStream_t s, int endBit = 64);

}  // namespace b2g
