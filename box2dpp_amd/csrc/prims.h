// Device-wide scan and sort used by the step (contact compaction, grid
// build, colour ordering, new-pair ordering). Backed by hipCUB's radix sort
// and scan; all calls are asynchronous on the given stream and take a
// caller-owned scratch buffer sized by prims_scratch_bytes().
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace b2g {

// Scratch bytes needed for arrays of up to n items.
size_t prims_scratch_bytes(int n);

// out[i] = sum(in[0..i-1]); in and out may alias.
hipError_t prims_exclusive_sum(void* scratch, size_t bytes, const int* in, int* out, int n, hipStream_t s);

// Stable sort of (key, value) pairs by the low `bits` bits of key.
hipError_t prims_sort_pairs_u32(void* scratch, size_t bytes, const uint32_t* kin, uint32_t* kout, const int* vin,
                                int* vout, int n, int bits, hipStream_t s);

// Sort of 64-bit keys.
hipError_t prims_sort_keys_u64(void* scratch, size_t bytes, const unsigned long long* kin, unsigned long long* kout,
                               int n, hipStream_t s);

}  // namespace b2g
