// Device primitives used by the MPM sort: exclusive scan of u32 and a stable LSD radix sort of
// (key, value) u32 pairs. Internal C++ interface of libmpm_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mpm {

// Scratch sizes (in u32 words) needed by the routines below for n elements.
size_t scan_scratch_words(size_t n);
size_t sort_scratch_words(size_t n);

// In-place exclusive prefix sum of n u32 (wrapping) on `stream`. scratch: scan_scratch_words(n).
hipError_t scan_exclusive(uint32_t* data, size_t n, uint32_t* scratch, hipStream_t stream);

// Stable LSD radix sort of (keys, vals) by the low `bits` bits of the keys, 8 bits per pass.
// keys_alt / vals_alt are ping-pong buffers of n entries. On return *result_keys / *result_vals
// point to whichever pair holds the sorted data. scratch: sort_scratch_words(n).
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, int bits,
                            uint32_t* scratch, hipStream_t stream, uint32_t** result_keys, uint32_t** result_vals);

// Histogram of keys into cnt[0..num_buckets) (cnt must be zeroed) with 32-bit atomics.
hipError_t bucket_count(const uint32_t* keys, size_t n, uint32_t* cnt, hipStream_t stream);

// vals[i] = i for i < n.
hipError_t iota_u32(uint32_t* vals, size_t n, hipStream_t stream);

}  // namespace mpm
