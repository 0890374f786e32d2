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
// The same from src into dst (dst may equal src; src is not written otherwise).
hipError_t scan_exclusive_to(const uint32_t* src, uint32_t* dst, size_t n, uint32_t* scratch, hipStream_t stream);

// Stable LSD radix sort of (keys, vals) by key bits [lo_bit, hi_bit) (bits outside the range do not
// take part: equal in-range bits keep the input order), in passes of up to 9 bits.
// keys_alt / vals_alt are ping-pong buffers of n entries. On return *result_keys / *result_vals
// point to whichever pair holds the sorted data. scratch: sort_scratch_words(n). vals_iota: the
// input values are taken to be 0..n-1 (vals is not read, only used as a buffer). nearly_sorted:
// the keys come nearly in order (a previous sort's order), so each pass writes its elements
// straight to their places (the results are the same either way).
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, int lo_bit,
                            int hi_bit, uint32_t* scratch, hipStream_t stream, uint32_t** result_keys,
                            uint32_t** result_vals, bool vals_iota = false, bool nearly_sorted = false);

// Histogram of keys into cnt[0..num_buckets) (cnt must be zeroed) with 32-bit atomics.
hipError_t bucket_count(const uint32_t* keys, size_t n, uint32_t* cnt, hipStream_t stream);

// vals[i] = i for i < n.
hipError_t iota_u32(uint32_t* vals, size_t n, hipStream_t stream);

// Uniform grid of SpatialGrid (src/prims/grid.js:14-21): cell of each point,
// gridIndex(clamp(floor((p - origin) * inv), 0, dims - 1)) in f32, points as float4 (xyz used).
struct GridSpec {
  float origin[3];
  float inv;
  uint32_t dims[3];
};
hipError_t grid_cells(const float4* points, size_t n, GridSpec g, uint32_t* cells, hipStream_t stream);
// Number of sorted-order entries in the 27 cells around each query (the 9 contiguous row ranges of
// SpatialGrid.neighbors, grid.js:101-117): count[q].
hipError_t grid_neighbor_count(const float4* queries, size_t m, GridSpec g, const uint32_t* start, uint32_t* count,
                               hipStream_t stream);
// The sorted-order indices j of those entries, written at offsets[q] in the visiting order (z, then
// y, then j over the row's x range).
hipError_t grid_neighbor_fill(const float4* queries, size_t m, GridSpec g, const uint32_t* start,
                              const uint32_t* offsets, uint32_t* out, hipStream_t stream);

}  // namespace mpm
