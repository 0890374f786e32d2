// Exclusive scan and stable radix sort for gfx950 (wave64).
//
// Scan: reduce-then-scan over 4096-element tiles (256 threads x 16 items), recursive on the tile
// partials — the scheme of src/prims/scan.js:252-320, written for 64-wide wavefronts (one wave
// scan with cross-lane shuffles, then one LDS step across the 4 waves of a workgroup).
//
// Radix sort: LSD, 8-bit digits. Per pass: a per-tile digit histogram stored digit-major
// (hist[d * tiles + t]), one exclusive scan over it (so tile t's slot for digit d follows every
// earlier tile's), then a scatter in which each element's rank among equal digits is found with
// eight wave ballots. Elements are visited in index order, so every pass is stable and the final
// order is a deterministic function of the keys (no atomics decide it).
#include "prims.hpp"

namespace mpm {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Workgroup exclusive scan of one value per thread (256 threads); returns the exclusive prefix and
// writes the workgroup total to *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; k++) {
    uint32_t s = wsum[k];
    if (k < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ src, size_t n,
                                                              uint32_t* __restrict__ partials) {
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    size_t i = base + (size_t)j * kScanThreads + threadIdx.x;
    if (i < n) s += src[i];
  }
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Each thread scans 16 contiguous items; tile offsets come from the scanned partials.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t* __restrict__ data, size_t n,
                                                            const uint32_t* __restrict__ partials, int use_offsets) {
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  if (base + kScanItems <= n && (((uintptr_t)(data + base)) & 15) == 0) {
    const uint4* p = reinterpret_cast<const uint4*>(data + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) {
      uint4 q = p[j];
      v[4 * j] = q.x; v[4 * j + 1] = q.y; v[4 * j + 2] = q.z; v[4 * j + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; j++) v[j] = (base + j < n) ? data[base + j] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kScanItems; j++) s += v[j];
  uint32_t tot;
  uint32_t acc = block_excl_scan(s, &tot);
  if (use_offsets) acc += partials[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    uint32_t a = v[j];
    v[j] = acc;
    acc += a;
  }
  if (base + kScanItems <= n && (((uintptr_t)(data + base)) & 15) == 0) {
    uint4* p = reinterpret_cast<uint4*>(data + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) p[j] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; j++)
      if (base + j < n) data[base + j] = v[j];
  }
}

hipError_t scan_level(uint32_t* data, size_t n, uint32_t* scratch, hipStream_t stream) {
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(kScanThreads), 0, stream, data, n, (const uint32_t*)nullptr, 0);
    return hipGetLastError();
  }
  uint32_t* partials = scratch;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream, (const uint32_t*)data, n,
                     partials);
  hipError_t e = scan_level(partials, tiles, scratch + ((tiles + 3) & ~(size_t)3), stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream, data, n,
                     (const uint32_t*)partials, 1);
  return hipGetLastError();
}

// ---- radix sort ----
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                             uint32_t tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; j++) {
    size_t i = base + (size_t)j * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                                size_t n, int shift, uint32_t tiles,
                                                                const uint32_t* __restrict__ hist) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kSortThreads / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  run[tid] = hist[(size_t)tid * tiles + blockIdx.x];
#pragma unroll
  for (int k = 0; k < kSortThreads / 64; k++) wcnt[k][tid] = 0;
  __syncthreads();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const size_t base = (size_t)blockIdx.x * kSortTile;
  for (int j = 0; j < kSortItems; j++) {
    const size_t i = base + (size_t)j * kSortThreads + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int k = 0; k < w; k++) pos += wcnt[k][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int k = 0; k < kSortThreads / 64; k++) {
      add += wcnt[k][tid];
      wcnt[k][tid] = 0;
    }
    run[tid] += add;
    __syncthreads();
  }
}

__global__ void k_bucket_count(const uint32_t* __restrict__ keys, size_t n, uint32_t* __restrict__ cnt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[keys[i]], 1u);
}

__global__ void k_iota(uint32_t* __restrict__ v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

}  // namespace

size_t scan_scratch_words(size_t n) {
  size_t total = 0;
  size_t m = n;
  while (m > (size_t)kScanTile) {
    m = (m + kScanTile - 1) / kScanTile;
    total += (m + 3) & ~(size_t)3;
  }
  return total + 4;
}

size_t sort_scratch_words(size_t n) {
  size_t tiles = (n + kSortTile - 1) / kSortTile;
  size_t h = 256 * tiles;
  return ((h + 3) & ~(size_t)3) + scan_scratch_words(h);
}

hipError_t scan_exclusive(uint32_t* data, size_t n, uint32_t* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  return scan_level(data, n, scratch, stream);
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, int bits,
                            uint32_t* scratch, hipStream_t stream, uint32_t** result_keys, uint32_t** result_vals) {
  *result_keys = keys;
  *result_vals = vals;
  if (n == 0) return hipSuccess;
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  uint32_t* hist = scratch;
  uint32_t* scan_tmp = scratch + (((size_t)256 * tiles + 3) & ~(size_t)3);
  uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, stream, (const uint32_t*)ka, n, shift, tiles,
                       hist);
    hipError_t e = scan_exclusive(hist, (size_t)256 * tiles, scan_tmp, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, stream, (const uint32_t*)ka,
                       (const uint32_t*)va, kb, vb, n, shift, tiles, (const uint32_t*)hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t* t;
    t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  *result_keys = ka;
  *result_vals = va;
  return hipSuccess;
}

hipError_t bucket_count(const uint32_t* keys, size_t n, uint32_t* cnt, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bucket_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, keys, n, cnt);
  return hipGetLastError();
}

__device__ __forceinline__ void grid_coord(float4 p, const GridSpec& g, int c[3]) {
  const float q[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int v = (int)floorf((q[a] - g.origin[a]) * g.inv);
    c[a] = min(max(v, 0), (int)g.dims[a] - 1);
  }
}

__global__ void k_grid_cells(const float4* __restrict__ pts, size_t n, GridSpec g, uint32_t* __restrict__ cells) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
  grid_coord(pts[i], g, c);
  cells[i] = (uint32_t)c[0] + g.dims[0] * ((uint32_t)c[1] + g.dims[1] * (uint32_t)c[2]);
}

// Visits the 9 row ranges [start[row + x0], start[row + x1 + 1]) around the query's cell.
template <class F>
__device__ __forceinline__ void grid_rows(float4 q, const GridSpec& g, const uint32_t* __restrict__ start, F&& f) {
  int c[3];
  grid_coord(q, g, c);
  const uint32_t x0 = (uint32_t)max(c[0] - 1, 0), x1 = (uint32_t)min(c[0] + 1, (int)g.dims[0] - 1);
  for (int z = max(c[2] - 1, 0); z <= min(c[2] + 1, (int)g.dims[2] - 1); z++)
    for (int y = max(c[1] - 1, 0); y <= min(c[1] + 1, (int)g.dims[1] - 1); y++) {
      const uint32_t row = g.dims[0] * ((uint32_t)y + g.dims[1] * (uint32_t)z);
      f(start[row + x0], start[row + x1 + 1]);
    }
}

__global__ void k_grid_neighbor_count(const float4* __restrict__ qs, size_t m, GridSpec g,
                                      const uint32_t* __restrict__ start, uint32_t* __restrict__ count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t n = 0;
  grid_rows(qs[i], g, start, [&](uint32_t b, uint32_t e) { n += e - b; });
  count[i] = n;
}

__global__ void k_grid_neighbor_fill(const float4* __restrict__ qs, size_t m, GridSpec g,
                                     const uint32_t* __restrict__ start, const uint32_t* __restrict__ offsets,
                                     uint32_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t o = offsets[i];
  grid_rows(qs[i], g, start, [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; j++) out[o++] = j;
  });
}

hipError_t grid_cells(const float4* points, size_t n, GridSpec g, uint32_t* cells, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_cells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, points, n, g, cells);
  return hipGetLastError();
}
hipError_t grid_neighbor_count(const float4* queries, size_t m, GridSpec g, const uint32_t* start, uint32_t* count,
                               hipStream_t stream) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_neighbor_count, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, queries, m, g,
                     start, count);
  return hipGetLastError();
}
hipError_t grid_neighbor_fill(const float4* queries, size_t m, GridSpec g, const uint32_t* start,
                              const uint32_t* offsets, uint32_t* out, hipStream_t stream) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_neighbor_fill, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, queries, m, g,
                     start, offsets, out);
  return hipGetLastError();
}

hipError_t iota_u32(uint32_t* vals, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, vals, n);
  return hipGetLastError();
}

}  // namespace mpm
