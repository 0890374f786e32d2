// Exclusive scan and stable radix sort for gfx950 (wave64).
//
// Scan: reduce-then-scan over 4096-element tiles (256 threads x 16 items), recursive on the tile
// partials — the scheme of src/prims/scan.js:252-320, written for 64-wide wavefronts (one wave
// scan with cross-lane shuffles, then one LDS step across the 4 waves of a workgroup).
//
// Radix sort: LSD over a bit range, in passes of up to 9-bit digits. Per pass: a per-tile digit
// histogram stored digit-major (hist[d * tiles + t]), one exclusive scan over it (so tile t's slot
// for digit d follows every earlier tile's), then a scatter that ranks the tile in LDS (an
// element's rank among equal digits comes from wave ballots, in index order), stages it in sorted
// order and writes runs of equal digits. Every pass is stable and the final order is a
// deterministic function of the keys (no atomic decides a position).
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

// Each thread scans 16 contiguous items of src into dst (the same array for an in-place scan; a
// thread reads its items before it writes them). Tile offsets: none (use_offsets 0, a single tile),
// the scanned partials (1), or the sum of the unscanned partials of the earlier tiles, added up here
// (2; at most kScanTile tiles, so each thread adds at most kScanItems of them).
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* src, uint32_t* dst, size_t n,
                                                            const uint32_t* __restrict__ partials, int use_offsets) {
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  if (base + kScanItems <= n && (((uintptr_t)(src + base)) & 15) == 0) {
    const uint4* p = reinterpret_cast<const uint4*>(src + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) {
      uint4 q = p[j];
      v[4 * j] = q.x; v[4 * j + 1] = q.y; v[4 * j + 2] = q.z; v[4 * j + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; j++) v[j] = (base + j < n) ? src[base + j] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kScanItems; j++) s += v[j];
  uint32_t tot;
  uint32_t acc = block_excl_scan(s, &tot);
  if (use_offsets == 1) acc += partials[blockIdx.x];
  if (use_offsets == 2) {
    uint32_t ps = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) ps += partials[t];
    uint32_t before;
    block_excl_scan(ps, &before);  // the workgroup total of ps
    acc += before;
  }
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    uint32_t a = v[j];
    v[j] = acc;
    acc += a;
  }
  if (base + kScanItems <= n && (((uintptr_t)(dst + base)) & 15) == 0) {
    uint4* p = reinterpret_cast<uint4*>(dst + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) p[j] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; j++)
      if (base + j < n) dst[base + j] = v[j];
  }
}

// Exclusive scan of n words of src into dst (may be the same array); the tile partials go to scratch.
hipError_t scan_level(const uint32_t* src, uint32_t* dst, size_t n, uint32_t* scratch, hipStream_t stream) {
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(kScanThreads), 0, stream, src, dst, n, (const uint32_t*)nullptr, 0);
    return hipGetLastError();
  }
  uint32_t* partials = scratch;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream, src, n, partials);
  if (tiles <= (size_t)kScanTile) {  // each tile adds up the partials before it: no third launch
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream, src, dst, n,
                       (const uint32_t*)partials, 2);
    return hipGetLastError();
  }
  hipError_t e = scan_level(partials, partials, tiles, scratch + ((tiles + 3) & ~(size_t)3), stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream, src, dst, n,
                     (const uint32_t*)partials, 1);
  return hipGetLastError();
}

// ---- radix sort ----
// Digits of up to kRadixMaxBits bits. A tile is kSortTile consecutive elements; wave w of a tile's
// workgroup owns the contiguous chunk [w * kSortChunk, (w + 1) * kSortChunk) of it, read 64 elements
// at a time in index order.
constexpr int kRadixMaxBits = 9;
constexpr int kRadixMaxBuckets = 1 << kRadixMaxBits;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortChunk = kSortTile / kSortWaves;  // 1024

// Lanes of the wave whose digit equals this lane's (among the lanes in `valid`).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, int bits, uint64_t valid) {
  uint64_t peers = valid;
  for (int b = 0; b < bits; b++) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Per-tile digit counts, digit-major (hist[d * tiles + t]). Each thread counts 16 consecutive keys
// (four 16-byte loads), adding a run of equal digits to the LDS histogram at once (the keys arrive
// nearly sorted). Workgroups are dealt to the 8 XCDs in turn, so workgroup b takes tile
// (b % 8) * per + b / 8 (per = ceil(tiles / 8)): each XCD counts a contiguous run of tiles, and a
// cache line of a digit's row (32 tiles) is written from one XCD's L2 instead of from all eight.
constexpr uint32_t kXcds = 8;
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                             int bits, uint32_t tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kRadixMaxBuckets];
  const uint32_t per = (tiles + kXcds - 1) / kXcds;
  const uint32_t tile = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (tile >= tiles) return;  // the whole workgroup: no barrier is reached
  const int nb = 1 << bits;
  const uint32_t mask = (uint32_t)nb - 1u;
  for (int d = threadIdx.x; d < nb; d += kSortThreads) h[d] = 0;
  __syncthreads();
  const size_t first = (size_t)tile * kSortTile + (size_t)threadIdx.x * kSortItems;
  uint32_t k[kSortItems];
  int valid = 0;
  if (first + kSortItems <= n) {  // keys is 16-byte aligned and first is a multiple of 16
    const uint4* q = reinterpret_cast<const uint4*>(keys + first);
#pragma unroll
    for (int j = 0; j < kSortItems / 4; j++) {
      const uint4 v = q[j];
      k[4 * j] = v.x; k[4 * j + 1] = v.y; k[4 * j + 2] = v.z; k[4 * j + 3] = v.w;
    }
    valid = kSortItems;
  } else {
#pragma unroll
    for (int j = 0; j < kSortItems; j++) k[j] = first + j < n ? keys[first + j] : 0u;
    valid = first < n ? (int)(n - first) : 0;
  }
  uint32_t run_d = (k[0] >> shift) & mask, run_n = 0;
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    if (j < valid) {
      const uint32_t d = (k[j] >> shift) & mask;
      if (d != run_d) {
        atomicAdd(&h[run_d], run_n);
        run_d = d;
        run_n = 0;
      }
      run_n++;
    }
  }
  if (run_n) atomicAdd(&h[run_d], run_n);
  __syncthreads();
  for (int d = threadIdx.x; d < nb; d += kSortThreads) hist[(size_t)d * tiles + tile] = h[d];
}

// Stable scatter of one tile. hist holds the exclusive scan of the digit-major counts, so
// hist[d * tiles + t] is where tile t's first element of digit d goes. Wave w ranks its chunk in
// index order (an element's rank among equal digits of the chunk: a running LDS count per digit plus
// its rank among the equal-digit lanes of its 64-element step, found with ballots); after the
// per-digit offsets of the tile and of the earlier waves are known, the tile is staged in LDS in
// sorted order and written out in runs of equal digits (coalesced). DIRECT: each element is written
// straight to its place, without the staging (for input already nearly in digit order, where the
// lanes of a step fall in few runs; less LDS, more workgroups per CU).
template <bool DIRECT>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                                size_t n, int shift, int bits, uint32_t tiles,
                                                                const uint32_t* __restrict__ hist, int vals_iota) {
  __shared__ uint32_t skey[DIRECT ? 1 : kSortTile];
  __shared__ uint32_t sval[DIRECT ? 1 : kSortTile];
  __shared__ uint32_t run[kSortWaves][kRadixMaxBuckets];  // chunk counts, then each wave's first local slot
  __shared__ uint32_t gofs[kRadixMaxBuckets];             // output index of local slot 0 of digit d
  __shared__ uint32_t wsum[kSortWaves];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // the histogram's tile order: each XCD takes a contiguous run of tiles (their offsets share lines)
  const uint32_t per = (tiles + kXcds - 1) / kXcds;
  const uint32_t tile = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (tile >= tiles) return;  // the whole workgroup: no barrier is reached
  const int nb = 1 << bits;
  const uint32_t mask = (uint32_t)nb - 1u;
  const size_t base = (size_t)tile * kSortTile;
  const uint32_t count = (uint32_t)((n - base) < (size_t)kSortTile ? (n - base) : (size_t)kSortTile);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int d = tid; d < nb; d += kSortThreads)
#pragma unroll
    for (int k = 0; k < kSortWaves; k++) run[k][d] = 0;
  __syncthreads();
  // 1. load the chunk (64 consecutive elements per step) and rank it within the chunk
  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const uint32_t e = (uint32_t)(w * kSortChunk + j * 64 + lane);
    const bool valid = e < count;
    key[j] = valid ? kin[base + e] : 0u;
    val[j] = valid ? (vals_iota ? (uint32_t)(base + e) : vin[base + e]) : 0u;
  }
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const uint32_t e = (uint32_t)(w * kSortChunk + j * 64 + lane);
    const uint64_t valid = __ballot(e < count);
    rank[j] = 0;
    if (!valid) continue;
    const uint32_t d = (key[j] >> shift) & mask;
    // the valid lanes are a prefix, so lane 0 is one; when every valid lane has its digit (runs of
    // equal keys), they are all peers, else the ballots sort them out
    const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)d, 0);
    const uint64_t peers = __ballot(e < count && d != d0) == 0 ? valid : digit_peers(d, bits, valid);
    if (e < count) rank[j] = run[w][d] + (uint32_t)__popcll(peers & below);
    // every lane has read run[w][d]; the first lane of each digit advances it
    if (e < count && (peers & below) == 0) run[w][d] += (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // 2. exclusive scan over digits of the tile counts; wave k's elements of digit d start after the
  // earlier waves' ones
  {
    constexpr int kPer = kRadixMaxBuckets / kSortThreads;  // digits per thread (2)
    uint32_t c[kPer], s = 0;
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int d = tid * kPer + q;
      c[q] = 0;
      if (d < nb)
#pragma unroll
        for (int k = 0; k < kSortWaves; k++) c[q] += run[k][d];
      s += c[q];
    }
    uint32_t inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = inc - s;
    for (int k = 0; k < w; k++) off += wsum[k];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int d = tid * kPer + q;
      if (d < nb) {
        gofs[d] = hist[(size_t)d * tiles + tile] - off;
        uint32_t o = off;
#pragma unroll
        for (int k = 0; k < kSortWaves; k++) {
          const uint32_t ck = run[k][d];
          run[k][d] = o;
          o += ck;
        }
      }
      off += c[q];
    }
  }
  __syncthreads();
  if (DIRECT) {  // local slot s of digit d goes to gofs[d] + s
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
      const uint32_t e = (uint32_t)(w * kSortChunk + j * 64 + lane);
      if (e < count) {
        const uint32_t d = (key[j] >> shift) & mask;
        const uint32_t pos = gofs[d] + run[w][d] + rank[j];
        kout[pos] = key[j];
        vout[pos] = val[j];
      }
    }
    return;
  }
  // 3. stage the tile in sorted order
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const uint32_t e = (uint32_t)(w * kSortChunk + j * 64 + lane);
    if (e < count) {
      const uint32_t slot = run[w][(key[j] >> shift) & mask] + rank[j];
      skey[slot] = key[j];
      sval[slot] = val[j];
    }
  }
  __syncthreads();
  // 4. write out: local slot s of digit d goes to gofs[d] + s
  for (uint32_t s = (uint32_t)tid; s < count; s += kSortThreads) {
    const uint32_t k = skey[s];
    const uint32_t pos = gofs[(k >> shift) & mask] + s;
    kout[pos] = k;
    vout[pos] = sval[s];
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
  size_t h = (size_t)kRadixMaxBuckets * tiles;
  return ((h + 3) & ~(size_t)3) + scan_scratch_words(h);
}

hipError_t scan_exclusive(uint32_t* data, size_t n, uint32_t* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  return scan_level(data, data, n, scratch, stream);
}

hipError_t scan_exclusive_to(const uint32_t* src, uint32_t* dst, size_t n, uint32_t* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  return scan_level(src, dst, n, scratch, stream);
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, size_t n, int lo_bit,
                            int hi_bit, uint32_t* scratch, hipStream_t stream, uint32_t** result_keys,
                            uint32_t** result_vals, bool vals_iota, bool nearly_sorted) {
  *result_keys = keys;
  *result_vals = vals;
  if (n == 0) return hipSuccess;
  if (hi_bit <= lo_bit) return vals_iota ? iota_u32(vals, n, stream) : hipSuccess;
  if (lo_bit < 0 || hi_bit > 32) return hipErrorInvalidValue;
  const int span = hi_bit - lo_bit;
  const int passes = (span + kRadixMaxBits - 1) / kRadixMaxBits;
  const int bits = (span + passes - 1) / passes;
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  uint32_t* hist = scratch;
  uint32_t* scan_tmp = scratch + (((size_t)kRadixMaxBuckets * tiles + 3) & ~(size_t)3);
  uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
  for (int shift = lo_bit; shift < hi_bit; shift += bits) {
    const int b = hi_bit - shift < bits ? hi_bit - shift : bits;
    hipLaunchKernelGGL(k_radix_hist, dim3((tiles + kXcds - 1) / kXcds * kXcds), dim3(kSortThreads), 0, stream,
                       (const uint32_t*)ka, n, shift, b, tiles, hist);
    hipError_t e = scan_exclusive(hist, ((size_t)1 << b) * tiles, scan_tmp, stream);
    if (e != hipSuccess) return e;
    const dim3 grid((tiles + kXcds - 1) / kXcds * kXcds);
    const int iota = vals_iota && shift == lo_bit ? 1 : 0;
    // nearly sorted input stays nearly sorted by the next digit after a stable pass (the elements of
    // one full key stay together), so both passes can write directly
    if (nearly_sorted)
      hipLaunchKernelGGL(k_radix_scatter<true>, grid, dim3(kSortThreads), 0, stream, (const uint32_t*)ka,
                         (const uint32_t*)va, kb, vb, n, shift, b, tiles, (const uint32_t*)hist, iota);
    else
      hipLaunchKernelGGL(k_radix_scatter<false>, grid, dim3(kSortThreads), 0, stream, (const uint32_t*)ka,
                         (const uint32_t*)va, kb, vb, n, shift, b, tiles, (const uint32_t*)hist, iota);
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
