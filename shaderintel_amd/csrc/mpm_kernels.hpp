// MLS-MPM substep kernels for gfx950 (P2G, grid update, G2P, binning, layout conversion).
//
// Device data layout
//  - Particles: structure of float4 streams. Stream k (k = 0..7) of a buffer holds vec4 k of every
//    particle record (x|mat, v|J, c0|t00, c1|t11, c2|t22, f0|t01, f1|t02, f2|t12), so a wavefront
//    reading one stream touches 64 consecutive 16-byte elements (1 KiB, fully coalesced).
//  - Grid: block-major. The domain is cut into 4x4x4-node blocks; block b = bx + NBX*(by + NBY*bz)
//    holds its 64 nodes contiguously (local = lx + 4*(ly + 4*lz)): 1 KiB of int4 (A, the i32
//    fixed-point grid that only out-of-window particles and the global-atomic P2G write) or of
//    float4 (V, velocity and mass after the grid update).
//  - Binning (rebuilt whenever the particles are sorted): particles sorted by the Morton code of
//    their stencil base cell form one contiguous run per 4^3-cell "particle block". Each occupied
//    particle block is a work item (start offset, Morton block id) and map[Morton block] = item.
//  - Tiles: P2G gives every work item one workgroup, which sums the 27 fixed-point contributions of
//    its particles in an LDS tile with LDS integer atomics and stores the tile's bounding box
//    (origin, extent) and nodes to tile slot `item` with plain stores. The grid update of node block
//    G pulls from the (at most 64) tiles of particle blocks G-2..G+1 per axis, so no global atomic
//    is needed on this path. A particle whose base drifted more than kDrift cells outside its
//    block since the binning (or any particle of an item beyond the tile slots) falls back to
//    global atomics on A and stamps the node blocks it touched.
//  - Active node blocks: S = node blocks within reach of any work item (from the binning);
//    D = node blocks stamped by fallback atomics in the current P2G (stamp[b] = epoch). The grid
//    update visits S and D and returns A to zero where it was written.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

struct Geo {
  int32_t X, Y, Z;        // cells (= nodes) per axis
  int32_t nbx, nby, nbz;  // 4^3 node blocks per axis
  int32_t mbits;          // bits per axis of a Morton particle-block coordinate
};

struct GridArgs {
  float g[3];
  float bound;
  float floor_friction;
  uint32_t ncol;
};

constexpr int kDrift = 3;                       // cells a base may drift outside its block
constexpr int kTileSpan = 4 + 2 + 2 * kDrift;   // nodes per axis a work item can touch (12)
constexpr int kTileNodes = kTileSpan * kTileSpan * kTileSpan;  // 1728 nodes = 27 KiB of int4

// Everything P2G, the grid update and the export need to know about the current binning.
struct BinView {
  const uint32_t* wl_start;   // [items + 1] particle offsets (in sorted order)
  const uint32_t* wl_block;   // [items] Morton block id
  const uint32_t* item_count; // device scalar
  const uint32_t* map;        // [Morton blocks] -> item or 0xffffffff
  const uint32_t* perm;       // sorted position -> particle index (nullptr: identity)
  int4* tiles;                // [tile_slots * kTileNodes]
  int4* tile_hdr;             // [tile_slots] (ox, oy, oz, ex | ey << 8 | ez << 16)
  uint32_t tile_slots;
  const uint32_t* slist;      // S list and its count
  const uint32_t* scount;
  const uint32_t* sstamp;     // per node block: binning epoch that put it in S
  uint32_t bin_epoch;
};

struct FallbackView {         // D: node blocks written through A in this P2G
  uint32_t* stamp;
  uint32_t* list;
  uint32_t* count;
  uint32_t epoch;
};

// Launch helpers (all enqueue on `s`; none synchronizes).
hipError_t launch_records_in(const float4* aos, float4* P, uint32_t cap, uint32_t offset, uint32_t n, hipStream_t s);
hipError_t launch_records_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, float4* aos, hipStream_t s);
hipError_t launch_rescale(float4* P, uint32_t cap, uint32_t n, float k, hipStream_t s);
hipError_t launch_keys(const float4* P, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s);
hipError_t launch_permute(const float4* src, float4* dst, uint32_t cap, uint32_t n, const uint32_t* perm, hipStream_t s);

// Binning from sorted keys: flags/scan scratch of n + 1 words (scan_scratch separate).
hipError_t launch_bin_flags(const uint32_t* keys_sorted, uint32_t n, uint32_t* flags, hipStream_t s);
hipError_t launch_bin_emit(const uint32_t* keys_sorted, uint32_t n, const uint32_t* scanned, uint32_t* wl_start,
                           uint32_t* wl_block, uint32_t* map, uint32_t* item_count, hipStream_t s);
// Reorders each block's sorted indices by (rank within cell, cell): out[start + pos] = in[start + j].
hipError_t launch_bin_transpose(const uint32_t* keys_sorted, const uint32_t* in, const uint32_t* wl_start,
                                const uint32_t* item_count, uint32_t* out, uint32_t max_items, hipStream_t s);
hipError_t launch_bin_dilate(const uint32_t* wl_block, const uint32_t* item_count, const Geo& geo, uint32_t* sstamp,
                             uint32_t* slist, uint32_t* scount, uint32_t bin_epoch, uint32_t max_items, hipStream_t s);

// P2G over the work items into tiles (+ fallback atomics on A).
hipError_t launch_p2g_binned(const float4* P, uint32_t cap, const float* mats, float vol4, const BinView& bv,
                             int4* A, const FallbackView& fb, const Geo& geo, uint32_t max_items, hipStream_t s);
// P2G with one global atomic per contribution (the reference's forceGlobal baseline).
hipError_t launch_p2g_global(const float4* P, uint32_t cap, uint32_t n, const float* mats, float vol4, int4* A,
                             const FallbackView& fb, const uint32_t* sstamp, uint32_t bin_epoch, const Geo& geo,
                             hipStream_t s);
// Grid update over S and D: pulls tiles + A, writes V and zeroes A where it was written.
hipError_t launch_grid_update(const BinView& bv, const FallbackView& fb, int4* A, float4* V, const Geo& geo,
                              const GridArgs& ga, const float* cols, uint32_t max_blocks, hipStream_t s);
// Folds the tiles into A over S (A then holds the whole P2G result, for reads and checkMPM).
hipError_t launch_resolve_tiles(const BinView& bv, const FallbackView& fb, int4* A, const Geo& geo,
                                uint32_t max_blocks, hipStream_t s);
hipError_t launch_clear_blocks(const uint32_t* list, const uint32_t* count, int4* A, uint32_t max_blocks, hipStream_t s);
hipError_t launch_g2p(float4* P, uint32_t cap, uint32_t n, const float4* V, const Geo& geo, const float* mats, float bound,
                      float vmax, hipStream_t s);
// G2P of this substep fused with the next substep's P2G (tiles + A for the next grid update).
hipError_t launch_g2p_p2g(float4* P, uint32_t cap, const float4* V, const Geo& geo, const float* mats, float bound,
                          float vmax, float vol4, const BinView& bv, int4* A, const FallbackView& fb, uint32_t max_items,
                          hipStream_t s);
// Reference-layout export: mode 0 = A (i32 words), mode 1 = V (f32 bits, 0 outside S and D).
hipError_t launch_grid_export(const int4* A, const float4* V, const uint32_t* sstamp, uint32_t bin_epoch,
                              const uint32_t* stamp, uint32_t epoch, int mode, const Geo& geo, int4* out, hipStream_t s);

}  // namespace mpm
