// MLS-MPM substep kernels for gfx950 (P2G, grid update, G2P, binning, layout conversion).
//
// One set of kernels advances any number of independent scenes at once ("a pool"): a plain
// simulation is a pool with one scene. Every phase is one launch over all scenes.
//
// Device data layout
//  - Particles: one pool of 8 structure-of-float4 streams (S0..S7, 128 bytes per particle), so a
//    wavefront reading one stream touches 64 consecutive 16-byte elements. The 32 floats of the
//    reference record are regrouped so that P2G reads 96 bytes (S0-S5) and G2P reads 64 (S0, S5,
//    S6, S7; fluids only S0 and S5):
//      S0 = x.xyz | row            S1 = v.xyz | t01
//      S2 = C col 0 | t00          S3 = C col 1 | t11          S4 = C col 2 | t22
//      S5 = t02, t12, J, F22       S6 = F00, F10, F20, F01     S7 = F11, F21, F02, F12
//    (t = the stress tau, F col-major: Frc = row r of column c; kRecSlot maps the reference record.)
//    Positions are in the scene's own lattice coordinates; x.w holds the particle's global material
//    row (scene * 8 + material), so row >> 3 is the particle's scene.
//  - Grid: a pool of 4x4x4-node blocks. Scene s owns blocks [block_base, block_base + NBX*NBY*NBZ),
//    block-major inside (block = bx + NBX*(by + NBY*bz), node = lx + 4*(ly + 4*lz)): 1 KiB of int4
//    (A, the i32 fixed-point grid written only by the global-atomic paths) or of float4 (V,
//    velocity and mass after the grid update).
//  - Binning (rebuilt at every sort): particles sorted by (scene, Morton code of their stencil base
//    cell) form one contiguous run per 4^3-cell particle block. Each occupied block is a work item
//    (start offset, key >> 6). Inside a block the order is (rank within cell, cell numbered x
//    fastest), so consecutive particles sit in different cells and a rank layer's z layers fill
//    the quarters of a wavefront.
//  - Tiles: P2G gives every work item one workgroup that sums the 27 fixed-point contributions of
//    its particles in an LDS tile (LDS integer atomics) sized to the bounding box of their stencils,
//    takes that many nodes from one region of a bump-allocated arena and stores the box and the
//    nodes there with plain stores (tile_hdr[item], tile_off[item], and by the block's slot in
//    mtile, stamped with the P2G epoch), node block by node block (tile_index, item_sched.hpp: the
//    part of a box that one node block holds is one contiguous run). The grid update of a node
//    block pulls from the tiles of particle blocks G-2..G+1 per axis (their mtile slots), so the
//    tiled path needs no global atomic. A particle whose base drifted more than kDrift cells
//    outside its block since the binning, or every particle of an item whose tile does not fit its
//    arena region, falls back to global atomics on A and stamps the node blocks it wrote.
//  - Active node blocks: S = node blocks within reach of any work item (rebuilt at each binning);
//    D = node blocks written through A in the current P2G. The grid update visits S and D and
//    returns A to zero where it was written.
//  - Counters (device words, so captured graphs can be replayed): P2G epoch, binning epoch, item
//    count, |D|, |S|, work counter, and the epochs the velocities in V belong to.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int kDrift = 3;                                      // cells a base may drift outside its block, at most
constexpr int kTileSpan = 4 + 2 + 2 * kDrift;                  // nodes per axis a work item can touch (12)
constexpr int kTileNodes = kTileSpan * kTileSpan * kTileSpan;  // 1728 nodes = 27 KiB of int4
// The window of a work item with drift D (2 or 3; MPM_TUNE_DRIFT): stencil bases within D cells of
// its 4^3-cell block, an LDS tile of kSpan^3 nodes (10^3 = 16 KiB or 12^3 = 27 KiB of int4). A particle
// that moves at most 1 cell per substep stays in the drift-3 window for the 3 substeps between
// sorts; one that leaves its window takes P2G's global-atomic path (the same sums).
template <int D>
struct Window {
  static_assert(D >= 2 && D <= kDrift, "drift 2 or 3");
  static constexpr int kSpan = 4 + 2 + 2 * D;
  static constexpr int kNodes = kSpan * kSpan * kSpan;
};
// Float f of the reference's 32-float record (x, v, c0, c1, c2, f0, f1, f2 as vec4f, mpm.js:54-63)
// is component kRecSlot[f] & 3 of stream kRecSlot[f] >> 2.
constexpr uint8_t kRecSlot[32] = {
    0,  1,  2,  3,   // x.xyz, row -> S0
    4,  5,  6,  22,  // v.xyz -> S1.xyz; J -> S5.z
    8,  9,  10, 11,  // c0.xyz, t00 -> S2
    12, 13, 14, 15,  // c1.xyz, t11 -> S3
    16, 17, 18, 19,  // c2.xyz, t22 -> S4
    24, 25, 26, 7,   // f0.xyz -> S6.xyz; t01 -> S1.w
    27, 28, 29, 20,  // f1.xyz -> S6.w, S7.x, S7.y; t02 -> S5.x
    30, 31, 23, 21,  // f2.xyz -> S7.z, S7.w, S5.w; t12 -> S5.y
};
constexpr int kMatRows = 8;                                    // material rows per scene (row >> 3 = scene)
constexpr int kColRows = 16;                                   // collider rows per scene
// The tile arena is split into kArenaRegions regions, each with its own bump counter on its own
// 64-byte line: workgroup w allocates from region w % kArenaRegions, so allocations do not contend
// on one address (a single shared counter cost 0.8 ms per P2G at 9k items on MI355X).
constexpr int kArenaRegions = 256;
constexpr int kArenaStride = 16;  // u32 words between region counters
// Work items of P2G are split into kWorkSplits contiguous ranges, each with its own counter on its
// own 128-byte line after the arena's region counters. Workgroup w starts on range w % kWorkSplits
// (the XCD it was dispatched to) and moves to the next range when its own is used up, so no single
// address takes every item's atomic (one shared counter cost 7% of P2G on MI355X).
constexpr int kWorkSplits = 8;
constexpr int kWorkStride = 32;
// Two sets of work counters: P2G's, then G2P's (which walks the same items).
constexpr int kArenaCtrWords = kArenaRegions * kArenaStride + 2 * kWorkSplits * kWorkStride;
static_assert(2 * kWorkSplits <= kArenaRegions, "k_begin_p2g clears the work counters with its region threads");

enum Counter : int {
  kCtrEpoch = 0,     // P2G epoch (stamps of D)
  kCtrBinEpoch = 1,  // binning epoch (stamps of S)
  kCtrItems = 2,     // work items of the current binning
  kCtrD = 3,         // |D|
  kCtrS = 4,         // |S|
  kCtrWork = 5,      // unused (P2G's work counters are in arena_ctr)
  kCtrVEpoch = 6,    // P2G epoch whose grid update wrote V
  kCtrVBinEpoch = 7, // unused (a grid update stamps the blocks it writes, vstamp)
  kCounters = 8
};

// Per-scene geometry and step parameters (device table, one entry per scene).
struct SceneGeo {
  int32_t X, Y, Z;        // cells (= nodes) per axis
  int32_t nbx, nby, nbz;  // 4^3 node blocks per axis
  int32_t mbits;          // bits per axis of a Morton particle-block coordinate (2^(3*mbits) slots)
  uint32_t block_base;    // first node block of the scene in the pool
  uint32_t map_base;      // first particle-block slot of the scene (scene << 3 * mbits)
  float bound;            // boundary cells (mpm.js:103)
  float g[3];             // lattice gravity (mpm.js:588)
  float floor_friction;   // mpm.js:594
  uint32_t ncol;          // colliders (rows scene * kColRows + k)
  float vol4;             // 4 * particle volume = 4 / ppc_axis^3 (mpm.js:610)
  uint32_t active;        // 0: the scene's particles are dropped at the next sort
};

// 64-bit words per row of a scene's particle-block occupancy bitmap (pb blocks per axis).
__host__ __device__ inline uint32_t occ_row_words(uint32_t pb) { return (pb + 63u) / 64u; }

// Pointers and sizes fixed for the pool's lifetime, passed by value to every kernel.
struct Pool {
  uint32_t cap;            // particle capacity of each P buffer
  uint32_t key_bits;       // bits of the per-scene cell key (3 * per-axis bits)
  uint32_t nblocks;        // node blocks in the pool
  uint32_t max_items;      // work items a binning can produce
  uint32_t tile_nodes;     // nodes in the tile arena (kArenaRegions regions of tile_nodes / kArenaRegions)
  int4* A;
  float4* V;
  const uint2* block_info; // node block -> (scene, bx | by << 10 | bz << 20) in the scene's blocks
  uint32_t* stamp;         // D stamps (P2G epoch)
  uint32_t* dlist;
  uint32_t* sstamp;        // S stamps (binning epoch)
  uint32_t* vstamp;        // V stamps: the P2G epoch of the grid update that last wrote the block's V
  uint32_t* slist;
  uint2* sinfo;            // [nblocks] block_info of slist's entries, written beside them
  uint4* mtile;            // [block slots] tile of the particle block at each slot (scene << (key_bits - 6)
                           // | Morton block): (arena node, ox | oy << 10 | oz << 20, ex | ey << 8 | ez << 16,
                           // P2G epoch that wrote it); an entry of another epoch, or arena node 0xffffffff,
                           // has no tile
  uint32_t* wl_start;      // [max_items + 1]
  uint32_t* wl_block;      // [max_items] key >> 6 (scene bits on top)
  int4* tiles;             // [tile_nodes] arena
  int4* tile_hdr;          // [max_items] (ox, oy, oz, ex | ey << 8 | ez << 16)
  uint32_t* tile_off;      // [max_items] first arena node of the item's tile, 0xffffffff: none
  const float* mats;       // [scenes * kMatRows * 12]
  const float* cols;       // [scenes * kColRows * 12]
  const SceneGeo* scenes;
  uint32_t* ctr;           // [kCounters]
  unsigned long long* occ;  // [max_scenes * pb * pb * occ_row_words(pb)] occupied particle blocks of the
                            // binning, one bit per block: row (by, bz) of a scene, bit bx
  uint32_t* arena_ctr;     // [kArenaCtrWords]: nodes handed out per region in the current P2G, then the
                           // work counters of P2G and of G2P
  unsigned long long* law; // [kLawSlots * kLawWords] grid-update momentum bookkeeping, nullptr: off
};

// Momentum bookkeeping of the grid update (enabled per pool; no counterpart in the reference). For
// every node with mass, with p its P2G momentum words, m and v its mass and velocity after the
// update and vf = p/m + g its velocity before the walls, floor and colliders act, in wrapping 64-bit
// words: [0] mass words, [1..3] momentum words (2^22 fixed point), [4..6] m v, [7..9] m g,
// [10..12] m (v - vf), each f32 product as rint(x 2^32); [13] nodes; [14] grid updates (slot 0).
// Workgroups add into slot blockIdx % kLawSlots; the host sums the slots.
constexpr int kLawSlots = 64;
constexpr int kLawWords = 16;

// Launch helpers; all enqueue on `s` and none synchronizes. P is the current particle buffer.
hipError_t launch_records_in(const float4* aos, float4* P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base,
                             hipStream_t s);
hipError_t launch_records_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base, float4* aos,
                              hipStream_t s);
hipError_t launch_rescale(float4* P, uint32_t cap, uint32_t offset, uint32_t n, float k, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s);
// block_info of the n = nbx*nby*nbz node blocks of a scene starting at block `base`.
hipError_t launch_block_info(uint2* info, uint32_t base, uint32_t n, uint32_t scene, uint32_t nbx, uint32_t nby,
                             hipStream_t s);
// ctr[kCtrEpoch]++, |D| = 0, work = 0, arena empty (start of a P2G)
hipError_t launch_begin_p2g(const Pool& pool, hipStream_t s);
// ctr[kCtrBinEpoch]++, |S| = 0, items = 0 (start of a binning)
hipError_t launch_begin_bin(const Pool& pool, hipStream_t s);
// launch_begin_bin + the occupancy bitmap cleared (one launch)
hipError_t launch_bin_reset(const Pool& pool, uint32_t occ_words32, hipStream_t s);

// Sort keys (scene << key_bits | Morton(base cell)) of the particles in their current order;
// particles of inactive scenes get key 0xffffffff (sorted behind every live particle).
hipError_t launch_keys(const Pool& pool, const float4* P, uint32_t n, uint32_t* keys, hipStream_t s);
// Keys 0 for particles of active scenes, 1 for removed ones (a stable sort on them is a stable
// partition); vals = index.
hipError_t launch_live_keys(const Pool& pool, const float4* P, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s);
hipError_t launch_permute(const float4* src, float4* dst, uint32_t cap, uint32_t n, const uint32_t* perm, hipStream_t s);
// Binning from the block-sorted keys: per-tile head counts (bin_tiles(n) + 1 words, the last one 0),
// then, after an exclusive scan of those words (the last becomes the item count), the work items;
// the within-block transposition of the sorted indices into `out`; and S.
uint32_t bin_tiles(uint32_t n);
hipError_t launch_bin_count(const uint32_t* keys_sorted, uint32_t n, uint32_t* tile_heads, hipStream_t s);
// scanned: tile_first holds the scanned head counts; otherwise the counts (bin_emit_scans(n) must hold)
bool bin_emit_scans(uint32_t n);
hipError_t launch_bin_emit(const Pool& pool, const uint32_t* keys_sorted, uint32_t n, const uint32_t* tile_first,
                           bool scanned, hipStream_t s);
hipError_t launch_bin_transpose(const Pool& pool, const uint32_t* keys_sorted, const uint32_t* in, uint32_t* out,
                                hipStream_t s);
// S from the occupancy bitmap: each node block of a scene's grid with an occupied particle block
// within [-(drift - 1), +1] per axis (the blocks whose drift-window tiles can reach it), appended to
// slist and stamped with the binning epoch.
hipError_t launch_bin_dilate(const Pool& pool, uint32_t max_scenes, int drift, hipStream_t s);

// P2G over the work items into tiles (+ fallback atomics on A). perm: sorted position -> particle
// (nullptr: identity).
// drift: the work items' window (Window<drift>, 2 or 3). keys_out (optional): the sort key of each
// record at its slot (as launch_keys computes them).
hipError_t launch_p2g_binned(const Pool& pool, const float4* P, const uint32_t* perm, int waves, int drift,
                             hipStream_t s, uint32_t* keys_out = nullptr);
// P2G with one global atomic per contribution (the reference's forceGlobal path), any order.
hipError_t launch_p2g_global(const Pool& pool, const float4* P, uint32_t n, hipStream_t s);
// Writes n (<= 64) 32-bit words to dst from a kernel argument (ordered with the stream's work).
hipError_t launch_write_words(uint32_t* dst, const uint32_t* src, uint32_t n, hipStream_t s);
// Grid update over S and D (tiles + A -> V, A back to zero). use_s = false: only D holds data. drift:
// the window of the P2G that made the tiles (its tiles' reach), as for S.
hipError_t launch_grid_update(const Pool& pool, bool use_s, int drift, hipStream_t s);
// Folds the tiles into A over S (A then holds the whole P2G result, for reads).
hipError_t launch_resolve_tiles(const Pool& pool, int drift, hipStream_t s);
// Zeroes A over S (use_s; with s_stamped only the S blocks the last P2G wrote through A) and D.
hipError_t launch_clear_a(const Pool& pool, bool use_s, bool s_stamped, hipStream_t s);
// G2P; ev_start / ev_stop (optional) are recorded by the dispatch itself at the kernel's start / end.
hipError_t launch_g2p(const Pool& pool, float4* P, uint32_t n, float vmax, int waves, hipStream_t s,
                      hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
// G2P over the work items of this substep's binned P2G, the velocities of each item's tile box
// staged in LDS. perm: sorted position -> record (nullptr: identity). sorted_out: slot i of dst
// takes record perm[i] of src (as launch_g2p_perm); otherwise in place (src == dst). keys_out
// (optional): the sort key of each new record at its slot (as launch_keys computes them after it).
hipError_t launch_g2p_items(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, bool sorted_out,
                            float vmax, int waves, hipStream_t s, hipEvent_t ev_start = nullptr,
                            hipEvent_t ev_stop = nullptr, uint32_t* keys_out = nullptr);
// G2P (as launch_g2p_items) fused with the next substep's P2G over the same work items: the new
// records are scattered into tiles (+ fallback atomics on A) without being read back. Needs
// launch_begin_p2g after the grid update; leaves the grid state of a binned P2G. keys_out: as
// launch_g2p_items (the sort keys of the new records).
// shape: 0 = 192-thread workgroups at 3 waves per SIMD, 1 = 256 threads at 4; drift: the window
// (Window<drift>, 2 or 3). Both result-neutral.
hipError_t launch_g2p_p2g(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, bool sorted_out,
                          float vmax, int shape, int drift, hipStream_t s, hipEvent_t ev_start = nullptr,
                          hipEvent_t ev_stop = nullptr, uint32_t* keys_out = nullptr);
// G2P that also applies a pending reorder: slot i of dst takes particle perm[i] of src.
hipError_t launch_g2p_perm(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, uint32_t n, float vmax,
                           int waves, hipStream_t s, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
// Reference-layout export of one scene: mode 0 = A (i32 words), mode 1 = V (f32 bits, 0 outside the
// node blocks the last grid update wrote).
hipError_t launch_grid_export(const Pool& pool, const SceneGeo& geo, int mode, int4* out, hipStream_t s);
// readPositions (mpm.js:624-630) on the device: pts[i] = (x * dx + origin in f64 rounded to f32,
// material id bits); vel (optional) = (v * vscale in f64 rounded to f32, J). Particles [offset, offset + n).
hipError_t launch_points_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base,
                             const double origin[3], double dx, double vscale, float4* pts, float4* vel, hipStream_t s);
// Writes one scene's velocities from reference-layout words (f32 bits of v.xyz, mass) into V.
hipError_t launch_grid_import(const Pool& pool, const SceneGeo& geo, const int4* in, hipStream_t s);

}  // namespace mpm
