// Runtime behind include/mpm.h: device context, particle/grid pools, substep orchestration.
//
// Replaces the WebGPU Context/Kernel machinery the reference path runs on (src/core/context.js,
// src/core/kernel.js) with one HIP stream per pool: every call enqueues, reads block. A pool holds
// any number of independent scenes that are advanced by the same launches (mpm_kernels.hpp); the
// reference's single MPM object (an mpm_sim) is a pool with one scene. The substep loop is
// MPM.step (src/physics/mpm.js:582-604).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mpm.h"
#include "mpm_kernels.hpp"
#include "pool_internal.hpp"
#include "prims.hpp"

using namespace mpm;

namespace {

thread_local std::string g_last_error;

mpm_status fail(mpm_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

mpm_status from_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return MPM_OK;
  mpm_status code = MPM_ERR_ERROR;
  if (e == hipErrorOutOfMemory) code = MPM_ERR_OUT_OF_MEMORY;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) code = MPM_ERR_NO_DEVICE;
  return fail(code, "%s: %s", where, hipGetErrorString(e));
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return from_hip(_e, #expr);  \
  } while (0)

#define ST_TRY(expr)              \
  do {                            \
    mpm_status _s = (expr);       \
    if (_s != MPM_OK) return _s;  \
  } while (0)

// What the grid buffers hold after the last call: nothing, the i32 P2G result in A, the P2G result
// split over tiles and A (inside a step only), or velocities in V.
enum GridState { kGridClean = 0, kGridA = 1, kGridTiles = 2, kGridVelocity = 3 };
enum BinMode { kBinOnly = 0, kBinReorder = 1, kBinReorderInG2P = 2 };


int ceil_log2(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

template <class T>
hipError_t dmalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

// A scene kept on the device for repeated admission (serving): its records in the reference layout
// (scene-local material ids) plus what mpm_scene_desc carries.
struct SceneTemplate {
  bool used = false;
  int32_t dims[3] = {0, 0, 0};
  int32_t ppc_axis = 2;
  float bound = 3.0f;
  float gravity[3] = {0, 0, 0};
  uint32_t nmat = 0;
  float mats[8 * 12] = {0};
  float4* rec = nullptr;  // count records, 8 float4 each
  uint32_t count = 0;
};

struct SceneHost {
  bool used = false;
  SceneGeo geo{};
  uint32_t count = 0;   // particles of the scene
  uint32_t offset = 0;  // first particle in the pool (scenes are contiguous)
  uint32_t nblocks = 0;
};

}  // namespace

struct mpm_ctx {
  int device = 0;
  char name[256] = {0};
  int cus = 0;
  hipStream_t stream = nullptr;
  uint32_t* scratch = nullptr;  // this device's address of `scratch` (below)
};

namespace {

// Scratch of mpm_scan_u32_noalloc (the scan's tile partial sums): a fixed array in the memory of
// every device the library runs on, held until the process ends, so that entry point calls no
// allocator. In host code `scratch` names a host-side handle, not a device address: its address
// on a device comes from hipGetSymbolAddress, taken once per device at mpm_ctx_create with that
// device current, and only that address is passed to kernels.
constexpr uint64_t kNoAllocScanMax = 1ull << 28;  // elements (MPM_SCAN_NOALLOC_MAX)
constexpr size_t kScratchWords = 65792;          // >= scan_scratch_words(kNoAllocScanMax) = 65556
static __device__ __align__(256) uint32_t scratch[kScratchWords];
std::mutex g_scratch_mu;
std::vector<uint32_t*> g_scratch_addr;  // per device index, nullptr until resolved

// This device's address of `scratch` (the device is current).
hipError_t scratch_address(int device, uint32_t** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)device >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_addr[(size_t)device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[(size_t)device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_addr[(size_t)device];
  return hipSuccess;
}

}  // namespace

struct mpm_pool {
  mpm_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  // configuration
  uint32_t cap = 0, max_scenes = 0, max_blocks = 0, sort_every = 4;
  int max_dim = 0;
  uint32_t key_bits = 0, sort_bits = 0, map_per_scene = 0, max_items = 0, tile_nodes = 0;
  // device memory
  float4* P[2] = {nullptr, nullptr};
  int cur = 0;
  int4* A = nullptr;
  float4* V = nullptr;
  uint2* block_info = nullptr;
  uint32_t *stamp = nullptr, *dlist = nullptr, *sstamp = nullptr, *slist = nullptr, *vstamp = nullptr;
  uint2* sinfo = nullptr;
  uint4* mtile = nullptr;
  uint32_t *wl_start = nullptr, *wl_block = nullptr;
  int4 *tiles = nullptr, *tile_hdr = nullptr;
  uint32_t* tile_off = nullptr;
  float *mats = nullptr, *cols = nullptr;
  SceneGeo* scenes_d = nullptr;
  uint32_t* ctr = nullptr;
  uint32_t* arena_ctr = nullptr;
  unsigned long long* occ = nullptr;
  uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr}, *order = nullptr, *flags = nullptr;
  uint32_t* sort_scratch = nullptr;
  unsigned long long* law = nullptr;  // grid-update momentum bookkeeping (kLawSlots x kLawWords), lazily
  bool law_on = false;
  float4* staging = nullptr;
  uint32_t staging_records = 0;
  int4* export_buf = nullptr;
  size_t export_nodes = 0;
  // host state
  std::vector<SceneHost> scenes;
  std::vector<uint32_t> mem_order;  // scenes in the order their particle runs sit in the pool
  std::vector<SceneTemplate> templates;
  std::vector<std::pair<uint32_t, uint32_t>> free_blocks;  // (start, length), sorted by start
  uint32_t count = 0;
  uint64_t step_count = 0;
  bool bin_valid = false;
  const uint32_t* perm = nullptr;  // sorted position -> particle (binning without reorder)
  bool reorder_in_g2p = false;     // perm is the sort's new order, applied by this substep's G2P
  bool needs_repack = false;       // a removed scene's particles are still in the pool
  int grid_state = kGridClean;
  bool a_dirty_s = false, a_dirty_d = false;  // A holds data over S / over D
  int g2p_waves = 3, p2g_waves = 4;
  int fuse_shape = 1;        // fused G2P + P2G workgroup shape (launch_g2p_p2g)
  int drift = 2;             // the work items' window (Window<drift>) in the binned and fused P2G
  float last_vmax = 0.9f;    // vmax of the last step (the autotuner's trials step with it)
  bool g2p_items = true;     // G2P by work item (LDS velocity boxes) after a binned P2G
  bool tiles_fresh = false;  // the tiles and work items are those of this substep's binned P2G
  bool fuse = true;          // G2P fused with the next substep's P2G when no sort comes between them
  bool p2g_ahead = false;    // the next substep's P2G is done (by a fused G2P of this mpm_step call, or
                             // of the previous one when it is carried)
  bool p2g_presort = false;  // ... before that substep's sort, over the items of the sort before it
  bool keys_ahead = false;   // keys[0] holds the sort keys of the current particles (written by the
                             // G2P before a sort in the same mpm_step call)
  // Work carried across calls: the last G2P of a step call also did the next substep's P2G (and,
  // before a sort, its keys). Kept only while nothing but reads and further steps follow; any call
  // that changes the pool discards it (drop_carry). carry_on: the knob; clean: no such call since the
  // last step call (a call that follows another this way carries its own last P2G forward).
  bool carry = false;
  bool carry_on = true;
  bool clean = false;
  bool use_graphs = false;
  mpm_grid_hook hook = nullptr;  // called after every grid update (not with graphs)
  void* hook_user = nullptr;
  bool mid_substep = false;      // between mpm_substep_begin and mpm_substep_end
  // one sort period, by parity of cur, whether its sort's P2G was done ahead (entry) and whether its
  // last G2P does the next one's (exit)
  hipGraphExec_t period_graph[2][2][2] = {};
  uint64_t graph_key = 0;                                // (count, scenes, vmax) the graphs were captured for
  // profiling 1: five events per profiled substep (sort, p2g, grid, g2p start; end);
  // profiling 2: two events per timed G2P launch, recorded by the dispatch itself (start, end)
  int profiling = 0;
  uint32_t prof_every = 1;  // profiling 2 times G2P launches 0, every, 2 every, ... since it was enabled
  uint32_t prof_seq = 0;    // G2P launches since profiling 2 was enabled
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  float phase_ms[4] = {0, 0, 0, 0};
  std::vector<uint8_t> ev_fused;    // profiling 2: per (start, end) pair, 1 for a fused G2P + P2G launch
  std::vector<uint8_t> ev_presort;  // profiling 1: per group of 5 marks, 1 when the grid update ran first
  float kernel_ms[2] = {0, 0};      // profiling 2: summed G2P launches {G2P alone, fused G2P + P2G}
  uint32_t kernel_launches[2] = {0, 0};
};

struct mpm_sim {
  mpm_pool* pool = nullptr;  // a pool with exactly one scene (id 0)
};

hipStream_t mpm::pool_stream(mpm_pool* pool) { return pool->stream; }

namespace {

mpm_status set_dev(mpm_ctx* c) {
  HIP_TRY(hipSetDevice(c->device));
  return MPM_OK;
}

Pool view(const mpm_pool* p) {
  Pool v;
  v.cap = p->cap;
  v.key_bits = p->key_bits;
  v.nblocks = p->max_blocks;
  v.max_items = p->max_items;
  v.tile_nodes = p->tile_nodes;
  v.A = p->A;
  v.V = p->V;
  v.block_info = p->block_info;
  v.stamp = p->stamp;
  v.dlist = p->dlist;
  v.sstamp = p->sstamp;
  v.vstamp = p->vstamp;
  v.slist = p->slist;
  v.sinfo = p->sinfo;
  v.mtile = p->mtile;
  v.wl_start = p->wl_start;
  v.wl_block = p->wl_block;
  v.tiles = p->tiles;
  v.tile_hdr = p->tile_hdr;
  v.tile_off = p->tile_off;
  v.mats = p->mats;
  v.cols = p->cols;
  v.scenes = p->scenes_d;
  v.ctr = p->ctr;
  v.arena_ctr = p->arena_ctr;
  v.occ = p->occ;
  v.law = p->law_on ? p->law : nullptr;
  return v;
}

void drop_graphs(mpm_pool* p) {
  for (auto& by_entry : p->period_graph)
    for (auto& by_exit : by_entry)
      for (auto& g : by_exit)
        if (g) {
          hipGraphExecDestroy(g);
          g = nullptr;
        }
}

void free_pool(mpm_pool* p) {
  if (!p) return;
  if (p->ctx) hipSetDevice(p->ctx->device);
  if (p->stream) hipStreamSynchronize(p->stream);
  drop_graphs(p);
  for (auto& t : p->templates)
    if (t.rec) hipFree(t.rec);
  for (hipEvent_t e : p->ev_pool) hipEventDestroy(e);
  void* ptrs[] = {p->P[0],   p->P[1],     p->A,        p->V,        p->block_info, p->stamp,   p->dlist,
                  p->sstamp, p->slist,    p->wl_start, p->wl_block,    p->tiles,   p->tile_hdr, p->tile_off,
                  p->mats,   p->cols,     p->scenes_d, p->ctr, p->arena_ctr,      p->occ,      p->keys[0],     p->keys[1], p->vals[0],
                  p->vals[1], p->order,   p->flags,    p->sort_scratch, p->staging, p->export_buf, p->law,
                  p->sinfo,   p->mtile,   p->vstamp};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (p->stream) hipStreamDestroy(p->stream);
  delete p;
}

// Device bytes a pool of this shape allocates (used for admission against an HBM budget).
// 64-bit words of one scene's particle-block occupancy bitmap (mps = pb^3 block slots).
uint32_t occ_words(uint32_t mps) {
  uint32_t pb = 1;
  while (pb * pb * pb < mps) pb *= 2;
  return pb * pb * occ_row_words(pb);
}

size_t pool_bytes(uint32_t cap, uint32_t max_scenes, uint32_t max_blocks, int max_dim, uint32_t* map_per_scene,
                  uint32_t* max_items, uint32_t* tile_nodes) {
  const int bits = std::max(3, ceil_log2((uint64_t)max_dim));
  const uint32_t mps = 1u << (3 * (bits - 2));
  const uint64_t items = std::min<uint64_t>((uint64_t)mps * max_scenes, cap);
  // tile arena: a full 4^3-cell block (512 particles) needs 216-343 nodes, a sparse surface block a
  // few nodes per particle; 4 nodes per particle (64 B) covers both, and an item whose tile does not
  // fit takes the fallback path
  const uint64_t arena = std::min<uint64_t>(0x7fffffffu, std::max<uint64_t>((uint64_t)4096 * kTileNodes, 4ull * cap)) /
                         kArenaRegions * kArenaRegions;
  if (map_per_scene) *map_per_scene = mps;
  if (max_items) *max_items = (uint32_t)items;
  if (tile_nodes) *tile_nodes = (uint32_t)arena;
  size_t b = 0;
  b += 2 * (size_t)cap * 128;                            // P[2]
  b += (size_t)max_blocks * 64 * (16 + 16);              // A, V
  b += (size_t)max_blocks * 4 * 7;                       // block_info, stamp, dlist, sstamp, slist, vstamp
  b += (size_t)max_blocks * 8;                           // sinfo
  b += (size_t)mps * max_scenes * 16;                    // mtile
  b += (size_t)occ_words(mps) * max_scenes * 8;          // occupancy bitmap
  b += (items + 1) * 4 * 2;                              // wl_start, wl_block
  b += arena * sizeof(int4) + items * 20;               // tile arena, headers, offsets
  b += (size_t)max_scenes * (kMatRows + kColRows) * 48;  // mats, cols
  b += (size_t)cap * 4 * 6 + 4;                          // keys, vals, order, flags
  b += sort_scratch_words(cap) * 4;
  b += (size_t)std::min<uint32_t>(cap, 1u << 20) * 128;  // staging
  return b;
}

mpm_status upload_scene(mpm_pool* p, uint32_t s) {
  static_assert(sizeof(SceneGeo) % 4 == 0 && sizeof(SceneGeo) <= 256, "SceneGeo words");
  HIP_TRY(launch_write_words(reinterpret_cast<uint32_t*>(p->scenes_d + s),
                             reinterpret_cast<const uint32_t*>(&p->scenes[s].geo), sizeof(SceneGeo) / 4, p->stream));
  drop_graphs(p);
  return MPM_OK;
}

// Material rows of scene s (n <= 8 rows of 12 floats) at rows [8s, 8s + 8).
mpm_status upload_materials(mpm_pool* p, uint32_t s, const float* tbl, uint32_t n) {
  float rows[kMatRows * 12];
  std::memset(rows, 0, sizeof rows);
  if (n) std::memcpy(rows, tbl, 48 * (size_t)n);
  uint32_t* dst = reinterpret_cast<uint32_t*>(p->mats + (size_t)s * kMatRows * 12);
  HIP_TRY(launch_write_words(dst, reinterpret_cast<const uint32_t*>(rows), 48, p->stream));
  HIP_TRY(launch_write_words(dst + 48, reinterpret_cast<const uint32_t*>(rows) + 48, 48, p->stream));
  return MPM_OK;
}

mpm_status upload_colliders(mpm_pool* p, uint32_t s, const float* tbl, uint32_t n) {
  float rows[kColRows * 12];
  std::memset(rows, 0, sizeof rows);
  if (n) std::memcpy(rows, tbl, 48 * (size_t)n);
  uint32_t* dst = reinterpret_cast<uint32_t*>(p->cols + (size_t)s * kColRows * 12);
  for (int part = 0; part < 3; part++)
    HIP_TRY(launch_write_words(dst + 64 * part, reinterpret_cast<const uint32_t*>(rows) + 64 * part, 64, p->stream));
  return MPM_OK;
}

bool alloc_blocks(mpm_pool* p, uint32_t n, uint32_t* base) {
  for (size_t i = 0; i < p->free_blocks.size(); i++) {
    auto& f = p->free_blocks[i];
    if (f.second >= n) {
      *base = f.first;
      f.first += n;
      f.second -= n;
      if (f.second == 0) p->free_blocks.erase(p->free_blocks.begin() + (long)i);
      return true;
    }
  }
  return false;
}

void release_blocks(mpm_pool* p, uint32_t base, uint32_t n) {
  auto& fb = p->free_blocks;
  auto it = std::lower_bound(fb.begin(), fb.end(), std::make_pair(base, 0u));
  it = fb.insert(it, {base, n});
  size_t i = (size_t)(it - fb.begin());
  if (i + 1 < fb.size() && fb[i].first + fb[i].second == fb[i + 1].first) {
    fb[i].second += fb[i + 1].second;
    fb.erase(fb.begin() + (long)i + 1);
  }
  if (i > 0 && fb[i - 1].first + fb[i - 1].second == fb[i].first) {
    fb[i - 1].second += fb[i].second;
    fb.erase(fb.begin() + (long)i);
  }
}

uint32_t free_block_total(const mpm_pool* p) {
  uint32_t t = 0;
  for (auto& f : p->free_blocks) t += f.second;
  return t;
}

// Scene offsets from the order the scenes' particle runs sit in.
void layout_offsets(mpm_pool* p) {
  uint32_t off = 0;
  for (uint32_t s : p->mem_order) {
    p->scenes[s].offset = off;
    off += p->scenes[s].count;
  }
}

// After a sort's reorder the particles are grouped by scene in id order (the scene is the top of
// the sort key).
void recompute_offsets(mpm_pool* p) {
  p->mem_order.clear();
  for (uint32_t s = 0; s < p->scenes.size(); s++)
    if (p->scenes[s].used) p->mem_order.push_back(s);
  layout_offsets(p);
}

// Return A to all-zero where the last P2G left data (before S or D change, and before the next P2G's
// epoch): a P2G's writes through A are in the blocks of D and, for a binned P2G, in blocks of S that
// carry its epoch's stamp (S blocks are not listed in D); after a fold of the tiles (a_dirty_s) A
// holds data over all of S. A P2G whose grid update never runs (a carried P2G dropped by a change
// between calls, a P2G read out by mpm_p2g) leaves such data behind.
mpm_status clean_a(mpm_pool* p) {
  if (p->a_dirty_s || p->a_dirty_d) HIP_TRY(launch_clear_a(view(p), true, !p->a_dirty_s, p->stream));
  p->a_dirty_d = p->a_dirty_s = false;
  return MPM_OK;
}

mpm_status apply_pending_reorder(mpm_pool* p) {
  if (!p->reorder_in_g2p) return MPM_OK;
  HIP_TRY(launch_permute(p->P[p->cur], p->P[p->cur ^ 1], p->cap, p->count, p->perm, p->stream));
  p->cur ^= 1;
  p->perm = nullptr;
  p->reorder_in_g2p = false;
  recompute_offsets(p);
  return MPM_OK;
}

// Sort keys, then reorder the particles now (MPM._sort, mpm.js:614-621), at this substep's G2P, or
// not at all (the sorted order is kept as a permutation), and rebuild work items and S.
mpm_status repack(mpm_pool* p);

// Discards the P2G carried from the last step call (its tiles, A and keys are left to the next P2G,
// which starts from a clean arena and clears A), and marks the pool changed since that call. Every
// entry point that changes particles, materials, scenes, the step count, the sort period, the hook,
// the grid or the kernel variants calls it first; reads and step calls do not.
void drop_carry(mpm_pool* p) {
  p->clean = false;
  if (!p->carry) return;
  p->carry = false;
  p->p2g_ahead = p->keys_ahead = p->p2g_presort = false;
  p->tiles_fresh = false;
  p->grid_state = kGridVelocity;  // V still holds the last grid update's velocities
}

mpm_status do_bin(mpm_pool* p, int mode) {
  if (p->needs_repack) {  // binning assumes every particle belongs to a live scene
    p->keys_ahead = false;
    ST_TRY(repack(p));
  }
  ST_TRY(clean_a(p));
  const Pool pv = view(p);
  p->perm = nullptr;
  p->reorder_in_g2p = false;
  if (p->count == 0) {
    HIP_TRY(launch_begin_bin(pv, p->stream));
    p->bin_valid = true;
    return MPM_OK;
  }
  if (!p->keys_ahead) HIP_TRY(launch_keys(pv, p->P[p->cur], p->count, p->keys[0], p->stream));
  p->keys_ahead = false;
  uint32_t *rk = nullptr, *rv = nullptr;
  // by block (key bits 6 and up), values the current indices; the cell order inside a block is made
  // by the transposition
  HIP_TRY(radix_sort_pairs(p->keys[0], p->vals[0], p->keys[1], p->vals[1], p->count, 6, (int)p->sort_bits,
                           p->sort_scratch, p->stream, &rk, &rv, true, true));
  // a new binning epoch, S and the items reset, the occupancy bitmap cleared
  HIP_TRY(launch_bin_reset(pv, 2 * occ_words(p->map_per_scene) * p->max_scenes, p->stream));
  {  // work items: heads per tile, scanned, then numbered (the sort's scratch is free again)
    const uint32_t tiles = bin_tiles(p->count);
    uint32_t* tile_first = p->sort_scratch;
    HIP_TRY(launch_bin_count(rk, p->count, tile_first, p->stream));
    const bool scan = !bin_emit_scans(p->count);  // few tiles: the emit adds up the counts itself
    if (scan) HIP_TRY(scan_exclusive(tile_first, (size_t)tiles + 1, p->sort_scratch + ((tiles + 4) & ~3u), p->stream));
    HIP_TRY(launch_bin_emit(pv, rk, p->count, tile_first, scan, p->stream));
  }
  HIP_TRY(launch_bin_transpose(pv, rk, rv, p->order, p->stream));
  if (mode == kBinReorder) {
    HIP_TRY(launch_permute(p->P[p->cur], p->P[p->cur ^ 1], p->cap, p->count, p->order, p->stream));
    p->cur ^= 1;
    recompute_offsets(p);
  } else {
    p->perm = p->order;
    p->reorder_in_g2p = mode == kBinReorderInG2P;
  }
  HIP_TRY(launch_bin_dilate(pv, p->max_scenes, p->drift, p->stream));
  p->bin_valid = true;
  return MPM_OK;
}

// Drops the particles of removed scenes with a stable partition (live before dead), so the live
// particles keep their order and the other scenes are not perturbed.
mpm_status repack(mpm_pool* p) {
  ST_TRY(clean_a(p));
  ST_TRY(apply_pending_reorder(p));
  const Pool pv = view(p);
  if (p->count) {
    HIP_TRY(launch_live_keys(pv, p->P[p->cur], p->count, p->keys[0], p->vals[0], p->stream));
    uint32_t *rk = nullptr, *rv = nullptr;
    HIP_TRY(radix_sort_pairs(p->keys[0], p->vals[0], p->keys[1], p->vals[1], p->count, 0, 1, p->sort_scratch, p->stream,
                             &rk, &rv));
    uint32_t live = 0;
    for (auto& sc : p->scenes)
      if (sc.used && sc.geo.active) live += sc.count;
    HIP_TRY(launch_permute(p->P[p->cur], p->P[p->cur ^ 1], p->cap, live, rv, p->stream));
    p->cur ^= 1;
    p->count = live;
  }
  for (uint32_t s = 0; s < p->scenes.size(); s++) {
    SceneHost& sc = p->scenes[s];
    if (sc.used && !sc.geo.active) {
      release_blocks(p, sc.geo.block_base, sc.nblocks);
      sc = SceneHost();
      p->mem_order.erase(std::find(p->mem_order.begin(), p->mem_order.end(), s));
    }
  }
  layout_offsets(p);
  p->needs_repack = false;
  p->bin_valid = false;
  p->grid_state = kGridClean;
  drop_graphs(p);
  return MPM_OK;
}

// P2G (mpm.js:607-612). Binned: into tiles (+ A for stragglers), grid state kGridTiles. Global: every
// contribution a global atomic on A, grid state kGridA.
// keys_out (binned only): the sort keys of the records, written by the P2G.
mpm_status do_p2g(mpm_pool* p, bool force_global, uint32_t* keys_out = nullptr) {
  if (p->p2g_ahead && !force_global) {  // done by the fused G2P of the previous substep
    p->p2g_ahead = false;
    return MPM_OK;
  }
  p->p2g_ahead = false;
  ST_TRY(clean_a(p));
  if (!force_global && !p->bin_valid) ST_TRY(do_bin(p, kBinOnly));
  if (force_global) ST_TRY(apply_pending_reorder(p));
  const Pool pv = view(p);
  HIP_TRY(launch_begin_p2g(pv, p->stream));
  if (force_global) {
    HIP_TRY(launch_p2g_global(pv, p->P[p->cur], p->count, p->stream));
    p->grid_state = kGridA;
    p->tiles_fresh = false;
  } else {
    HIP_TRY(launch_p2g_binned(pv, p->P[p->cur], p->perm, p->p2g_waves, p->drift, p->stream, keys_out));
    p->grid_state = kGridTiles;
    p->tiles_fresh = true;
  }
  p->a_dirty_d = true;
  return MPM_OK;
}

// Grid update (mpm.js:593-595) from tiles + A into V; A returns to zero.
mpm_status do_grid(mpm_pool* p) {
  HIP_TRY(launch_grid_update(view(p), p->grid_state == kGridTiles, p->drift, p->stream));
  p->a_dirty_d = false;
  p->grid_state = kGridVelocity;
  return MPM_OK;
}

// The next unused profiling event of the pool.
mpm_status next_event(mpm_pool* p, hipEvent_t* out) {
  if (p->ev_used == p->ev_pool.size()) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    p->ev_pool.push_back(e);
  }
  *out = p->ev_pool[p->ev_used++];
  return MPM_OK;
}

mpm_status mark(mpm_pool* p) {
  if (p->profiling != 1) return MPM_OK;
  hipEvent_t e;
  ST_TRY(next_event(p, &e));
  HIP_TRY(hipEventRecord(e, p->stream));
  return MPM_OK;
}

// G2P (mpm.js:598-600). fuse_next: the next substep follows in the same call and does not sort, so
// its P2G may be done by this G2P (same work items). keys_next: the next substep sorts; this G2P
// writes the sort keys. presort_next (with keys_next): the next substep follows in the same call, and
// its P2G may be done by this G2P too, over the current work items, ahead of its sort (P2G's sums
// do not depend on the particle order or on the items they are gathered by); that substep then
// runs its grid update before the sort rebuilds the work items and S.
mpm_status do_g2p(mpm_pool* p, float vmax, bool fuse_next, bool keys_next, bool presort_next) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (p->profiling == 2 && p->prof_seq++ % p->prof_every == 0) {
    ST_TRY(next_event(p, &e0));
    ST_TRY(next_event(p, &e1));
  }
  const bool items = p->g2p_items && p->tiles_fresh;
  p->tiles_fresh = false;  // one G2P per P2G: the G2P work counters are those begin_p2g cleared
  const bool presort = items && keys_next && presort_next && p->fuse && !p->hook;
  if (e0) p->ev_fused.push_back(items && (fuse_next || presort) && p->fuse ? 1 : 0);
  if (items && (fuse_next || presort) && p->fuse) {
    ST_TRY(clean_a(p));
    HIP_TRY(launch_begin_p2g(view(p), p->stream));  // the next P2G's epoch, arena and work counters
    const bool sorted_out = p->reorder_in_g2p;
    HIP_TRY(launch_g2p_p2g(view(p), p->P[p->cur], p->P[p->cur ^ (sorted_out ? 1 : 0)], p->perm, sorted_out, vmax,
                           p->fuse_shape, p->drift, p->stream, e0, e1, presort ? p->keys[0] : nullptr));
    p->keys_ahead = presort;
    p->p2g_presort = presort;
    if (sorted_out) {
      p->cur ^= 1;
      p->perm = nullptr;
      p->reorder_in_g2p = false;
      recompute_offsets(p);
    }
    p->grid_state = kGridTiles;
    p->a_dirty_d = true;
    p->tiles_fresh = true;
    p->p2g_ahead = true;
    return MPM_OK;
  }
  if (p->reorder_in_g2p) {  // the sort's reorder, applied by the record writes
    if (items)
      HIP_TRY(launch_g2p_items(view(p), p->P[p->cur], p->P[p->cur ^ 1], p->perm, true, vmax, p->g2p_waves, p->stream,
                               e0, e1));
    else
      HIP_TRY(launch_g2p_perm(view(p), p->P[p->cur], p->P[p->cur ^ 1], p->perm, p->count, vmax, p->g2p_waves, p->stream,
                              e0, e1));
    p->cur ^= 1;
    p->perm = nullptr;
    p->reorder_in_g2p = false;
    recompute_offsets(p);
    return MPM_OK;
  }
  if (items) {  // in place; perm (kBinOnly) maps the items' sorted positions to the records
    HIP_TRY(launch_g2p_items(view(p), p->P[p->cur], p->P[p->cur], p->perm, false, vmax, p->g2p_waves, p->stream, e0, e1,
                             keys_next ? p->keys[0] : nullptr));
    p->keys_ahead = keys_next;
  }
  else
    HIP_TRY(launch_g2p(view(p), p->P[p->cur], p->count, vmax, p->g2p_waves, p->stream, e0, e1));
  return MPM_OK;
}


// First part of a substep (MPM.step's loop body up to the grid update and the coupling call,
// mpm.js:591-596): sort when due, P2G, grid update, then the after-grid-update hook if one is set.
mpm_status substep_front(mpm_pool* p, uint32_t index) {
  if (p->p2g_presort) {  // P2G done ahead of this substep's sort (no hook is set): grid update, then sort
    p->p2g_presort = false;
    ST_TRY(do_p2g(p, false));  // consumes p2g_ahead; nothing is launched
    if (p->profiling == 1) p->ev_presort.push_back(1);  // marks: grid, (P2G), sort
    ST_TRY(mark(p));
    ST_TRY(do_grid(p));
    ST_TRY(mark(p));
    ST_TRY(mark(p));
    ST_TRY(do_bin(p, kBinReorderInG2P));  // the keys came with the P2G
    // the new items' velocity boxes are their blocks' stencil boxes (written with the items)
    p->tiles_fresh = true;
    return MPM_OK;
  }
  if (p->profiling == 1) p->ev_presort.push_back(0);
  ST_TRY(mark(p));
  if (p->step_count % p->sort_every == 0) {
    ST_TRY(do_bin(p, kBinReorderInG2P));
  } else if (!p->bin_valid) {
    ST_TRY(do_bin(p, kBinOnly));
  }
  ST_TRY(mark(p));
  ST_TRY(do_p2g(p, false));
  ST_TRY(mark(p));
  const bool use_s = p->grid_state == kGridTiles;
  ST_TRY(do_grid(p));
  if (p->hook) {
    const SceneGeo& geo = p->scenes[0].geo;
    mpm_grid_view v;
    std::memset(&v, 0, sizeof v);
    v.velocity = p->V;
    v.block_base = geo.block_base;
    v.blocks[0] = geo.nbx; v.blocks[1] = geo.nby; v.blocks[2] = geo.nbz;
    v.dims[0] = geo.X; v.dims[1] = geo.Y; v.dims[2] = geo.Z;
    v.s_blocks = use_s ? p->slist : nullptr;
    v.d_blocks = p->dlist;
    v.counts = p->ctr + kCtrD;
    v.stream = p->stream;
    v.substep = index;
    v.step_count = p->step_count;
    p->hook(p->hook_user, &v);
  }
  return MPM_OK;
}

// Rest of a substep: G2P (mpm.js:598-601).
mpm_status substep_back(mpm_pool* p, float vmax, bool fuse_next = false, bool keys_next = false,
                        bool presort_next = false) {
  ST_TRY(mark(p));
  ST_TRY(do_g2p(p, vmax, fuse_next, keys_next, presort_next));
  ST_TRY(mark(p));
  p->step_count++;
  return MPM_OK;
}

// One substep on the stream (MPM.step's loop body, mpm.js:591-601).
// more: another substep follows in this call. presort_ok: that substep runs eagerly (not from a
// captured period), so its P2G may be done ahead of its sort.
mpm_status substep(mpm_pool* p, float vmax, uint32_t index = 0, bool more = false, bool presort_ok = false) {
  ST_TRY(substep_front(p, index));
  // another substep follows in this call: fuse this G2P with its P2G when it does not sort, else
  // let this G2P write the sort keys (and, eagerly, do the P2G ahead of the sort as well)
  const bool next_sorts = (p->step_count + 1) % p->sort_every == 0;
  return substep_back(p, vmax, more && !next_sorts, more && next_sorts, more && next_sorts && presort_ok);
}

// One sort period (sort_every substeps starting at a sort) captured once per parity of the current
// particle buffer, entry (the sort's P2G already done by the G2P before it) and exit (the period's
// last G2P does the next sort's P2G), and replayed; valid while the pool's particle count, scenes
// and vmax do not change.
mpm_status period_graph(mpm_pool* p, float vmax, bool exit_ahead, hipGraphExec_t* out) {
  uint32_t vbits;
  std::memcpy(&vbits, &vmax, 4);
  const uint64_t key = ((uint64_t)p->count << 32) ^ (uint64_t)vbits;
  if (key != p->graph_key) {
    drop_graphs(p);
    p->graph_key = key;
  }
  const int parity = p->cur, entry = p->p2g_presort ? 1 : 0, exit = exit_ahead ? 1 : 0;
  hipGraphExec_t& slot = p->period_graph[parity][entry][exit];
  if (slot) {
    *out = slot;
    return MPM_OK;
  }
  // capture one period from the current host-side state, then restore that state: the launch below
  // replays exactly what was captured
  const uint64_t step0 = p->step_count;
  const bool ahead0 = p->p2g_ahead, keys0 = p->keys_ahead, presort0 = p->p2g_presort, fresh0 = p->tiles_fresh;
  const bool dirty_d0 = p->a_dirty_d, dirty_s0 = p->a_dirty_s, bin0 = p->bin_valid, reorder0 = p->reorder_in_g2p;
  const int grid0 = p->grid_state;
  const uint32_t* const perm0 = p->perm;
  HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  mpm_status st = MPM_OK;
  for (uint32_t k = 0; k < p->sort_every && st == MPM_OK; k++) {
    const bool last = k + 1 == p->sort_every;
    st = substep(p, vmax, k, !last || exit_ahead, last && exit_ahead);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(p->stream, &g);
  p->step_count = step0;
  p->cur = parity;
  p->p2g_ahead = ahead0;
  p->keys_ahead = keys0;
  p->p2g_presort = presort0;
  p->tiles_fresh = fresh0;
  p->a_dirty_d = dirty_d0;
  p->a_dirty_s = dirty_s0;
  p->bin_valid = bin0;
  p->reorder_in_g2p = reorder0;
  p->grid_state = grid0;
  p->perm = perm0;
  recompute_offsets(p);
  if (st != MPM_OK) {
    if (g) hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return from_hip(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&slot, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return from_hip(e, "hipGraphInstantiate");
  *out = slot;
  return MPM_OK;
}

mpm_status pool_step(mpm_pool* p, uint32_t substeps, float vmax) {
  if (substeps == 0) return MPM_OK;
  // work ahead is only kept within one call, or carried from the previous call (see carry)
  if (!p->carry) p->p2g_ahead = p->keys_ahead = p->p2g_presort = false;
  p->carry = false;
  // a call that follows another with nothing but reads between them carries its last P2G forward
  const bool carry_out = p->clean && p->carry_on && p->fuse && !p->hook && p->profiling != 1;
  p->clean = false;
  p->last_vmax = vmax;
  if (p->needs_repack) ST_TRY(repack(p));
  if (p->count == 0) return MPM_OK;  // mpm.js:584
  // A call that starts with a sort while the previous sort's work items still cover the records in
  // their order does that substep's P2G over those items first (as a G2P before a sort does within a
  // call), which also writes the sort keys; the substep then runs its grid update before the sort. The
  // P2G reads the records in order instead of through the new permutation; its sums do not depend
  // on the order or on the items.
  if (!p->p2g_ahead && p->fuse && p->g2p_items && !p->hook && p->profiling != 1 && p->bin_valid && !p->perm &&
      !p->reorder_in_g2p && p->step_count % p->sort_every == 0) {
    ST_TRY(do_p2g(p, false, p->keys[0]));
    p->p2g_ahead = p->keys_ahead = p->p2g_presort = true;
  }
  uint32_t k = 0;
  while (k < substeps) {
    const bool at_sort = p->step_count % p->sort_every == 0;
    // a sort period from a captured graph
    if (p->use_graphs && !p->profiling && !p->hook && at_sort && substeps - k >= p->sort_every) {
      // the period's last G2P does the next sort's P2G when another substep follows in this call (or
      // the next call's, carried), if G2P fuses at all (with MPM_TUNE_FUSE or MPM_TUNE_G2P_ITEMS off it
      // does not, and the next substep runs its own P2G)
      const bool exit_ahead = (substeps - k > p->sort_every || carry_out) && p->fuse && p->g2p_items;
      if (!p->p2g_presort) {  // (otherwise A holds part of the P2G done ahead, and its keys are set)
        ST_TRY(clean_a(p));
        ST_TRY(apply_pending_reorder(p));
        p->keys_ahead = false;  // the captured period computes its own sort keys
      }
      hipGraphExec_t ge = nullptr;
      ST_TRY(period_graph(p, vmax, exit_ahead, &ge));
      HIP_TRY(hipGraphLaunch(ge, p->stream));
      // host-side effects of one period: the sort's reorder swapped the buffers once
      p->cur ^= 1;
      p->perm = nullptr;
      p->reorder_in_g2p = false;
      p->bin_valid = true;
      p->a_dirty_s = false;
      p->grid_state = exit_ahead ? kGridTiles : kGridVelocity;
      p->a_dirty_d = exit_ahead;
      p->tiles_fresh = exit_ahead;
      p->p2g_ahead = p->keys_ahead = p->p2g_presort = exit_ahead;
      recompute_offsets(p);
      p->step_count += p->sort_every;
      k += p->sort_every;
    } else {
      const bool more = k + 1 < substeps || carry_out;
      ST_TRY(substep(p, vmax, k, more, more));
      k++;
    }
  }
  p->carry = carry_out && p->p2g_ahead;
  p->clean = true;
  return MPM_OK;
}

mpm_status create_pool(mpm_ctx* c, uint32_t cap, uint32_t max_scenes, uint32_t max_blocks, int max_dim,
                       uint32_t sort_every, mpm_pool** out) {
  *out = nullptr;
  if (cap == 0 || cap > (1u << 30)) return fail(MPM_ERR_INVALID_ARGUMENT, "pool: max_particles = %u outside [1, 2^30]", cap);
  if (max_scenes == 0 || max_scenes > 4096) return fail(MPM_ERR_INVALID_ARGUMENT, "pool: max_scenes = %u outside [1, 4096]", max_scenes);
  if (max_dim < 8 || max_dim > 1024) return fail(MPM_ERR_INVALID_ARGUMENT, "pool: max_dim = %d outside [8, 1024]", max_dim);
  if (sort_every == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "pool: sort_every = 0");
  const int bits = std::max(3, ceil_log2((uint64_t)max_dim));
  const int scene_bits = ceil_log2(max_scenes);
  if (3 * bits + scene_bits > 32)
    return fail(MPM_ERR_INVALID_ARGUMENT, "pool: %u scenes of edge %d do not fit a 32-bit sort key", max_scenes, max_dim);
  ST_TRY(set_dev(c));
  mpm_pool* p = new (std::nothrow) mpm_pool;
  if (!p) return fail(MPM_ERR_OUT_OF_MEMORY, "pool: host allocation");
  p->ctx = c;
  p->cap = cap;
  p->max_scenes = max_scenes;
  p->max_blocks = max_blocks;
  p->max_dim = max_dim;
  p->sort_every = sort_every;
  p->key_bits = (uint32_t)(3 * bits);
  p->sort_bits = p->key_bits + (uint32_t)scene_bits;
  pool_bytes(cap, max_scenes, max_blocks, max_dim, &p->map_per_scene, &p->max_items, &p->tile_nodes);
  p->scenes.resize(max_scenes);
  p->free_blocks.push_back({0u, max_blocks});
  const size_t nodes = (size_t)max_blocks * 64;
  hipError_t e = hipStreamCreate(&p->stream);
  if (e == hipSuccess) e = dmalloc(&p->P[0], (size_t)cap * 8);
  if (e == hipSuccess) e = dmalloc(&p->P[1], (size_t)cap * 8);
  if (e == hipSuccess) e = dmalloc(&p->A, nodes);
  if (e == hipSuccess) e = dmalloc(&p->V, nodes);
  if (e == hipSuccess) e = dmalloc(&p->block_info, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->stamp, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->dlist, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->sstamp, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->vstamp, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->slist, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->sinfo, max_blocks);
  if (e == hipSuccess) e = dmalloc(&p->mtile, (size_t)p->map_per_scene * max_scenes);
  if (e == hipSuccess) e = dmalloc(&p->occ, (size_t)occ_words(p->map_per_scene) * max_scenes);
  if (e == hipSuccess) e = dmalloc(&p->wl_start, (size_t)p->max_items + 1);
  if (e == hipSuccess) e = dmalloc(&p->wl_block, p->max_items);
  if (e == hipSuccess) e = dmalloc(&p->tiles, (size_t)p->tile_nodes);
  if (e == hipSuccess) e = dmalloc(&p->tile_hdr, p->max_items);
  if (e == hipSuccess) e = dmalloc(&p->tile_off, p->max_items);
  if (e == hipSuccess) e = dmalloc(&p->mats, (size_t)max_scenes * kMatRows * 12);
  if (e == hipSuccess) e = dmalloc(&p->cols, (size_t)max_scenes * kColRows * 12);
  if (e == hipSuccess) e = dmalloc(&p->scenes_d, max_scenes);
  if (e == hipSuccess) e = dmalloc(&p->ctr, kCounters);
  if (e == hipSuccess) e = dmalloc(&p->arena_ctr, (size_t)kArenaCtrWords);
  if (e == hipSuccess) e = dmalloc(&p->keys[0], cap);
  if (e == hipSuccess) e = dmalloc(&p->keys[1], cap);
  if (e == hipSuccess) e = dmalloc(&p->vals[0], cap);
  if (e == hipSuccess) e = dmalloc(&p->vals[1], cap);
  if (e == hipSuccess) e = dmalloc(&p->order, cap);
  if (e == hipSuccess) e = dmalloc(&p->flags, (size_t)cap + 1);
  if (e == hipSuccess)
    e = dmalloc(&p->sort_scratch, std::max(sort_scratch_words(cap), scan_scratch_words((size_t)cap + 1)));
  if (e == hipSuccess) {
    p->staging_records = std::min<uint32_t>(cap, 1u << 20);
    e = dmalloc(&p->staging, (size_t)p->staging_records * 8);
  }
  if (e == hipSuccess) e = hipMemsetAsync(p->A, 0, nodes * sizeof(int4), p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->V, 0, nodes * sizeof(float4), p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->block_info, 0xff, (size_t)max_blocks * 8, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->stamp, 0xff, (size_t)max_blocks * 4, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->sstamp, 0xff, (size_t)max_blocks * 4, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->vstamp, 0xff, (size_t)max_blocks * 4, p->stream);
  if (e == hipSuccess)  // epoch 0xffffffff: no P2G has written the entry
    e = hipMemsetAsync(p->mtile, 0xff, (size_t)p->map_per_scene * max_scenes * sizeof(uint4), p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->ctr, 0, kCounters * 4, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->arena_ctr, 0, (size_t)kArenaCtrWords * 4, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->mats, 0, (size_t)max_scenes * kMatRows * 48, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->cols, 0, (size_t)max_scenes * kColRows * 48, p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->scenes_d, 0, (size_t)max_scenes * sizeof(SceneGeo), p->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->P[0], 0, (size_t)cap * 128, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) {
    mpm_status code = from_hip(e, "pool allocation");
    free_pool(p);
    return code;
  }
  *out = p;
  return MPM_OK;
}

// Adds a scene: geometry, node blocks, material rows (particles are written by the caller).
mpm_status add_scene(mpm_pool* p, const int32_t dims[3], int32_t ppc_axis, float bound, const float g[3],
                     const float* mats, uint32_t nmat, uint32_t* id) {
  for (int a = 0; a < 3; a++)
    if (dims[a] < 8 || dims[a] > p->max_dim)
      return fail(MPM_ERR_INVALID_ARGUMENT, "scene: dims[%d] = %d outside [8, %d]", a, dims[a], p->max_dim);
  if (ppc_axis < 1 || ppc_axis > 8) return fail(MPM_ERR_INVALID_ARGUMENT, "scene: ppc_axis = %d outside [1, 8]", ppc_axis);
  // bound >= 3 keeps every stencil inside the grid: G2P clamps x to [bound-1, dims-bound+1]
  // (mpm.js:502), so the base cell floor(x - 1/2) stays in [1, dims-3].
  if (!(bound >= 3.0f && bound <= 16.0f)) return fail(MPM_ERR_INVALID_ARGUMENT, "scene: bound = %g outside [3, 16]", (double)bound);
  if (nmat > (uint32_t)kMatRows) return fail(MPM_ERR_INVALID_ARGUMENT, "scene: at most 8 materials (got %u)", nmat);
  uint32_t s = 0;
  while (s < p->scenes.size() && p->scenes[s].used) s++;
  if (s == p->scenes.size()) return fail(MPM_ERR_OUT_OF_MEMORY, "pool: all %u scene slots are in use", p->max_scenes);
  SceneHost sc;
  SceneGeo& geo = sc.geo;
  geo.X = dims[0]; geo.Y = dims[1]; geo.Z = dims[2];
  geo.nbx = (dims[0] + 3) / 4; geo.nby = (dims[1] + 3) / 4; geo.nbz = (dims[2] + 3) / 4;
  geo.mbits = std::max(1, ceil_log2((uint64_t)std::max(dims[0], std::max(dims[1], dims[2]))) - 2);
  sc.nblocks = (uint32_t)geo.nbx * geo.nby * geo.nbz;
  if (!alloc_blocks(p, sc.nblocks, &geo.block_base))
    return fail(MPM_ERR_OUT_OF_MEMORY, "pool: no run of %u free node blocks (%u free in total)", sc.nblocks,
                free_block_total(p));
  geo.map_base = s * p->map_per_scene;
  geo.bound = bound;
  geo.g[0] = g[0]; geo.g[1] = g[1]; geo.g[2] = g[2];
  geo.floor_friction = 0.4f;
  geo.ncol = 0;
  geo.vol4 = (float)(4.0 * (1.0 / std::pow((double)ppc_axis, 3)));  // mpm.js:610
  geo.active = 1;
  sc.used = true;
  sc.count = 0;
  sc.offset = p->count;
  p->scenes[s] = sc;
  p->mem_order.push_back(s);
  ST_TRY(upload_scene(p, s));
  ST_TRY(upload_materials(p, s, mats, nmat));
  ST_TRY(upload_colliders(p, s, nullptr, 0));
  HIP_TRY(launch_block_info(p->block_info, geo.block_base, sc.nblocks, s, (uint32_t)geo.nbx, (uint32_t)geo.nby, p->stream));
  HIP_TRY(launch_fill_u32(p->stamp + geo.block_base, 0xffffffffu, sc.nblocks, p->stream));
  HIP_TRY(launch_fill_u32(p->sstamp + geo.block_base, 0xffffffffu, sc.nblocks, p->stream));
  HIP_TRY(launch_fill_u32(p->vstamp + geo.block_base, 0xffffffffu, sc.nblocks, p->stream));
  *id = s;
  return MPM_OK;
}

// Positions inside [0.5, dims - 1.5) (the range MPM.add and G2P keep them in) and material ids < 8.
mpm_status validate_positions(const int32_t dims[3], const float* rec, uint32_t n, uint32_t first = 0) {
  for (uint32_t i = 0; i < n; i++) {
    const float* r = rec + 32 * (size_t)i;
    for (int a = 0; a < 3; a++)
      if (!(r[a] >= 0.5f && r[a] < (float)dims[a] - 1.5f))
        return fail(MPM_ERR_INVALID_ARGUMENT, "record %u position[%d] = %g outside [0.5, %d)", first + i, a,
                    (double)r[a], dims[a] - 1);
    uint32_t mid;
    std::memcpy(&mid, r + 3, 4);
    if (mid >= (uint32_t)kMatRows) return fail(MPM_ERR_INVALID_ARGUMENT, "record %u material %u >= 8", first + i, mid);
  }
  return MPM_OK;
}

mpm_status validate_records(const mpm_pool* p, uint32_t s, const float* rec, uint32_t n, uint32_t first) {
  const SceneGeo& geo = p->scenes[s].geo;
  const int32_t dims[3] = {geo.X, geo.Y, geo.Z};
  return validate_positions(dims, rec, n, first);
}

// Writes n records (reference layout, scene-local material ids) to pool slots [at, at + n).
mpm_status write_records(mpm_pool* p, uint32_t s, const float* rec, uint32_t at, uint32_t n) {
  ST_TRY(apply_pending_reorder(p));
  for (uint32_t done = 0; done < n;) {
    const uint32_t k = std::min(n - done, p->staging_records);
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(p->staging, rec + 32 * (size_t)done, (size_t)k * 128, hipMemcpyHostToDevice));
    HIP_TRY(launch_records_in(p->staging, p->P[p->cur], p->cap, at + done, k, s * kMatRows, p->stream));
    done += k;
  }
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->bin_valid = false;
  return MPM_OK;
}

mpm_status read_records(mpm_pool* p, uint32_t s, float* rec, uint32_t at, uint32_t n) {
  ST_TRY(apply_pending_reorder(p));
  for (uint32_t done = 0; done < n;) {
    const uint32_t k = std::min(n - done, p->staging_records);
    HIP_TRY(launch_records_out(p->P[p->cur], p->cap, at + done, k, s * kMatRows, p->staging, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(rec + 32 * (size_t)done, p->staging, (size_t)k * 128, hipMemcpyDeviceToHost));
    done += k;
  }
  return MPM_OK;
}

mpm_status check_materials(const float* mats, uint32_t n, const char* what) {
  if (n && !mats) return fail(MPM_ERR_INVALID_ARGUMENT, "%s: materials NULL", what);
  if (n > (uint32_t)kMatRows) return fail(MPM_ERR_INVALID_ARGUMENT, "%s: at most 8 materials (got %u)", what, n);
  for (uint32_t k = 0; k < n; k++) {
    uint32_t kind;
    std::memcpy(&kind, mats + 12 * k, 4);
    if (kind > 3) return fail(MPM_ERR_INVALID_ARGUMENT, "%s: material %u has kind %u (0-3)", what, k, kind);
  }
  return MPM_OK;
}

// Admits a scene: slot, node blocks, geometry, materials, then its records appended to the pool
// from host memory (validated) or from a device template. Out-of-memory leaves the pool unchanged.
mpm_status admit(mpm_pool* p, const int32_t dims[3], int32_t ppc, float bound, const float g[3], const float* mats,
                 uint32_t nmat, const float* host_rec, const float4* dev_rec, uint32_t count, uint32_t* scene) {
  if (p->needs_repack) ST_TRY(repack(p));
  if ((uint64_t)p->count + count > p->cap)
    return fail(MPM_ERR_OUT_OF_MEMORY, "pool: %u particles do not fit (%u of %u used)", count, p->count, p->cap);
  uint32_t id = 0;
  ST_TRY(add_scene(p, dims, ppc, bound, g, mats, nmat, &id));
  mpm_status st = host_rec ? validate_records(p, id, host_rec, count, 0) : MPM_OK;
  if (st == MPM_OK && count) {
    if (host_rec) {
      st = write_records(p, id, host_rec, p->count, count);
    } else {
      st = apply_pending_reorder(p);
      if (st == MPM_OK) {
        hipError_t e = launch_records_in(dev_rec, p->P[p->cur], p->cap, p->count, count, id * kMatRows, p->stream);
        st = from_hip(e, "template admission");
        p->bin_valid = false;
      }
    }
  }
  if (st != MPM_OK) {
    release_blocks(p, p->scenes[id].geo.block_base, p->scenes[id].nblocks);
    p->scenes[id] = SceneHost();
    p->mem_order.pop_back();
    return st;
  }
  p->scenes[id].offset = p->count;
  p->scenes[id].count = count;
  p->count += count;
  p->grid_state = kGridClean;
  drop_graphs(p);
  *scene = id;
  return MPM_OK;
}

// Lattice gravity and floor friction of scene 0 (a simulation's only scene).
mpm_status set_step_params(mpm_pool* p, const float g[3], float ff) {
  SceneGeo& geo = p->scenes[0].geo;
  if (geo.g[0] != g[0] || geo.g[1] != g[1] || geo.g[2] != g[2] || geo.floor_friction != ff) {
    geo.g[0] = g[0]; geo.g[1] = g[1]; geo.g[2] = g[2];
    geo.floor_friction = ff;
    ST_TRY(upload_scene(p, 0));
  }
  return MPM_OK;
}

#define CHECK_SIM(s)                                                        \
  do {                                                                      \
    if (!(s)) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");  \
    ST_TRY(set_dev((s)->pool->ctx));                                        \
  } while (0)

#define CHECK_POOL(p)                                                 \
  do {                                                                \
    if (!(p)) return fail(MPM_ERR_INVALID_ARGUMENT, "pool is NULL");  \
    ST_TRY(set_dev((p)->ctx));                                        \
  } while (0)

}  // namespace

extern "C" {

const char* mpm_status_string(mpm_status s) {
  switch (s) {
    case MPM_OK: return "ok";
    case MPM_ERR_INVALID_ARGUMENT: return "invalid-argument";
    case MPM_ERR_UNSUPPORTED: return "unsupported";
    case MPM_ERR_NO_DEVICE: return "no-device";
    case MPM_ERR_ERROR: return "error";
    case MPM_ERR_OUT_OF_MEMORY: return "out-of-memory";
    case MPM_ERR_DESTROYED: return "destroyed";
    default: return "unknown";
  }
}

const char* mpm_last_error_message(void) { return g_last_error.c_str(); }

mpm_status mpm_ctx_create(int device, mpm_ctx** out) {
  if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(MPM_ERR_NO_DEVICE, "mpm_ctx_create: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(MPM_ERR_NO_DEVICE, "mpm_ctx_create: device %d of %d", device, n);
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, device));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return fail(MPM_ERR_UNSUPPORTED, "mpm_ctx_create: device %d is %s; this library is built for gfx950 only", device,
                p.gcnArchName);
  uint32_t* scr = nullptr;
  HIP_TRY(scratch_address(device, &scr));
  mpm_ctx* c = new (std::nothrow) mpm_ctx;
  if (!c) return fail(MPM_ERR_OUT_OF_MEMORY, "mpm_ctx_create: host allocation");
  c->device = device;
  c->scratch = scr;
  if (p.name[0])  // the marketing name can be empty (then the device index stands in)
    std::snprintf(c->name, sizeof c->name, "%s (%s)", p.name, p.gcnArchName);
  else
    std::snprintf(c->name, sizeof c->name, "HIP device %d (%s)", device, p.gcnArchName);
  c->cus = p.multiProcessorCount;
  e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    delete c;
    return from_hip(e, "hipStreamCreate");
  }
  *out = c;
  return MPM_OK;
}

void mpm_ctx_destroy(mpm_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) {
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
  }
  delete c;
}

mpm_status mpm_ctx_info(mpm_ctx* c, char* name, size_t name_len, int32_t* num_cus) {
  if (!c) return fail(MPM_ERR_INVALID_ARGUMENT, "context is NULL");
  if (name && name_len) {
    std::strncpy(name, c->name, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (num_cus) *num_cus = c->cus;
  return MPM_OK;
}

mpm_status mpm_sim_create(mpm_ctx* c, const mpm_config* cfg, mpm_sim** out) {
  mpm_pool* p = nullptr;
  if (!c || !cfg || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: NULL argument");
  *out = nullptr;
  for (int a = 0; a < 3; a++)
    if (cfg->dims[a] < 8 || cfg->dims[a] > 1024)
      return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: dims[%d] = %d outside [8, 1024]", a, cfg->dims[a]);
  if (cfg->ppc_axis < 1 || cfg->ppc_axis > 8)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: ppc_axis = %d outside [1, 8]", cfg->ppc_axis);
  if (!(cfg->bound >= 3.0f && cfg->bound <= 16.0f))
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: bound = %g outside [3, 16]", (double)cfg->bound);
  const int maxd = std::max(cfg->dims[0], std::max(cfg->dims[1], cfg->dims[2]));
  const uint32_t blocks = (uint32_t)((cfg->dims[0] + 3) / 4) * ((cfg->dims[1] + 3) / 4) * ((cfg->dims[2] + 3) / 4);
  ST_TRY(create_pool(c, cfg->max_particles, 1, blocks, maxd, cfg->sort_every, &p));
  const float g0[3] = {0, 0, 0};
  uint32_t id = 0;
  mpm_status st = add_scene(p, cfg->dims, cfg->ppc_axis, cfg->bound, g0, nullptr, 0, &id);
  if (st != MPM_OK) {
    free_pool(p);
    return st;
  }
  mpm_sim* s = new (std::nothrow) mpm_sim;
  if (!s) {
    free_pool(p);
    return fail(MPM_ERR_OUT_OF_MEMORY, "mpm_sim_create: host allocation");
  }
  s->pool = p;
  *out = s;
  return MPM_OK;
}

void mpm_sim_destroy(mpm_sim* s) {
  if (!s) return;
  free_pool(s->pool);
  delete s;
}

mpm_status mpm_set_materials(mpm_sim* s, const float* tbl, uint32_t n) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (n > 8) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_materials: at most 8 materials (got %u)", n);
  if (n && !tbl) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_materials: table is NULL");
  for (uint32_t k = 0; k < n; k++) {
    uint32_t kind;
    std::memcpy(&kind, tbl + 12 * k, 4);
    if (kind > 3) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_materials: row %u has kind %u (0-3)", k, kind);
  }
  return upload_materials(s->pool, 0, tbl, n);
}

mpm_status mpm_set_colliders(mpm_sim* s, const float* tbl, uint32_t n) {
  CHECK_SIM(s);
  if (n > 16) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_colliders: at most 16 colliders (got %u)", n);
  if (n && !tbl) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_colliders: table is NULL");
  mpm_pool* p = s->pool;
  ST_TRY(upload_colliders(p, 0, tbl, n));
  if (p->scenes[0].geo.ncol != n) {
    p->scenes[0].geo.ncol = n;
    ST_TRY(upload_scene(p, 0));
  }
  return MPM_OK;
}

mpm_status mpm_write_particles(mpm_sim* s, const float* rec, uint32_t offset, uint32_t n) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (s->pool->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: inside a substep begun with mpm_substep_begin");
  if (n == 0) return MPM_OK;
  if (!rec) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: records NULL");
  mpm_pool* p = s->pool;
  if ((uint64_t)offset + n > p->cap)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: offset %u + n %u exceeds capacity %u", offset, n, p->cap);
  ST_TRY(validate_records(p, 0, rec, n, offset));
  return write_records(p, 0, rec, offset, n);
}

mpm_status mpm_read_particles(mpm_sim* s, float* rec, uint32_t offset, uint32_t n) {
  CHECK_SIM(s);
  if (n == 0) return MPM_OK;
  if (!rec) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_particles: records NULL");
  if ((uint64_t)offset + n > s->pool->cap)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_particles: offset %u + n %u exceeds capacity %u", offset, n,
                s->pool->cap);
  return read_records(s->pool, 0, rec, offset, n);
}

mpm_status mpm_set_count(mpm_sim* s, uint32_t n) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (s->pool->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_count: inside a substep begun with mpm_substep_begin");
  mpm_pool* p = s->pool;
  if (n > p->cap) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_count: %u exceeds capacity %u", n, p->cap);
  if (n != p->count) {
    ST_TRY(apply_pending_reorder(p));
    p->bin_valid = false;
    drop_graphs(p);
  }
  p->count = n;
  p->scenes[0].count = n;
  return MPM_OK;
}

mpm_status mpm_get_count(mpm_sim* s, uint32_t* n) {
  if (!s || !n) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_get_count: NULL argument");
  *n = s->pool->count;
  return MPM_OK;
}

mpm_status mpm_set_step_count(mpm_sim* s, uint64_t n) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  drop_carry(s->pool);
  s->pool->step_count = n;
  return MPM_OK;
}

mpm_status mpm_set_sort_every(mpm_sim* s, uint32_t n) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (n == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_sort_every: 0");
  if (s->pool->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_sort_every: inside a substep begun with mpm_substep_begin");
  if (s->pool->sort_every != n) {
    s->pool->sort_every = n;
    drop_graphs(s->pool);  // a captured period is sort_every substeps long
  }
  return MPM_OK;
}

mpm_status mpm_get_step_count(mpm_sim* s, uint64_t* n) {
  if (!s || !n) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_get_step_count: NULL argument");
  *n = s->pool->step_count;
  return MPM_OK;
}

mpm_status mpm_step(mpm_sim* s, uint32_t substeps, const float g[3], float ff, float vmax) {
  CHECK_SIM(s);
  if (!g) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_step: gravity is NULL");
  mpm_pool* p = s->pool;
  if (p->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_step: inside a substep begun with mpm_substep_begin");
  ST_TRY(set_step_params(p, g, ff));
  return pool_step(p, substeps, vmax);
}

mpm_status mpm_p2g(mpm_sim* s, int32_t force_global) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (s->pool->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_p2g: inside a substep begun with mpm_substep_begin");
  mpm_pool* p = s->pool;
  ST_TRY(do_p2g(p, force_global != 0));
  if (p->grid_state == kGridTiles) {  // fold the tiles into A so the grid reads as the i32 P2G result
    HIP_TRY(launch_resolve_tiles(view(p), p->drift, p->stream));
    p->grid_state = kGridA;
    p->a_dirty_s = true;
  }
  return MPM_OK;
}

mpm_status mpm_p2g_to(mpm_sim* s, int32_t force_global, int32_t* words) {
  CHECK_SIM(s);
  if (!words) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_p2g_to: words NULL");
  ST_TRY(mpm_p2g(s, force_global));
  return mpm_read_grid(s, words);
}

mpm_status mpm_sort(mpm_sim* s) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (s->pool->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sort: inside a substep begun with mpm_substep_begin");
  return do_bin(s->pool, kBinReorder);
}

mpm_status mpm_rescale(mpm_sim* s, float k) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (s->pool->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_rescale: inside a substep begun with mpm_substep_begin");
  mpm_pool* p = s->pool;
  ST_TRY(apply_pending_reorder(p));
  HIP_TRY(launch_rescale(p->P[p->cur], p->cap, 0, p->count, k, p->stream));
  return MPM_OK;
}

mpm_status mpm_set_grid_hook(mpm_sim* s, mpm_grid_hook hook, void* user) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  s->pool->hook = hook;
  s->pool->hook_user = user;
  drop_graphs(s->pool);
  return MPM_OK;
}

mpm_status mpm_substep_begin(mpm_sim* s, const float g[3], float ff) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (!g) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_substep_begin: gravity is NULL");
  mpm_pool* p = s->pool;
  if (p->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_substep_begin: the previous substep was not ended");
  ST_TRY(set_step_params(p, g, ff));
  if (p->needs_repack) ST_TRY(repack(p));
  if (p->count == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_substep_begin: no particles");
  p->p2g_ahead = p->keys_ahead = p->p2g_presort = false;  // work ahead is only kept within one mpm_step
  ST_TRY(substep_front(p, 0));
  p->mid_substep = true;
  return MPM_OK;
}

mpm_status mpm_substep_end(mpm_sim* s, float vmax) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  mpm_pool* p = s->pool;
  if (!p->mid_substep) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_substep_end: no substep was begun");
  p->mid_substep = false;
  return substep_back(p, vmax);
}

mpm_status mpm_write_grid(mpm_sim* s, const int32_t* words) {
  CHECK_SIM(s);
  drop_carry(s->pool);
  if (!words) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_grid: words NULL");
  mpm_pool* p = s->pool;
  if (p->grid_state != kGridVelocity)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_grid: the grid holds no velocities (call after a grid update)");
  const SceneGeo& geo = p->scenes[0].geo;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  if (p->export_nodes < nodes) {
    if (p->export_buf) hipFree(p->export_buf);
    p->export_buf = nullptr;
    HIP_TRY(dmalloc(&p->export_buf, nodes));
    p->export_nodes = nodes;
  }
  HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipMemcpy(p->export_buf, words, nodes * 16, hipMemcpyHostToDevice));
  HIP_TRY(launch_grid_import(view(p), geo, p->export_buf, p->stream));
  return MPM_OK;
}

mpm_status mpm_read_points(mpm_sim* s, float* pts, float* vel, uint32_t offset, uint32_t n, const double origin[3],
                           double dx, double vscale) {
  CHECK_SIM(s);
  if (n == 0) return MPM_OK;
  if (!pts || !origin) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_points: NULL argument");
  mpm_pool* p = s->pool;
  if ((uint64_t)offset + n > p->count)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_points: offset %u + n %u exceeds count %u", offset, n, p->count);
  ST_TRY(apply_pending_reorder(p));
  // staging holds staging_records * 8 float4: room for 4 * staging_records points and velocities
  const uint32_t chunk = p->staging_records * 4;
  for (uint32_t done = 0; done < n;) {
    const uint32_t k = std::min(n - done, chunk);
    float4* dp = p->staging;
    float4* dv = vel ? p->staging + chunk : nullptr;
    HIP_TRY(launch_points_out(p->P[p->cur], p->cap, offset + done, k, 0, origin, dx, vscale, dp, dv, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(pts + 4 * (size_t)done, dp, (size_t)k * 16, hipMemcpyDeviceToHost));
    if (vel) HIP_TRY(hipMemcpy(vel + 4 * (size_t)done, dv, (size_t)k * 16, hipMemcpyDeviceToHost));
    done += k;
  }
  return MPM_OK;
}

mpm_status mpm_particle_streams(mpm_sim* s, void* streams[8], uint32_t* capacity) {
  CHECK_SIM(s);
  if (!streams) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_particle_streams: NULL argument");
  mpm_pool* p = s->pool;
  ST_TRY(apply_pending_reorder(p));
  HIP_TRY(hipStreamSynchronize(p->stream));
  for (int k = 0; k < 8; k++) streams[k] = p->P[p->cur] + (size_t)k * p->cap;
  if (capacity) *capacity = p->cap;
  return MPM_OK;
}

// ---- reads that resolve later (P.read / grid.read are Promises in the reference) ----
struct mpm_read {
  int device = 0;
  void* staging = nullptr;  // device memory, stream-ordered allocation on the simulation's stream
  size_t bytes = 0;
  hipEvent_t done = nullptr;
};

}  // extern "C"

namespace {

// Allocates `bytes` of staging on the simulation's stream, lets `fill` enqueue the snapshot into it,
// and records the event the finish waits for.
template <class Fill>
mpm_status begin_read(mpm_pool* p, size_t bytes, mpm_read** out, Fill&& fill) {
  mpm_read* r = new (std::nothrow) mpm_read;
  if (!r) return fail(MPM_ERR_OUT_OF_MEMORY, "read: host allocation");
  r->device = p->ctx->device;
  r->bytes = bytes;
  hipError_t e = hipEventCreateWithFlags(&r->done, hipEventDisableTiming);
  if (e == hipSuccess && bytes) e = hipMallocAsync(&r->staging, bytes, p->stream);
  if (e == hipSuccess && bytes) e = fill(r->staging);
  if (e == hipSuccess) e = hipEventRecord(r->done, p->stream);
  if (e != hipSuccess) {
    if (r->staging) hipFreeAsync(r->staging, p->stream);
    if (r->done) hipEventDestroy(r->done);
    delete r;
    return from_hip(e, "read");
  }
  *out = r;
  return MPM_OK;
}

}  // namespace

extern "C" {

mpm_status mpm_read_particles_begin(mpm_sim* s, uint32_t offset, uint32_t n, mpm_read** out) {
  CHECK_SIM(s);
  if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_particles_begin: out is NULL");
  mpm_pool* p = s->pool;
  if ((uint64_t)offset + n > p->cap)
    return fail(MPM_ERR_INVALID_ARGUMENT, "P.read: offset %u + n %u exceeds capacity %u", offset, n, p->cap);
  ST_TRY(apply_pending_reorder(p));
  return begin_read(p, (size_t)n * 128, out, [&](void* stg) {
    return launch_records_out(p->P[p->cur], p->cap, offset, n, 0, static_cast<float4*>(stg), p->stream);
  });
}

mpm_status mpm_read_grid_begin(mpm_sim* s, mpm_read** out) {
  CHECK_SIM(s);
  if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_grid_begin: out is NULL");
  mpm_pool* p = s->pool;
  const SceneGeo& geo = p->scenes[0].geo;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  const int state = p->carry ? kGridVelocity : p->grid_state;  // a carried P2G leaves V as it was
  return begin_read(p, nodes * 16, out, [&](void* stg) {
    if (state != kGridA && state != kGridVelocity)
      return launch_fill_u32(static_cast<uint32_t*>(stg), 0u, (uint32_t)(nodes * 4), p->stream);
    return launch_grid_export(view(p), geo, state == kGridA ? 0 : 1, static_cast<int4*>(stg), p->stream);
  });
}

mpm_status mpm_read_points_begin(mpm_sim* s, uint32_t offset, uint32_t n, const double origin[3], double dx,
                                 double vscale, int32_t with_velocity, mpm_read** out) {
  CHECK_SIM(s);
  if (!out || !origin) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_points_begin: NULL argument");
  mpm_pool* p = s->pool;
  if ((uint64_t)offset + n > p->count)
    return fail(MPM_ERR_INVALID_ARGUMENT, "readPoints: offset %u + n %u exceeds count %u", offset, n, p->count);
  ST_TRY(apply_pending_reorder(p));
  return begin_read(p, (size_t)n * 16 * (with_velocity ? 2 : 1), out, [&](void* stg) {
    float4* pts = static_cast<float4*>(stg);
    return launch_points_out(p->P[p->cur], p->cap, offset, n, 0, origin, dx, vscale, pts,
                             with_velocity ? pts + n : nullptr, p->stream);
  });
}

mpm_status mpm_read_finish(mpm_read* r, void* dst) {
  if (!r) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_finish: NULL read");
  hipError_t e = hipSetDevice(r->device);
  hipStream_t cs = nullptr;
  if (e == hipSuccess) e = hipEventSynchronize(r->done);
  if (e == hipSuccess && r->bytes) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e == hipSuccess && r->bytes && dst) e = hipMemcpyAsync(dst, r->staging, r->bytes, hipMemcpyDeviceToHost, cs);
  if (r->staging) {  // freed in any case; without the copy stream, by a synchronising free
    const hipError_t f = cs ? hipFreeAsync(r->staging, cs) : hipFree(r->staging);
    if (e == hipSuccess) e = f;
  }
  if (cs) {
    const hipError_t f = hipStreamSynchronize(cs);
    if (e == hipSuccess) e = f;
    hipStreamDestroy(cs);
  }
  hipEventDestroy(r->done);
  delete r;
  return from_hip(e, "read");
}

mpm_status mpm_read_grid(mpm_sim* s, int32_t* words) {
  CHECK_SIM(s);
  if (!words) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_grid: words NULL");
  mpm_pool* p = s->pool;
  const SceneGeo& geo = p->scenes[0].geo;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  const int state = p->carry ? kGridVelocity : p->grid_state;  // a carried P2G leaves V as it was
  if (state != kGridA && state != kGridVelocity) {
    std::memset(words, 0, nodes * 16);
    return MPM_OK;
  }
  if (p->export_nodes < nodes) {
    if (p->export_buf) hipFree(p->export_buf);
    p->export_buf = nullptr;
    HIP_TRY(dmalloc(&p->export_buf, nodes));
    p->export_nodes = nodes;
  }
  HIP_TRY(launch_grid_export(view(p), geo, state == kGridA ? 0 : 1, p->export_buf, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipMemcpy(words, p->export_buf, nodes * 16, hipMemcpyDeviceToHost));
  return MPM_OK;
}

mpm_status mpm_sync(mpm_sim* s) {
  CHECK_SIM(s);
  HIP_TRY(hipStreamSynchronize(s->pool->stream));
  return MPM_OK;
}

mpm_status mpm_set_tuning(mpm_sim* s, int32_t knob, int32_t value) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  return mpm_pool_set_tuning(s->pool, knob, value);
}

mpm_status mpm_set_profiling(mpm_sim* s, int32_t enable) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  mpm_pool* p = s->pool;
  if (enable < 0 || enable > 2) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_profiling: mode %d not in 0..2", enable);
  drop_carry(p);
  p->profiling = enable;
  p->prof_seq = 0;
  p->ev_used = 0;
  p->ev_fused.clear();
  p->ev_presort.clear();
  for (float& m : p->phase_ms) m = 0.0f;
  p->kernel_ms[0] = p->kernel_ms[1] = 0.0f;
  p->kernel_launches[0] = p->kernel_launches[1] = 0;
  return MPM_OK;
}

mpm_status mpm_set_profiling_every(mpm_sim* s, uint32_t every) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  if (every == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_profiling_every: every = 0");
  s->pool->prof_every = every;
  s->pool->prof_seq = 0;
  return MPM_OK;
}

// Profiling 2: fold the recorded (start, end) pairs of the G2P launches into kernel_ms / phase_ms.
mpm_status collect_kernel_events(mpm_pool* p) {
  const size_t pairs = p->ev_used / 2;
  if (pairs) HIP_TRY(hipEventSynchronize(p->ev_pool[pairs * 2 - 1]));
  for (size_t g = 0; g < pairs; g++) {
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, p->ev_pool[2 * g], p->ev_pool[2 * g + 1]));
    const int f = g < p->ev_fused.size() ? p->ev_fused[g] : 0;
    p->kernel_ms[f] += t;
    p->kernel_launches[f]++;
    p->phase_ms[3] += t;
  }
  p->ev_used = 0;
  p->ev_fused.clear();
  return MPM_OK;
}

mpm_status mpm_kernel_times(mpm_sim* s, float ms[2], uint32_t launches[2]) {
  if (!s || !ms || !launches) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_kernel_times: NULL argument");
  mpm_pool* p = s->pool;
  ST_TRY(set_dev(p->ctx));
  if (p->profiling == 2) ST_TRY(collect_kernel_events(p));
  for (int k = 0; k < 2; k++) {
    ms[k] = p->kernel_ms[k];
    launches[k] = p->kernel_launches[k];
  }
  return MPM_OK;
}

mpm_status mpm_phase_times(mpm_sim* s, float ms[4]) {
  if (!s || !ms) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_phase_times: NULL argument");
  mpm_pool* p = s->pool;
  ST_TRY(set_dev(p->ctx));
  if (p->profiling == 2) {  // (start, end) of each G2P launch
    ST_TRY(collect_kernel_events(p));
  } else {
    const size_t groups = p->ev_used / 5;
    if (groups) HIP_TRY(hipEventSynchronize(p->ev_pool[groups * 5 - 1]));
    for (size_t g = 0; g < groups; g++) {
      float t[4] = {0, 0, 0, 0};
      for (int ph = 0; ph < 4; ph++) HIP_TRY(hipEventElapsedTime(&t[ph], p->ev_pool[g * 5 + ph], p->ev_pool[g * 5 + ph + 1]));
      if (g < p->ev_presort.size() && p->ev_presort[g]) std::swap(t[0], t[2]);  // that group ran grid, P2G, sort
      for (int ph = 0; ph < 4; ph++) p->phase_ms[ph] += t[ph];
    }
  }
  p->ev_used = 0;
  p->ev_presort.clear();
  for (int ph = 0; ph < 4; ph++) ms[ph] = p->phase_ms[ph];
  return MPM_OK;
}

// ---- pools ----
size_t mpm_pool_bytes(const mpm_pool_config* cfg) {
  if (!cfg) return 0;
  return pool_bytes(cfg->max_particles, cfg->max_scenes, cfg->max_node_blocks, cfg->max_dim, nullptr, nullptr, nullptr);
}

mpm_status mpm_pool_create(mpm_ctx* c, const mpm_pool_config* cfg, mpm_pool** out) {
  if (!c || !cfg || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_create: NULL argument");
  if (cfg->max_node_blocks == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_create: max_node_blocks = 0");
  return create_pool(c, cfg->max_particles, cfg->max_scenes, cfg->max_node_blocks, cfg->max_dim,
                     cfg->sort_every ? cfg->sort_every : 4, out);
}

void mpm_pool_destroy(mpm_pool* p) { free_pool(p); }

mpm_status mpm_pool_add_scene(mpm_pool* p, const mpm_scene_desc* d, uint32_t* scene) {
  CHECK_POOL(p);
  drop_carry(p);
  if (!d || !scene) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_add_scene: NULL argument");
  if (d->count && !d->records) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_add_scene: records NULL");
  ST_TRY(check_materials(d->materials, d->num_materials, "mpm_pool_add_scene"));
  return admit(p, d->dims, d->ppc_axis, d->bound, d->gravity, d->materials, d->num_materials, d->records, nullptr,
               d->count, scene);
}

mpm_status mpm_pool_add_template(mpm_pool* p, const mpm_scene_desc* d, uint32_t* tmpl) {
  CHECK_POOL(p);
  drop_carry(p);
  if (!d || !tmpl) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_add_template: NULL argument");
  if (d->count && !d->records) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_add_template: records NULL");
  ST_TRY(check_materials(d->materials, d->num_materials, "mpm_pool_add_template"));
  for (int a = 0; a < 3; a++)
    if (d->dims[a] < 8 || d->dims[a] > p->max_dim)
      return fail(MPM_ERR_INVALID_ARGUMENT, "template: dims[%d] = %d outside [8, %d]", a, d->dims[a], p->max_dim);
  ST_TRY(validate_positions(d->dims, d->records, d->count));
  SceneTemplate t;
  t.used = true;
  std::memcpy(t.dims, d->dims, sizeof t.dims);
  t.ppc_axis = d->ppc_axis;
  t.bound = d->bound;
  std::memcpy(t.gravity, d->gravity, sizeof t.gravity);
  t.nmat = d->num_materials;
  if (t.nmat) std::memcpy(t.mats, d->materials, 48 * (size_t)t.nmat);
  t.count = d->count;
  HIP_TRY(dmalloc(&t.rec, (size_t)d->count * 8));
  hipError_t e = d->count ? hipMemcpy(t.rec, d->records, (size_t)d->count * 128, hipMemcpyHostToDevice) : hipSuccess;
  if (e != hipSuccess) {
    hipFree(t.rec);
    return from_hip(e, "mpm_pool_add_template");
  }
  uint32_t k = 0;
  while (k < p->templates.size() && p->templates[k].used) k++;
  if (k == p->templates.size()) p->templates.push_back(t);
  else p->templates[k] = t;
  *tmpl = k;
  return MPM_OK;
}

mpm_status mpm_pool_remove_template(mpm_pool* p, uint32_t tmpl) {
  CHECK_POOL(p);
  drop_carry(p);
  if (tmpl >= p->templates.size() || !p->templates[tmpl].used)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_remove_template: no template %u", tmpl);
  HIP_TRY(hipStreamSynchronize(p->stream));
  hipFree(p->templates[tmpl].rec);
  p->templates[tmpl] = SceneTemplate();
  return MPM_OK;
}

mpm_status mpm_pool_add_scene_from_template(mpm_pool* p, uint32_t tmpl, uint32_t* scene) {
  CHECK_POOL(p);
  drop_carry(p);
  if (!scene) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_add_scene_from_template: NULL argument");
  if (tmpl >= p->templates.size() || !p->templates[tmpl].used)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_add_scene_from_template: no template %u", tmpl);
  const SceneTemplate& t = p->templates[tmpl];
  return admit(p, t.dims, t.ppc_axis, t.bound, t.gravity, t.mats, t.nmat, nullptr, t.rec, t.count, scene);
}

mpm_status mpm_pool_remove_scene(mpm_pool* p, uint32_t scene) {
  CHECK_POOL(p);
  drop_carry(p);
  if (scene >= p->scenes.size() || !p->scenes[scene].used || !p->scenes[scene].geo.active)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_remove_scene: no scene %u", scene);
  p->scenes[scene].geo.active = 0;
  ST_TRY(upload_scene(p, scene));
  p->needs_repack = true;
  return MPM_OK;
}

mpm_status mpm_pool_step(mpm_pool* p, uint32_t substeps, float floor_friction, float vmax) {
  CHECK_POOL(p);
  for (uint32_t s = 0; s < p->scenes.size(); s++) {
    SceneHost& sc = p->scenes[s];
    if (sc.used && sc.geo.active && sc.geo.floor_friction != floor_friction) {
      sc.geo.floor_friction = floor_friction;
      ST_TRY(upload_scene(p, s));
    }
  }
  return pool_step(p, substeps, vmax);
}

mpm_status mpm_pool_scene_count(mpm_pool* p, uint32_t scene, uint32_t* n) {
  if (!p || !n) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_scene_count: NULL argument");
  if (scene >= p->scenes.size() || !p->scenes[scene].used || !p->scenes[scene].geo.active)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_scene_count: no scene %u", scene);
  *n = p->scenes[scene].count;
  return MPM_OK;
}

mpm_status mpm_pool_read_scene(mpm_pool* p, uint32_t scene, float* rec, uint32_t capacity, uint32_t* n) {
  CHECK_POOL(p);
  if (scene >= p->scenes.size() || !p->scenes[scene].used || !p->scenes[scene].geo.active)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_read_scene: no scene %u", scene);
  ST_TRY(apply_pending_reorder(p));
  const SceneHost& sc = p->scenes[scene];
  if (n) *n = sc.count;
  if (!rec) return MPM_OK;
  return read_records(p, scene, rec, sc.offset, std::min(capacity, sc.count));
}

mpm_status mpm_pool_info(mpm_pool* p, uint32_t* particles, uint32_t* scenes, uint32_t* free_node_blocks) {
  if (!p) return fail(MPM_ERR_INVALID_ARGUMENT, "pool is NULL");
  uint32_t live = 0, parts = 0;
  for (auto& sc : p->scenes)
    if (sc.used && sc.geo.active) {
      live++;
      parts += sc.count;
    }
  if (particles) *particles = parts;
  if (scenes) *scenes = live;
  if (free_node_blocks) *free_node_blocks = free_block_total(p);
  return MPM_OK;
}

mpm_status mpm_pool_sync(mpm_pool* p) {
  CHECK_POOL(p);
  HIP_TRY(hipStreamSynchronize(p->stream));
  return MPM_OK;
}

mpm_status mpm_pool_use_graphs(mpm_pool* p, int32_t enable) {
  if (!p) return fail(MPM_ERR_INVALID_ARGUMENT, "pool is NULL");
  p->use_graphs = enable != 0;
  if (!p->use_graphs) drop_graphs(p);
  return MPM_OK;
}

mpm_status mpm_pool_autotune(mpm_pool* p, uint32_t trial_substeps, int32_t* g2p_waves, int32_t* p2g_waves) {
  CHECK_POOL(p);
  drop_carry(p);
  if (trial_substeps == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_autotune: trial_substeps = 0");
  const uint32_t trial = (trial_substeps + p->sort_every - 1) / p->sort_every * p->sort_every;
  const bool graphs = p->use_graphs;
  const int prof = p->profiling;
  const float vmax = p->last_vmax;  // the trials step as the caller does
  const bool carry_on = p->carry_on;
  p->use_graphs = false;
  p->profiling = 0;
  p->carry_on = false;  // every trial opens with its own P2G and ends with its own G2P
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  mpm_status st = MPM_OK;
  // one trial of `trial` substeps with the current knobs; returns its milliseconds
  auto time_trial = [&](float* ms) -> mpm_status {
    hipEventRecord(e0, p->stream);
    ST_TRY(pool_step(p, trial, vmax));
    hipEventRecord(e1, p->stream);
    hipEventSynchronize(e1);
    hipEventElapsedTime(ms, e0, e1);
    return MPM_OK;
  };
  // align to a sort boundary so every trial covers the same phases
  while (st == MPM_OK && p->step_count % p->sort_every != 0) st = pool_step(p, 1, vmax);
  // the fused G2P + P2G (3 substeps of 4) first: its workgroup shape; then the work items' window
  // (P2G's tile: drift 2 or 3); then the occupancy of the G2P without a fused P2G (a call's last
  // substep) and of the binned P2G (a call's first), each tried with the others at their best so far
  float best = 1e30f;
  int bf = p->fuse_shape;
  for (int f = 0; f < 2 && st == MPM_OK; f++) {
    p->fuse_shape = f;
    float ms = 0;
    st = time_trial(&ms);
    if (st == MPM_OK && ms < best) { best = ms; bf = f; }
  }
  p->fuse_shape = bf;
  best = 1e30f;
  int bd = p->drift;
  for (int d = 2; d <= 3 && st == MPM_OK; d++) {
    if (p->drift != d) p->bin_valid = false;  // S and the tiles' reach depend on the window
    p->drift = d;
    float ms = 0;
    st = time_trial(&ms);
    if (st == MPM_OK && ms < best) { best = ms; bd = d; }
  }
  if (p->drift != bd) p->bin_valid = false;
  p->drift = bd;
  best = 1e30f;
  int bg = p->g2p_waves, bp = p->p2g_waves;
  for (int g = 3; g <= 6 && st == MPM_OK; g++)
    for (int q = 4; q <= 6 && st == MPM_OK; q++) {
      p->g2p_waves = g;
      p->p2g_waves = q;
      float ms = 0;
      st = time_trial(&ms);
      if (st == MPM_OK && ms < best) { best = ms; bg = g; bp = q; }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  p->g2p_waves = bg;
  p->p2g_waves = bp;
  p->use_graphs = graphs;
  p->profiling = prof;
  p->carry_on = carry_on;
  drop_graphs(p);
  if (g2p_waves) *g2p_waves = bg;
  if (p2g_waves) *p2g_waves = bp;
  return st;
}

mpm_status mpm_pool_tuning(mpm_pool* p, int32_t knob, int32_t* value) {
  if (!p || !value) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_tuning: NULL argument");
  switch (knob) {
    case MPM_TUNE_G2P_WAVES: *value = p->g2p_waves; return MPM_OK;
    case MPM_TUNE_P2G_WAVES: *value = p->p2g_waves; return MPM_OK;
    case MPM_TUNE_G2P_ITEMS: *value = p->g2p_items ? 1 : 0; return MPM_OK;
    case MPM_TUNE_FUSE: *value = p->fuse ? 1 : 0; return MPM_OK;
    case MPM_TUNE_FUSE_SHAPE: *value = p->fuse_shape; return MPM_OK;
    case MPM_TUNE_DRIFT: *value = p->drift; return MPM_OK;
    case MPM_TUNE_CARRY: *value = p->carry_on ? 1 : 0; return MPM_OK;
    default: return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_tuning: unknown knob %d", knob);
  }
}

mpm_status mpm_pool_set_grid_law(mpm_pool* p, int32_t enable) {
  CHECK_POOL(p);
  const uint32_t words = 2u * kLawSlots * kLawWords;  // u32 words
  if (enable) {
    if (!p->law) HIP_TRY(dmalloc(&p->law, (size_t)kLawSlots * kLawWords));
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(p->law), 0u, words, p->stream));
  }
  if (p->law_on != (enable != 0)) drop_graphs(p);  // the grid update's arguments change
  p->law_on = enable != 0;
  return MPM_OK;
}

mpm_status mpm_pool_read_grid_law(mpm_pool* p, mpm_grid_law* out, int32_t reset) {
  CHECK_POOL(p);
  if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_grid_law: NULL argument");
  std::memset(out, 0, sizeof *out);
  if (!p->law) return MPM_OK;
  std::vector<unsigned long long> h((size_t)kLawSlots * kLawWords);
  HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipMemcpy(h.data(), p->law, h.size() * 8, hipMemcpyDeviceToHost));
  unsigned long long t[kLawWords] = {0};
  for (int s = 0; s < kLawSlots; s++)
    for (int k = 0; k < kLawWords; k++) t[k] += h[(size_t)s * kLawWords + k];  // wrapping sums
  out->mass_words = (int64_t)t[0];
  for (int c = 0; c < 3; c++) {
    out->p2g[c] = (int64_t)t[1 + c];
    out->momentum[c] = (int64_t)t[4 + c];
    out->gravity[c] = (int64_t)t[7 + c];
    out->boundary[c] = (int64_t)t[10 + c];
  }
  out->nodes = t[13];
  out->updates = t[14];
  if (reset) HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(p->law), 0u, 2u * kLawSlots * kLawWords, p->stream));
  return MPM_OK;
}

mpm_status mpm_set_grid_law(mpm_sim* s, int32_t enable) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  return mpm_pool_set_grid_law(s->pool, enable);
}

mpm_status mpm_read_grid_law(mpm_sim* s, mpm_grid_law* out, int32_t reset) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  return mpm_pool_read_grid_law(s->pool, out, reset);
}

mpm_status mpm_pool_diagnostics(mpm_pool* p, mpm_diagnostics* out) {
  CHECK_POOL(p);
  if (!out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_pool_diagnostics: NULL argument");
  uint32_t c[kCounters];
  std::vector<uint32_t> a((size_t)kArenaRegions * kArenaStride);
  HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipMemcpy(c, p->ctr, sizeof c, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(a.data(), p->arena_ctr, a.size() * 4, hipMemcpyDeviceToHost));
  uint64_t used = 0;
  for (int r = 0; r < kArenaRegions; r++) used += a[(size_t)r * kArenaStride];
  // the P2G kernels give every work item a slot of the window tile's size when all slots fit
  const uint64_t slots = (uint64_t)c[kCtrItems] * (p->drift == 2 ? Window<2>::kNodes : Window<3>::kNodes);
  if (slots <= p->tile_nodes) used = slots;
  out->work_items = c[kCtrItems];
  out->s_blocks = c[kCtrS];
  out->d_blocks = c[kCtrD];
  out->tile_nodes_used = (uint32_t)std::min<uint64_t>(used, 0xffffffffu);
  out->tile_nodes_capacity = p->tile_nodes;
  out->particles = p->count;
  return MPM_OK;
}

mpm_status mpm_diagnostics_of(mpm_sim* s, mpm_diagnostics* out) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  return mpm_pool_diagnostics(s->pool, out);
}

mpm_status mpm_pool_set_tuning(mpm_pool* p, int32_t knob, int32_t value) {
  if (!p) return fail(MPM_ERR_INVALID_ARGUMENT, "pool is NULL");
  drop_carry(p);
  switch (knob) {
    case MPM_TUNE_G2P_WAVES:
      if (value < 3 || value > 6) break;
      p->g2p_waves = value;
      drop_graphs(p);
      return MPM_OK;
    case MPM_TUNE_FUSE:
      if (value < 0 || value > 1) break;
      p->fuse = value != 0;
      drop_graphs(p);
      return MPM_OK;
    case MPM_TUNE_G2P_ITEMS:
      if (value < 0 || value > 1) break;
      p->g2p_items = value != 0;
      drop_graphs(p);
      return MPM_OK;
    case MPM_TUNE_P2G_WAVES:
      if (value < 4 || value > 6) break;
      p->p2g_waves = value;
      drop_graphs(p);
      return MPM_OK;
    case MPM_TUNE_FUSE_SHAPE:
      if (value < 0 || value > 1) break;
      p->fuse_shape = value;
      drop_graphs(p);
      return MPM_OK;
    case MPM_TUNE_DRIFT:
      if (value < 2 || value > 3) break;
      if (p->drift != value) p->bin_valid = false;  // S and the tiles' reach depend on the window
      p->drift = value;
      drop_graphs(p);
      return MPM_OK;
    case MPM_TUNE_CARRY:
      if (value < 0 || value > 1) break;
      p->carry_on = value != 0;
      return MPM_OK;
    default:
      return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_tuning: unknown knob %d", knob);
  }
  return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_tuning: value %d out of range for knob %d", value, knob);
}

mpm_status mpm_scan_u32(mpm_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n) {
  if (!c) return fail(MPM_ERR_INVALID_ARGUMENT, "context is NULL");
  if (n == 0) return MPM_OK;
  if (!in || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_scan_u32: NULL array");
  if (n > (1ull << 32)) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_scan_u32: n too large");
  mpm_status st = set_dev(c);
  if (st) return st;
  uint32_t *d = nullptr, *tmp = nullptr;
  hipError_t e = dmalloc(&d, n);
  if (e == hipSuccess) e = dmalloc(&tmp, scan_scratch_words(n));
  if (e == hipSuccess) e = hipMemcpy(d, in, n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = scan_exclusive(d, n, tmp, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d, n * 4, hipMemcpyDeviceToHost);
  if (d) hipFree(d);
  if (tmp) hipFree(tmp);
  return from_hip(e, "mpm_scan_u32");
}

mpm_status mpm_scan_u32_noalloc(mpm_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, void* stream) {
  if (!c) return fail(MPM_ERR_INVALID_ARGUMENT, "context is NULL");
  if (n == 0) return MPM_OK;
  if (!in || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_scan_u32_noalloc: NULL array");
  if (n > kNoAllocScanMax || scan_scratch_words(n) > kScratchWords)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_scan_u32_noalloc: n = %llu exceeds the maximum %llu (2^28)",
                (unsigned long long)n, (unsigned long long)kNoAllocScanMax);
  if (!c->scratch) return fail(MPM_ERR_ERROR, "mpm_scan_u32_noalloc: no scratch address for device %d", c->device);
  ST_TRY(set_dev(c));
  HIP_TRY(scan_exclusive_to(in, out, n, c->scratch, static_cast<hipStream_t>(stream)));
  return MPM_OK;
}

mpm_status mpm_counting_sort(mpm_ctx* c, const uint32_t* keys, uint64_t n, uint32_t num_buckets, uint32_t* order,
                             uint32_t* start) {
  if (!c) return fail(MPM_ERR_INVALID_ARGUMENT, "context is NULL");
  if (num_buckets == 0 || !start) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_counting_sort: need num_buckets >= 1 and start");
  if (n && (!keys || !order)) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_counting_sort: NULL array");
  if (n >= (1ull << 32)) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_counting_sort: n too large");
  for (uint64_t i = 0; i < n; i++)
    if (keys[i] >= num_buckets)
      return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_counting_sort: keys[%llu] = %u >= num_buckets %u", (unsigned long long)i,
                  keys[i], num_buckets);
  mpm_status st = set_dev(c);
  if (st) return st;
  int bits = 0;
  while (bits < 32 && (1ull << bits) < num_buckets) bits++;
  const size_t nb1 = (size_t)num_buckets + 1;
  uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *cnt = nullptr, *scr = nullptr;
  hipError_t e = dmalloc(&k0, n);
  if (e == hipSuccess) e = dmalloc(&k1, n);
  if (e == hipSuccess) e = dmalloc(&v0, n);
  if (e == hipSuccess) e = dmalloc(&v1, n);
  if (e == hipSuccess) e = dmalloc(&cnt, nb1);
  if (e == hipSuccess) e = dmalloc(&scr, std::max(sort_scratch_words(n), scan_scratch_words(nb1)));
  if (e == hipSuccess && n) e = hipMemcpy(k0, keys, n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, nb1 * 4, c->stream);
  if (e == hipSuccess) e = bucket_count(k0, n, cnt, c->stream);
  if (e == hipSuccess) e = scan_exclusive(cnt, nb1, scr, c->stream);
  uint32_t *rk = k0, *rv = v0;
  if (e == hipSuccess) e = iota_u32(v0, n, c->stream);
  if (e == hipSuccess) e = radix_sort_pairs(k0, v0, k1, v1, n, 0, bits, scr, c->stream, &rk, &rv);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && n) e = hipMemcpy(order, rv, n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(start, cnt, nb1 * 4, hipMemcpyDeviceToHost);
  for (uint32_t* p : {k0, k1, v0, v1, cnt, scr})
    if (p) hipFree(p);
  return from_hip(e, "mpm_counting_sort");
}

mpm_status mpm_spatial_grid(mpm_ctx* c, const float* points, uint32_t n, const double lo[3], const double hi[3],
                            double cell_size, const float* queries, uint32_t m, uint32_t* order, uint32_t* start,
                            uint32_t* offsets, uint32_t* neighbors, uint64_t capacity, uint64_t* total) {
  if (!c) return fail(MPM_ERR_INVALID_ARGUMENT, "context is NULL");
  if (!lo || !hi || !(cell_size > 0) || (n && !points) || (m && (!queries || !offsets)))
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_spatial_grid: bad argument");
  GridSpec g;
  uint64_t cells = 1;
  for (int a = 0; a < 3; a++) {  // grid.js:31-33
    g.origin[a] = (float)lo[a];
    g.dims[a] = (uint32_t)std::max(1.0, std::ceil((hi[a] - lo[a]) / cell_size));
    cells *= g.dims[a];
  }
  if (cells > (1ull << 26)) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_spatial_grid: %llu cells is too many; increase cell_size",
                                        (unsigned long long)cells);
  g.inv = (float)(1.0 / cell_size);
  ST_TRY(set_dev(c));
  const size_t nb1 = (size_t)cells + 1;
  float4 *dpts = nullptr, *dq = nullptr;
  uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *cnt = nullptr, *scr = nullptr, *qoff = nullptr,
           *nb = nullptr;
  uint32_t *rk = nullptr, *rv = nullptr;
  uint64_t tot = 0;
  int bits = 0;
  while (bits < 32 && (1ull << bits) < cells) bits++;
  hipError_t e = dmalloc(&dpts, n);
  if (e == hipSuccess) e = dmalloc(&dq, m);
  if (e == hipSuccess) e = dmalloc(&k0, n);
  if (e == hipSuccess) e = dmalloc(&k1, n);
  if (e == hipSuccess) e = dmalloc(&v0, n);
  if (e == hipSuccess) e = dmalloc(&v1, n);
  if (e == hipSuccess) e = dmalloc(&cnt, nb1);
  if (e == hipSuccess) e = dmalloc(&qoff, (size_t)m + 1);
  if (e == hipSuccess)
    e = dmalloc(&scr, std::max(sort_scratch_words(n), std::max(scan_scratch_words(nb1), scan_scratch_words((size_t)m + 1))));
  if (e == hipSuccess && n) e = hipMemcpy(dpts, points, (size_t)n * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess && m) e = hipMemcpy(dq, queries, (size_t)m * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, nb1 * 4, c->stream);
  if (e == hipSuccess) e = grid_cells(dpts, n, g, k0, c->stream);
  if (e == hipSuccess) e = bucket_count(k0, n, cnt, c->stream);
  if (e == hipSuccess) e = scan_exclusive(cnt, nb1, scr, c->stream);  // start (grid.js:74-75)
  rk = k0;
  rv = v0;
  if (e == hipSuccess) e = iota_u32(v0, n, c->stream);
  if (e == hipSuccess) e = radix_sort_pairs(k0, v0, k1, v1, n, 0, bits, scr, c->stream, &rk, &rv);  // order, stable
  if (e == hipSuccess) e = grid_neighbor_count(dq, m, g, cnt, qoff, c->stream);
  if (e == hipSuccess && m) e = scan_exclusive(qoff, (size_t)m + 1, scr, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && m) {
    uint32_t t32 = 0;
    e = hipMemcpy(&t32, qoff + m, 4, hipMemcpyDeviceToHost);
    tot = t32;
  }
  if (e == hipSuccess && m && neighbors && capacity >= tot && tot > 0) {
    e = dmalloc(&nb, tot);
    if (e == hipSuccess) e = grid_neighbor_fill(dq, m, g, cnt, qoff, nb, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(neighbors, nb, tot * 4, hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess && order && n) e = hipMemcpy(order, rv, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && start) e = hipMemcpy(start, cnt, nb1 * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && m) e = hipMemcpy(offsets, qoff, ((size_t)m + 1) * 4, hipMemcpyDeviceToHost);
  for (void* q : {(void*)dpts, (void*)dq, (void*)k0, (void*)k1, (void*)v0, (void*)v1, (void*)cnt, (void*)scr,
                  (void*)qoff, (void*)nb})
    if (q) hipFree(q);
  if (total) *total = tot;
  if (e == hipSuccess && m && neighbors && capacity < tot)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_spatial_grid: %llu neighbour entries exceed capacity %llu",
                (unsigned long long)tot, (unsigned long long)capacity);
  return from_hip(e, "mpm_spatial_grid");
}

}  // extern "C"
