// Runtime behind include/mpm.h: device context, simulation state, substep orchestration.
//
// Replaces the WebGPU Context/Kernel machinery the reference path runs on (src/core/context.js,
// src/core/kernel.js) with one HIP stream per simulation: every call enqueues, reads block. The
// substep loop is MPM.step (src/physics/mpm.js:582-604); see mpm_kernels.hpp for the device layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mpm.h"
#include "mpm_kernels.hpp"
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

// What the grid buffers hold after the last call: nothing, the i32 P2G result in A, the P2G result
// split over tiles and A (inside mpm_step only), or velocities in V.
enum GridState { kGridClean = 0, kGridA = 1, kGridTiles = 2, kGridVelocity = 3 };

constexpr size_t kTileBudgetBytes = size_t(8) << 30;  // tile slots beyond this use the fallback path

// Small tables are written by a one-thread-per-float kernel whose argument holds the data: the copy
// is ordered with the work already on the stream and needs no host buffer to outlive the call (the
// reference's writeOrdered, src/core/context.js:228-241).
struct Table192 { float v[192]; };
__global__ void k_write_table(Table192 t, float* dst, int n) {
  int i = threadIdx.x;
  if (i < n) dst[i] = t.v[i];
}

}  // namespace

struct mpm_ctx {
  int device = 0;
  char name[256] = {0};
  int cus = 0;
  hipStream_t stream = nullptr;
};

struct mpm_sim {
  mpm_ctx* ctx = nullptr;
  mpm_config cfg{};
  Geo geo{};
  uint32_t cap = 0, nblocks = 0;
  int key_bits = 0;
  uint32_t nbm = 0;         // Morton particle blocks (map entries)
  uint32_t max_items = 0;   // work items a binning can produce
  uint32_t tile_slots = 0;
  hipStream_t stream = nullptr;
  float4* P[2] = {nullptr, nullptr};
  int cur = 0;
  int4* A = nullptr;
  float4* V = nullptr;
  // D (fallback writes of the current P2G)
  uint32_t* stamp = nullptr;
  uint32_t* dlist = nullptr;
  uint32_t* dcount = nullptr;
  uint32_t epoch = 0;
  // binning
  uint32_t* wl_start = nullptr;
  uint32_t* wl_block = nullptr;
  uint32_t* item_count = nullptr;
  uint32_t* map = nullptr;
  uint32_t* flags = nullptr;
  uint32_t* slist = nullptr;
  uint32_t* scount = nullptr;
  uint32_t* sstamp = nullptr;
  int4* tiles = nullptr;
  int4* tile_hdr = nullptr;
  const uint32_t* perm = nullptr;
  uint32_t bin_epoch = 0;
  bool bin_valid = false;
  // materials / colliders
  float* mats_d = nullptr;
  float* cols_d = nullptr;
  uint32_t ncol = 0;
  // sort
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* sort_scratch = nullptr;
  // transfers
  float4* staging = nullptr;
  uint32_t staging_records = 0;
  int4* export_buf = nullptr;
  uint32_t count = 0;
  uint64_t step_count = 0;
  int grid_state = kGridClean;
  bool a_dirty_s = false, a_dirty_d = false;  // A holds data over S / over D
  uint32_t v_bin_epoch = 0, v_epoch = 0;      // binning and P2G the velocities in V belong to
  bool profiling = false;
  // Profiling: five events per profiled substep (sort, p2g, grid, g2p start; end), recorded on the
  // stream without host synchronization and resolved by mpm_phase_times.
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  float phase_ms[4] = {0, 0, 0, 0};
  uint32_t phase_substeps = 0;
};

namespace {

mpm_status set_dev(mpm_ctx* c) {
  HIP_TRY(hipSetDevice(c->device));
  return MPM_OK;
}

template <class T>
hipError_t dmalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

void free_sim(mpm_sim* s) {
  if (!s) return;
  if (s->ctx) hipSetDevice(s->ctx->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  for (hipEvent_t e : s->ev_pool) hipEventDestroy(e);
  void* ptrs[] = {s->P[0],     s->P[1],     s->A,        s->V,          s->stamp,    s->dlist,    s->dcount,
                  s->wl_start, s->wl_block, s->item_count, s->map,    s->flags,    s->slist,    s->scount,
                  s->sstamp,   s->tiles,    s->tile_hdr, s->mats_d,     s->cols_d,   s->keys[0],  s->keys[1],
                  s->vals[0],  s->vals[1],  s->sort_scratch, s->staging, s->export_buf};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

mpm_status write_table(mpm_sim* s, float* dst, const float* src, int n) {
  Table192 t;
  std::memset(&t, 0, sizeof t);
  std::memcpy(t.v, src, sizeof(float) * (size_t)n);
  hipLaunchKernelGGL(k_write_table, dim3(1), dim3(192), 0, s->stream, t, dst, 192);
  HIP_TRY(hipGetLastError());
  return MPM_OK;
}

BinView bin_view(const mpm_sim* s) {
  BinView bv;
  bv.wl_start = s->wl_start;
  bv.wl_block = s->wl_block;
  bv.item_count = s->item_count;
  bv.map = s->map;
  bv.perm = s->perm;
  bv.tiles = s->tiles;
  bv.tile_hdr = s->tile_hdr;
  bv.tile_slots = s->tile_slots;
  bv.slist = s->slist;
  bv.scount = s->scount;
  bv.sstamp = s->sstamp;
  bv.bin_epoch = s->bin_epoch;
  return bv;
}

FallbackView fallback_view(const mpm_sim* s) {
  FallbackView fb;
  fb.stamp = s->stamp;
  fb.list = s->dlist;
  fb.count = s->dcount;
  fb.epoch = s->epoch;
  return fb;
}

// Return A to all-zero where the last standalone P2G left data (before S or D change).
mpm_status clean_a(mpm_sim* s) {
  if (s->a_dirty_d) HIP_TRY(launch_clear_blocks(s->dlist, s->dcount, s->A, s->nblocks, s->stream));
  if (s->a_dirty_s) HIP_TRY(launch_clear_blocks(s->slist, s->scount, s->A, s->nblocks, s->stream));
  s->a_dirty_d = s->a_dirty_s = false;
  return MPM_OK;
}

// Sort keys, then either reorder the particles (MPM._sort, mpm.js:614-621) or only record the sorted
// order as a permutation, and rebuild the work items, the block map and S from the sorted keys.
mpm_status do_bin(mpm_sim* s, bool reorder) {
  mpm_status st = clean_a(s);
  if (st) return st;
  s->bin_epoch++;
  HIP_TRY(hipMemsetAsync(s->scount, 0, sizeof(uint32_t), s->stream));
  HIP_TRY(hipMemsetAsync(s->item_count, 0, sizeof(uint32_t), s->stream));
  s->perm = nullptr;
  if (s->count == 0) {
    s->bin_valid = true;
    return MPM_OK;
  }
  HIP_TRY(launch_keys(s->P[s->cur], s->count, s->keys[0], s->vals[0], s->stream));
  uint32_t *rk = nullptr, *rv = nullptr;
  HIP_TRY(radix_sort_pairs(s->keys[0], s->vals[0], s->keys[1], s->vals[1], s->count, s->key_bits, s->sort_scratch,
                           s->stream, &rk, &rv));
  HIP_TRY(hipMemsetAsync(s->map, 0xff, (size_t)s->nbm * sizeof(uint32_t), s->stream));
  HIP_TRY(launch_bin_flags(rk, s->count, s->flags, s->stream));
  HIP_TRY(scan_exclusive(s->flags, (size_t)s->count + 1, s->sort_scratch, s->stream));
  HIP_TRY(launch_bin_emit(rk, s->count, s->flags, s->wl_start, s->wl_block, s->map, s->item_count, s->stream));
  // final order inside each block: (rank in cell, cell); written to whichever vals buffer rv is not
  uint32_t* order = rv == s->vals[0] ? s->vals[1] : s->vals[0];
  HIP_TRY(launch_bin_transpose(rk, rv, s->wl_start, s->item_count, order, s->max_items, s->stream));
  if (reorder) {
    HIP_TRY(launch_permute(s->P[s->cur], s->P[s->cur ^ 1], s->cap, s->count, order, s->stream));
    s->cur ^= 1;
  } else {
    s->perm = order;
  }
  HIP_TRY(launch_bin_dilate(s->wl_block, s->item_count, s->geo, s->sstamp, s->slist, s->scount, s->bin_epoch,
                            s->max_items, s->stream));
  s->bin_valid = true;
  return MPM_OK;
}

float vol4_of(const mpm_sim* s) { return (float)(4.0 * (1.0 / std::pow((double)s->cfg.ppc_axis, 3))); }  // mpm.js:610

// P2G (mpm.js:607-612). Binned: into tiles (+ A for stragglers), grid state kGridTiles. Global: every
// contribution a global atomic on A, grid state kGridA.
mpm_status do_p2g(mpm_sim* s, bool force_global) {
  mpm_status st = clean_a(s);
  if (st) return st;
  if (!force_global && !s->bin_valid && (st = do_bin(s, false))) return st;
  s->epoch++;
  HIP_TRY(hipMemsetAsync(s->dcount, 0, sizeof(uint32_t), s->stream));
  const FallbackView fb = fallback_view(s);
  if (force_global) {
    HIP_TRY(launch_p2g_global(s->P[s->cur], s->cap, s->count, s->mats_d, vol4_of(s), s->A, fb, nullptr, 0, s->geo,
                              s->stream));
    s->grid_state = kGridA;
  } else {
    HIP_TRY(launch_p2g_binned(s->P[s->cur], s->cap, s->mats_d, vol4_of(s), bin_view(s), s->A, fb, s->geo, s->max_items,
                              s->stream));
    s->grid_state = kGridTiles;
  }
  s->a_dirty_d = true;
  return MPM_OK;
}

// Grid update (mpm.js:593-595) from tiles + A into V; A returns to zero.
mpm_status do_grid(mpm_sim* s, const float g[3], float ff) {
  GridArgs ga;
  ga.g[0] = g[0]; ga.g[1] = g[1]; ga.g[2] = g[2];
  ga.bound = s->cfg.bound;
  ga.floor_friction = ff;
  ga.ncol = s->ncol;
  BinView bv = bin_view(s);
  if (s->grid_state == kGridA) bv.scount = nullptr;  // global-atomic P2G: only D holds data
  HIP_TRY(launch_grid_update(bv, fallback_view(s), s->A, s->V, s->geo, ga, s->cols_d, 2 * s->nblocks, s->stream));
  s->a_dirty_d = false;
  s->grid_state = kGridVelocity;
  s->v_bin_epoch = bv.scount ? s->bin_epoch : 0xffffffffu;
  s->v_epoch = s->epoch;
  return MPM_OK;
}

// G2P of this substep fused with the next substep's P2G (grid state kGridTiles for the next grid
// update). Valid only when the last P2G used the current binning (its tile boxes seed the new ones).
mpm_status do_g2p_p2g(mpm_sim* s, float vmax) {
  s->epoch++;
  HIP_TRY(hipMemsetAsync(s->dcount, 0, sizeof(uint32_t), s->stream));
  HIP_TRY(launch_g2p_p2g(s->P[s->cur], s->cap, s->V, s->geo, s->mats_d, s->cfg.bound, vmax, vol4_of(s), bin_view(s),
                         s->A, fallback_view(s), s->max_items, s->stream));
  s->grid_state = kGridTiles;
  s->a_dirty_d = true;
  return MPM_OK;
}

mpm_status do_g2p(mpm_sim* s, float vmax) {  // mpm.js:598-600
  HIP_TRY(launch_g2p(s->P[s->cur], s->cap, s->count, s->V, s->geo, s->mats_d, s->cfg.bound, vmax, s->stream));
  return MPM_OK;
}

#define CHECK_SIM(s)                                                            \
  do {                                                                          \
    if (!(s)) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");      \
    mpm_status _st = set_dev((s)->ctx);                                         \
    if (_st != MPM_OK) return _st;                                              \
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
  mpm_ctx* c = new (std::nothrow) mpm_ctx;
  if (!c) return fail(MPM_ERR_OUT_OF_MEMORY, "mpm_ctx_create: host allocation");
  c->device = device;
  std::snprintf(c->name, sizeof c->name, "%s (%s)", p.name, p.gcnArchName);
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
  if (!c || !cfg || !out) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: NULL argument");
  *out = nullptr;
  for (int a = 0; a < 3; a++)
    if (cfg->dims[a] < 8 || cfg->dims[a] > 1024)
      return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: dims[%d] = %d outside [8, 1024]", a, cfg->dims[a]);
  if (cfg->max_particles == 0 || cfg->max_particles > (1u << 30))
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: max_particles = %u outside [1, 2^30]", cfg->max_particles);
  if (cfg->ppc_axis < 1 || cfg->ppc_axis > 8)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: ppc_axis = %d outside [1, 8]", cfg->ppc_axis);
  // bound >= 3 keeps every stencil inside the grid: G2P clamps x to [bound-1, dims-bound+1]
  // (mpm.js:502), so the base cell floor(x - 1/2) stays in [1, dims-3].
  if (!(cfg->bound >= 3.0f && cfg->bound <= 16.0f))
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: bound = %g outside [3, 16]", (double)cfg->bound);
  if (cfg->sort_every == 0) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_sim_create: sort_every = 0");
  mpm_status st = set_dev(c);
  if (st != MPM_OK) return st;
  mpm_sim* s = new (std::nothrow) mpm_sim;
  if (!s) return fail(MPM_ERR_OUT_OF_MEMORY, "mpm_sim_create: host allocation");
  s->ctx = c;
  s->cfg = *cfg;
  s->cap = cfg->max_particles;
  s->geo.X = cfg->dims[0]; s->geo.Y = cfg->dims[1]; s->geo.Z = cfg->dims[2];
  s->geo.nbx = (cfg->dims[0] + 3) / 4; s->geo.nby = (cfg->dims[1] + 3) / 4; s->geo.nbz = (cfg->dims[2] + 3) / 4;
  s->nblocks = (uint32_t)s->geo.nbx * s->geo.nby * s->geo.nbz;
  int maxd = std::max(cfg->dims[0], std::max(cfg->dims[1], cfg->dims[2]));
  int bits = 0;
  while ((1 << bits) < maxd) bits++;
  s->key_bits = 3 * bits;
  s->geo.mbits = std::max(1, bits - 2);
  s->nbm = 1u << (3 * s->geo.mbits);
  s->max_items = std::min<uint32_t>(s->nbm, s->cap);
  s->tile_slots = (uint32_t)std::min<size_t>(s->max_items, kTileBudgetBytes / (sizeof(int4) * kTileNodes));
  const size_t nodes = (size_t)s->nblocks * 64;
  hipError_t e = hipStreamCreate(&s->stream);
  if (e == hipSuccess) e = dmalloc(&s->P[0], (size_t)s->cap * 8);
  if (e == hipSuccess) e = dmalloc(&s->P[1], (size_t)s->cap * 8);
  if (e == hipSuccess) e = dmalloc(&s->A, nodes);
  if (e == hipSuccess) e = dmalloc(&s->V, nodes);
  if (e == hipSuccess) e = dmalloc(&s->stamp, s->nblocks);
  if (e == hipSuccess) e = dmalloc(&s->dlist, s->nblocks);
  if (e == hipSuccess) e = dmalloc(&s->dcount, 1);
  if (e == hipSuccess) e = dmalloc(&s->wl_start, (size_t)s->max_items + 1);
  if (e == hipSuccess) e = dmalloc(&s->wl_block, s->max_items);
  if (e == hipSuccess) e = dmalloc(&s->item_count, 1);
  if (e == hipSuccess) e = dmalloc(&s->map, s->nbm);
  if (e == hipSuccess) e = dmalloc(&s->flags, (size_t)s->cap + 1);
  if (e == hipSuccess) e = dmalloc(&s->slist, s->nblocks);
  if (e == hipSuccess) e = dmalloc(&s->scount, 1);
  if (e == hipSuccess) e = dmalloc(&s->sstamp, s->nblocks);
  if (e == hipSuccess) e = dmalloc(&s->tiles, (size_t)s->tile_slots * kTileNodes);
  if (e == hipSuccess) e = dmalloc(&s->tile_hdr, s->tile_slots);
  if (e == hipSuccess) e = dmalloc(&s->mats_d, 192);
  if (e == hipSuccess) e = dmalloc(&s->cols_d, 192);
  if (e == hipSuccess) e = dmalloc(&s->keys[0], s->cap);
  if (e == hipSuccess) e = dmalloc(&s->keys[1], s->cap);
  if (e == hipSuccess) e = dmalloc(&s->vals[0], s->cap);
  if (e == hipSuccess) e = dmalloc(&s->vals[1], s->cap);
  if (e == hipSuccess)
    e = dmalloc(&s->sort_scratch, std::max(sort_scratch_words(s->cap), scan_scratch_words((size_t)s->cap + 1)));
  if (e == hipSuccess) {
    s->staging_records = std::min<uint32_t>(s->cap, 1u << 20);
    e = dmalloc(&s->staging, (size_t)s->staging_records * 8);
  }
  if (e == hipSuccess) e = hipMemsetAsync(s->A, 0, nodes * sizeof(int4), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->V, 0, nodes * sizeof(float4), s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->stamp, 0xff, (size_t)s->nblocks * 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->sstamp, 0xff, (size_t)s->nblocks * 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->dcount, 0, 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->scount, 0, 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->item_count, 0, 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->mats_d, 0, 192 * 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->cols_d, 0, 192 * 4, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->P[0], 0, (size_t)s->cap * 128, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    mpm_status code = from_hip(e, "mpm_sim_create");
    free_sim(s);
    return code;
  }
  *out = s;
  return MPM_OK;
}

void mpm_sim_destroy(mpm_sim* s) { free_sim(s); }

mpm_status mpm_set_materials(mpm_sim* s, const float* tbl, uint32_t n) {
  CHECK_SIM(s);
  if (n > 8) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_materials: at most 8 materials (got %u)", n);
  if (n && !tbl) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_materials: table is NULL");
  for (uint32_t k = 0; k < n; k++) {
    uint32_t kind;
    std::memcpy(&kind, tbl + 12 * k, 4);
    if (kind > 3) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_materials: row %u has kind %u (0-3)", k, kind);
  }
  float t[192] = {0};
  if (n) std::memcpy(t, tbl, 48 * (size_t)n);
  return write_table(s, s->mats_d, t, 96);
}

mpm_status mpm_set_colliders(mpm_sim* s, const float* tbl, uint32_t n) {
  CHECK_SIM(s);
  if (n > 16) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_colliders: at most 16 colliders (got %u)", n);
  if (n && !tbl) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_colliders: table is NULL");
  float t[192] = {0};
  if (n) std::memcpy(t, tbl, 48 * (size_t)n);
  s->ncol = n;
  return write_table(s, s->cols_d, t, 192);
}

mpm_status mpm_write_particles(mpm_sim* s, const float* rec, uint32_t offset, uint32_t n) {
  CHECK_SIM(s);
  if (n == 0) return MPM_OK;
  if (!rec) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: records NULL");
  if ((uint64_t)offset + n > s->cap)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: offset %u + n %u exceeds capacity %u", offset, n, s->cap);
  const int dims[3] = {s->geo.X, s->geo.Y, s->geo.Z};
  for (uint32_t i = 0; i < n; i++) {
    const float* r = rec + 32 * (size_t)i;
    for (int a = 0; a < 3; a++)
      if (!(r[a] >= 0.5f && r[a] < (float)dims[a] - 1.5f))
        return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: record %u position[%d] = %g outside [0.5, %d)",
                    offset + i, a, (double)r[a], dims[a] - 1);
    uint32_t mid;
    std::memcpy(&mid, r + 3, 4);
    if (mid >= 8) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_write_particles: record %u material %u >= 8", offset + i, mid);
  }
  for (uint32_t done = 0; done < n;) {
    const uint32_t k = std::min(n - done, s->staging_records);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->staging, rec + 32 * (size_t)done, (size_t)k * 128, hipMemcpyHostToDevice));
    HIP_TRY(launch_records_in(s->staging, s->P[s->cur], s->cap, offset + done, k, s->stream));
    done += k;
  }
  s->bin_valid = false;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

mpm_status mpm_read_particles(mpm_sim* s, float* rec, uint32_t offset, uint32_t n) {
  CHECK_SIM(s);
  if (n == 0) return MPM_OK;
  if (!rec) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_particles: records NULL");
  if ((uint64_t)offset + n > s->cap)
    return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_particles: offset %u + n %u exceeds capacity %u", offset, n, s->cap);
  for (uint32_t done = 0; done < n;) {
    const uint32_t k = std::min(n - done, s->staging_records);
    HIP_TRY(launch_records_out(s->P[s->cur], s->cap, offset + done, k, s->staging, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(rec + 32 * (size_t)done, s->staging, (size_t)k * 128, hipMemcpyDeviceToHost));
    done += k;
  }
  return MPM_OK;
}

mpm_status mpm_set_count(mpm_sim* s, uint32_t n) {
  CHECK_SIM(s);
  if (n > s->cap) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_set_count: %u exceeds capacity %u", n, s->cap);
  if (n != s->count) s->bin_valid = false;
  s->count = n;
  return MPM_OK;
}

mpm_status mpm_get_count(mpm_sim* s, uint32_t* n) {
  if (!s || !n) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_get_count: NULL argument");
  *n = s->count;
  return MPM_OK;
}

mpm_status mpm_set_step_count(mpm_sim* s, uint64_t n) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  s->step_count = n;
  return MPM_OK;
}

mpm_status mpm_get_step_count(mpm_sim* s, uint64_t* n) {
  if (!s || !n) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_get_step_count: NULL argument");
  *n = s->step_count;
  return MPM_OK;
}

}  // extern "C"

namespace {
mpm_status next_event(mpm_sim* s, hipEvent_t* out) {
  if (s->ev_used == s->ev_pool.size()) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    s->ev_pool.push_back(e);
  }
  *out = s->ev_pool[s->ev_used++];
  return MPM_OK;
}
mpm_status mark(mpm_sim* s) {
  if (!s->profiling) return MPM_OK;
  hipEvent_t e;
  mpm_status st = next_event(s, &e);
  if (st) return st;
  HIP_TRY(hipEventRecord(e, s->stream));
  return MPM_OK;
}
}  // namespace

extern "C" {

mpm_status mpm_step(mpm_sim* s, uint32_t substeps, const float g[3], float ff, float vmax) {
  CHECK_SIM(s);
  if (!g) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_step: gravity is NULL");
  if (s->count == 0) return MPM_OK;  // mpm.js:584
  const bool fuse = !(s->cfg.flags & MPM_FLAG_NO_FUSE);
  bool p2g_ready = false;  // this substep's P2G was already done by the previous fused kernel
  for (uint32_t k = 0; k < substeps; k++) {
    mpm_status st;
    if ((st = mark(s))) return st;
    if (!p2g_ready) {
      if (s->step_count % s->cfg.sort_every == 0) {
        if ((st = do_bin(s, true))) return st;
      } else if (!s->bin_valid && (st = do_bin(s, false))) {
        return st;
      }
    }
    if ((st = mark(s))) return st;
    if (!p2g_ready && (st = do_p2g(s, false))) return st;
    if ((st = mark(s))) return st;
    if ((st = do_grid(s, g, ff))) return st;
    if ((st = mark(s))) return st;
    // fuse this G2P with the next P2G unless the call ends here or the next substep sorts
    p2g_ready = fuse && k + 1 < substeps && (s->step_count + 1) % s->cfg.sort_every != 0;
    if ((st = p2g_ready ? do_g2p_p2g(s, vmax) : do_g2p(s, vmax))) return st;
    if ((st = mark(s))) return st;
    s->step_count++;
  }
  return MPM_OK;
}

mpm_status mpm_p2g(mpm_sim* s, int32_t force_global) {
  CHECK_SIM(s);
  mpm_status st = do_p2g(s, force_global != 0);
  if (st) return st;
  if (s->grid_state == kGridTiles) {  // fold the tiles into A so the grid reads as the i32 P2G result
    HIP_TRY(launch_resolve_tiles(bin_view(s), fallback_view(s), s->A, s->geo, 2 * s->nblocks, s->stream));
    s->grid_state = kGridA;
    s->a_dirty_s = true;
  }
  return MPM_OK;
}

mpm_status mpm_sort(mpm_sim* s) {
  CHECK_SIM(s);
  return do_bin(s, true);
}

mpm_status mpm_rescale(mpm_sim* s, float k) {
  CHECK_SIM(s);
  HIP_TRY(launch_rescale(s->P[s->cur], s->cap, s->count, k, s->stream));
  return MPM_OK;
}

mpm_status mpm_read_grid(mpm_sim* s, int32_t* words) {
  CHECK_SIM(s);
  if (!words) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_read_grid: words NULL");
  const size_t nodes = (size_t)s->geo.X * s->geo.Y * s->geo.Z;
  if (!s->export_buf) HIP_TRY(dmalloc(&s->export_buf, nodes));
  if (s->grid_state != kGridA && s->grid_state != kGridVelocity) {
    std::memset(words, 0, nodes * 16);
    return MPM_OK;
  }
  HIP_TRY(launch_grid_export(s->A, s->V, s->sstamp, s->v_bin_epoch, s->stamp, s->v_epoch,
                             s->grid_state == kGridA ? 0 : 1, s->geo, s->export_buf, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(words, s->export_buf, nodes * 16, hipMemcpyDeviceToHost));
  return MPM_OK;
}

mpm_status mpm_sync(mpm_sim* s) {
  CHECK_SIM(s);
  HIP_TRY(hipStreamSynchronize(s->stream));
  return MPM_OK;
}

mpm_status mpm_set_profiling(mpm_sim* s, int32_t enable) {
  if (!s) return fail(MPM_ERR_INVALID_ARGUMENT, "simulation is NULL");
  s->profiling = enable != 0;
  s->ev_used = 0;
  for (float& m : s->phase_ms) m = 0.0f;
  s->phase_substeps = 0;
  return MPM_OK;
}

mpm_status mpm_phase_times(mpm_sim* s, float ms[4]) {
  if (!s || !ms) return fail(MPM_ERR_INVALID_ARGUMENT, "mpm_phase_times: NULL argument");
  mpm_status st = set_dev(s->ctx);
  if (st) return st;
  const size_t groups = s->ev_used / 5;
  if (groups) HIP_TRY(hipEventSynchronize(s->ev_pool[groups * 5 - 1]));
  for (size_t g = 0; g < groups; g++)
    for (int p = 0; p < 4; p++) {
      float t = 0;
      HIP_TRY(hipEventElapsedTime(&t, s->ev_pool[g * 5 + p], s->ev_pool[g * 5 + p + 1]));
      s->phase_ms[p] += t;
    }
  s->phase_substeps += (uint32_t)groups;
  s->ev_used = 0;
  for (int p = 0; p < 4; p++) ms[p] = s->phase_ms[p];
  return MPM_OK;
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
  if (e == hipSuccess) e = radix_sort_pairs(k0, v0, k1, v1, n, bits, scr, c->stream, &rk, &rv);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && n) e = hipMemcpy(order, rv, n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(start, cnt, nb1 * 4, hipMemcpyDeviceToHost);
  for (uint32_t* p : {k0, k1, v0, v1, cnt, scr})
    if (p) hipFree(p);
  return from_hip(e, "mpm_counting_sort");
}

}  // extern "C"
