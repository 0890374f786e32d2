/*
 * mpm.h — C ABI of the MI355X MLS-MPM library (libmpm_hip.so).
 *
 * It replaces the WebGPU path behind ShaderIntel's `MPM` class (src/physics/mpm.js:82-680, type
 * contract types/shaderintel.d.ts:832-900) and the prefix-sum / counting-sort primitives that path
 * depends on (src/prims/scan.js:292-320, src/prims/grid.js:70-79). Each entry point cites the
 * reference method it stands in for.
 *
 * Conventions
 *  - Every call returns an mpm_status; 0 is success. The codes map one-to-one onto the reference's
 *    SIError codes (src/core/errors.js:3-11) plus out-of-memory and destroyed.
 *    mpm_last_error_message() returns the calling thread's last message.
 *  - The library owns all device memory. Host pointers are borrowed for the duration of a call.
 *  - Each simulation owns one HIP stream. Calls enqueue work on it and return; mpm_read_* and
 *    mpm_sync block until the stream has drained (the reference's async read, made synchronous).
 *  - A simulation is not thread-safe: one caller thread at a time. Different simulations may be
 *    driven from different threads.
 *  - Particle records at this boundary use the reference layout (mpm.js:54-63): 32 floats per
 *    particle = 8 vec4f (x.xyz, material id as u32 bits | v.xyz, J/Jp | c0.xyz, tau00 | c1.xyz,
 *    tau11 | c2.xyz, tau22 | f0.xyz, tau01 | f1.xyz, tau02 | f2.xyz, tau12). Inside the library the
 *    particles live in a structure-of-arrays layout; the conversion happens on the device.
 *  - Grid words at this boundary use the reference layout (mpm.js:115, 317): 4 x 32-bit words per
 *    node, node = x + X*(y + Y*z). After mpm_p2g they are i32 fixed point (2^22) momentum and
 *    mass; after mpm_step they are the f32 bits of (velocity, mass) written by the grid update.
 */
#ifndef MPM_H_
#define MPM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mpm_status;
enum {
  MPM_OK = 0,
  MPM_ERR_INVALID_ARGUMENT = 1, /* SIError 'invalid-argument' */
  MPM_ERR_UNSUPPORTED = 2,      /* SIError 'unsupported' */
  MPM_ERR_NO_DEVICE = 3,        /* SIError 'no-device' */
  MPM_ERR_ERROR = 4,            /* SIError 'error' */
  MPM_ERR_OUT_OF_MEMORY = 5,
  MPM_ERR_DESTROYED = 6
};

typedef struct mpm_ctx mpm_ctx;
typedef struct mpm_sim mpm_sim;
typedef struct mpm_pool mpm_pool;

/* Geometry and capacity of one simulation (mpm.js:89-127). */
typedef struct mpm_config {
  int32_t dims[3];        /* grid cells per axis, 8..1024 (mpm.js:95) */
  uint32_t max_particles; /* capacity (mpm.js:91) */
  int32_t ppc_axis;       /* particles per cell per axis, default 2 (mpm.js:102) */
  float bound;            /* boundary cells, 3..16; 3 in the reference (mpm.js:103) */
  uint32_t sort_every;    /* sort every this many substeps, default 4 (mpm.js:101) */
  uint32_t flags;         /* reserved, 0 */
} mpm_config;

/* ---- device context (replaces SI.init / Context: src/core/context.js:32-100) ---- */
mpm_status mpm_ctx_create(int device, mpm_ctx** out);
void mpm_ctx_destroy(mpm_ctx* ctx);
/* Device name and CU count, for reports. */
mpm_status mpm_ctx_info(mpm_ctx* ctx, char* name, size_t name_len, int32_t* num_cus);

/* ---- simulation (class MPM: src/physics/mpm.js:82-680) ---- */
/* new MPM(ctx, o) — mpm.js:89-127 */
mpm_status mpm_sim_create(mpm_ctx* ctx, const mpm_config* cfg, mpm_sim** out);
/* MPM.destroy — mpm.js:262-267. Idempotent at the host level: passing NULL is a no-op. */
void mpm_sim_destroy(mpm_sim* sim);
/* matBuf.write after _uploadMaterials — mpm.js:170-199. tbl: n rows of 12 floats (n <= 8). */
mpm_status mpm_set_materials(mpm_sim* sim, const float* tbl, uint32_t n);
/* colBuf.writeOrdered after _uploadColliders — mpm.js:276-293. tbl: n rows of 12 floats (n <= 16). */
mpm_status mpm_set_colliders(mpm_sim* sim, const float* tbl, uint32_t n);
/* P.write(data, offset*8) — mpm.js:229. n records of 32 floats, reference layout. Positions must be
 * finite and inside [0.5, dims-1.5) (the range MPM.add and G2P keep them in, mpm.js:221, 502). */
mpm_status mpm_write_particles(mpm_sim* sim, const float* rec, uint32_t offset, uint32_t n);
/* MPM.count — mpm.js:104, 230, 239. */
mpm_status mpm_set_count(mpm_sim* sim, uint32_t n);
mpm_status mpm_get_count(mpm_sim* sim, uint32_t* n);
/* MPM.stepCount — mpm.js:105, 601 (drives the sort phase). */
mpm_status mpm_set_step_count(mpm_sim* sim, uint64_t n);
mpm_status mpm_get_step_count(mpm_sim* sim, uint64_t* n);
/* MPM.sortEvery — mpm.js:101, 591: sort every n substeps from now on (n >= 1). */
mpm_status mpm_set_sort_every(mpm_sim* sim, uint32_t n);
/* MPM.step's substep loop — mpm.js:590-602. g_lattice = gravity*dt^2/dx as float (mpm.js:588);
 * the reference passes floor_friction 0.4 and vmax 0.9 (mpm.js:594, 599). */
mpm_status mpm_step(mpm_sim* sim, uint32_t substeps, const float g_lattice[3], float floor_friction, float vmax);
/* MPM.p2g(P, forceGlobal) — mpm.js:607-612: clear the grid and scatter; the grid is left in i32
 * fixed point for mpm_read_grid. force_global skips the LDS tiles (the reference's baseline). */
mpm_status mpm_p2g(mpm_sim* sim, int32_t force_global);
/* MPM.p2g(P, forceGlobal, grid) with a grid array other than the simulation's (mpm.js:607-612;
 * physics.test.js:39 compares two such grids): the P2G result in reference layout (4 i32 words per
 * node) is written to `words` (numNodes * 4). Unlike the reference, the simulation's own grid then
 * holds the same P2G result (as after mpm_p2g). Blocking. */
mpm_status mpm_p2g_to(mpm_sim* sim, int32_t force_global, int32_t* words);
/* MPM._sort — mpm.js:614-621: stable counting sort of the particles by stencil cell. */
mpm_status mpm_sort(mpm_sim* sim);
/* rescaleK — mpm.js:405-413: v and C times k (used when substeps / timeScale change). */
mpm_status mpm_rescale(mpm_sim* sim, float k);
/* P.read(n*8) — mpm.js:625. Blocking; reference layout out. */
mpm_status mpm_read_particles(mpm_sim* sim, float* rec, uint32_t offset, uint32_t n);
/* The reference's reads are Promises that resolve later (P.read / grid.read, src/core/context.js:253-270).
 * mpm_read_*_begin enqueues the snapshot on the simulation's stream now, ordered after the work
 * enqueued before it and before the work enqueued after it, into library-owned device staging
 * (stream-ordered allocation), and returns at once. mpm_read_finish, which may run on any thread
 * (it touches the read, not the simulation), waits for that snapshot, copies it to `dst` (NULL:
 * discard it) and frees the read. Every begun read must be finished. Sizes: particles n * 128 bytes
 * (reference records), grid numNodes * 16 (reference words), points n * 16 (n * 32 with velocities:
 * the points, then the velocities). */
typedef struct mpm_read mpm_read;
mpm_status mpm_read_particles_begin(mpm_sim* sim, uint32_t offset, uint32_t n, mpm_read** out);
mpm_status mpm_read_grid_begin(mpm_sim* sim, mpm_read** out);
mpm_status mpm_read_points_begin(mpm_sim* sim, uint32_t offset, uint32_t n, const double origin[3], double dx,
                                 double vscale, int32_t with_velocity, mpm_read** out);
mpm_status mpm_read_finish(mpm_read* read, void* dst);
/* grid.read() — src/verify/index.js:178. Blocking; 4 words per node, reference layout. */
mpm_status mpm_read_grid(mpm_sim* sim, int32_t* words);
/* readPositions (mpm.js:624-630) without the full records: for particles [offset, offset + n),
 * pts[4i..] = (x * dx + origin computed in f64 and rounded to f32, material id as u32 bits) — the
 * position and material slots the render layers read (stride-8 records at position 0, material in
 * x.w, mpm.js:642-679); vel (optional, may be NULL) = (v * vscale in f64 -> f32, J), e.g.
 * vscale = dx / substepDt for world units per second. 16 (or 32) bytes per particle cross PCIe
 * instead of 128. Blocking. */
mpm_status mpm_read_points(mpm_sim* sim, float* pts, float* vel, uint32_t offset, uint32_t n, const double origin[3],
                           double dx, double vscale);
/* Device pointers of the 8 particle streams S0..S7 (capacity float4 entries each), for consumers on
 * the same GPU that read them (e.g. an interop renderer; changes go through mpm_write_particles, since
 * a P2G carried from the last step call, MPM_TUNE_CARRY, does not see writes through these pointers). Float f of a reference record (mpm.js:54-63) is component
 * MPM_REC_SLOT[f] & 3 of stream MPM_REC_SLOT[f] >> 2: S0 = (x.xyz, material), S1 = (v.xyz, tau01),
 * S2..S4 = (C column k, tau_kk), S5 = (tau02, tau12, J, F22), S6 = (F00, F10, F20, F01),
 * S7 = (F11, F21, F02, F12) with Frc = row r of column c (the reference's f0..f2 are F's columns).
 * Positions and the material slot are S0 as the render layers read them (mpm.js:642-679). Valid until the next
 * call that steps, sorts or writes: a sort's reorder swaps the particle buffers, as P and P2 swap in
 * the reference (types/shaderintel.d.ts:873-874). */
#define MPM_REC_SLOT                                                                                   \
  {0, 1, 2, 3, 4, 5, 6, 22, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 24, 25, 26, 7, 27, 28, 29, 20, 30, 31, 23, 21}
mpm_status mpm_particle_streams(mpm_sim* sim, void* streams[8], uint32_t* capacity);

/* The grid between the grid update and G2P — the point where the reference calls its coupling
 * (this.coupling?.substep(s), mpm.js:596; coupling.js:203-221 reads and writes the f32 (v, m) grid).
 * Device layout of the velocities: float4 (v.xyz, mass) per node in 4x4x4-node blocks;
 * node (x, y, z) is element 64 * (block_base + x/4 + NBX * (y/4 + NBY * (z/4))) + (x%4 + 4 * (y%4 + 4 * (z%4))).
 * Only the blocks of the two device lists hold this substep's values (others are stale). */
typedef struct mpm_grid_view {
  void* velocity;               /* device float4 array */
  uint32_t block_base;          /* first block of the simulation */
  int32_t blocks[3];            /* NBX, NBY, NBZ = ceil(dims / 4) */
  int32_t dims[3];
  const uint32_t* s_blocks;     /* device list of blocks the particles reach (NULL after a global-atomic P2G) */
  const uint32_t* d_blocks;     /* device list of further blocks written by the fallback path */
  const uint32_t* counts;       /* device [2]: |d_blocks|, |s_blocks| */
  void* stream;                 /* the simulation's hipStream_t: enqueue work on it, do not synchronize */
  uint32_t substep;             /* index of the substep within the current mpm_step call */
  uint64_t step_count;
} mpm_grid_view;
typedef void (*mpm_grid_hook)(void* user, const mpm_grid_view* view);
/* Native after-grid-update hook (the coupling call site, mpm.js:596): called on the caller's thread
 * at every substep, after the grid update has been enqueued; work the hook enqueues on view->stream
 * runs before G2P. NULL removes it. While a hook is set, mpm_step runs without captured graphs. */
mpm_status mpm_set_grid_hook(mpm_sim* sim, mpm_grid_hook hook, void* user);
/* One substep split at the same point, for callers that edit the grid from the host: begin = sort
 * when due, P2G and grid update (mpm.js:591-595); then mpm_read_grid / mpm_write_grid; end = G2P
 * (mpm.js:598-601). */
mpm_status mpm_substep_begin(mpm_sim* sim, const float g_lattice[3], float floor_friction);
mpm_status mpm_substep_end(mpm_sim* sim, float vmax);
/* Writes every node's (v.xyz, mass) from f32 bits in the reference grid layout (4 words per node,
 * node = x + X*(y + Y*z)); valid after a grid update (between mpm_substep_begin and _end). */
mpm_status mpm_write_grid(mpm_sim* sim, const int32_t* words);
/* ctx.idle() — src/core/context.js:198-204. */
mpm_status mpm_sync(mpm_sim* sim);
/* ctx.region timings (src/core/profiler.js): milliseconds per phase {sort, p2g, grid, g2p} summed
 * over the substeps stepped since profiling was enabled or the times were last read, measured with
 * HIP events. enable = 1: events between the phases of every substep (all four phases); enable = 2:
 * only G2P, timed by its own dispatch (start and end recorded by the launch, no extra packets on
 * the stream, so the substep runs as it does unprofiled); 0: off. */
mpm_status mpm_set_profiling(mpm_sim* sim, int32_t enable);
mpm_status mpm_phase_times(mpm_sim* sim, float ms[4]);
/* Profiling mode 2 (no counterpart in the reference): milliseconds and launch counts of the G2P
 * launches since profiling was enabled, split into {G2P alone, G2P fused with the next substep's
 * P2G}. */
mpm_status mpm_kernel_times(mpm_sim* sim, float ms[2], uint32_t launches[2]);
/* Profiling mode 2 times only G2P launches 0, every, 2 every, ... counted from when profiling was
 * enabled (default every = 1: all of them); the launches in between run without events. No
 * counterpart in the reference. */
mpm_status mpm_set_profiling_every(mpm_sim* sim, uint32_t every);

/* Tuning knobs (no counterpart in the reference; results do not depend on them — every variant
 * computes the same integer grid sums and the same per-particle float sequence). */
#define MPM_TUNE_G2P_WAVES 2 /* occupancy target of the G2P kernel without a fused P2G (a call's last
                              * substep), 3..6 waves per SIMD (default 3) */
#define MPM_TUNE_P2G_WAVES 4 /* occupancy target of the P2G kernel, 4..6 (default 4) */
#define MPM_TUNE_G2P_ITEMS 5 /* 1: G2P walks P2G's work items with each item's velocity box in LDS
                              * (default); 0: one thread per particle, gathering from the grid */
#define MPM_TUNE_FUSE 6      /* 1: G2P fused with the next substep's P2G within one mpm_step call, before
                              * a sort too (over the current work items; default; needs
                              * MPM_TUNE_G2P_ITEMS); 0: separate launches */
#define MPM_TUNE_FUSE_SHAPE 7 /* workgroup shape of the fused G2P + P2G: 0 = 192 threads at 3 waves per
                               * SIMD, 1 = 256 threads at 4 (default) */
#define MPM_TUNE_DRIFT 8      /* P2G's tile shape: the cells a particle's stencil base may drift outside
                               * its work item's 4^3-cell block before it takes the global-atomic path:
                               * 2 (a 10^3-node LDS tile, default) or 3 (12^3; no particle slower than
                               * 1 cell per substep leaves it between sorts) */
#define MPM_TUNE_CARRY 9      /* 1 (default): a step call that follows another with only reads (or
                               * collider changes) between them ends with a G2P fused with the next
                               * substep's P2G, which the next step call takes up instead of running
                               * its own; any other call in between discards it. 0: every call opens
                               * with its own P2G and ends with a plain G2P */
mpm_status mpm_set_tuning(mpm_sim* sim, int32_t knob, int32_t value);

/* ---- scene pools: many independent simulations advanced by the same launches ----
 * No single reference counterpart: the reference runs one MPM object per scene and steps each with
 * its own dispatches (mpm.js:582-604). A pool holds many scenes (each with its own grid, gravity,
 * materials and particles) in one particle pool and one node-block pool, and every substep phase is
 * one launch over all of them; results per scene are the same as for a separate simulation of that
 * scene. Used for batched serving. Records at this boundary use the reference layout with
 * scene-local material ids (0..7). */
typedef struct mpm_pool_config {
  uint32_t max_particles;   /* particle capacity over all scenes */
  uint32_t max_scenes;      /* scene slots, 1..4096 */
  uint32_t max_node_blocks; /* 4x4x4-node blocks over all scenes; a scene needs ceil(X/4)*ceil(Y/4)*ceil(Z/4) */
  int32_t max_dim;          /* largest grid edge of any scene, 8..1024 */
  uint32_t sort_every;      /* 0: default 4 */
  uint32_t flags;           /* reserved, 0 */
} mpm_pool_config;

typedef struct mpm_scene_desc {
  int32_t dims[3];         /* grid cells per axis, 8..max_dim */
  int32_t ppc_axis;        /* particles per cell per axis (particle volume, mpm.js:102, 610) */
  float bound;             /* boundary cells, 3..16 */
  float gravity[3];        /* lattice gravity g*dt^2/dx (mpm.js:588) */
  uint32_t num_materials;  /* rows of `materials`, <= 8 */
  const float* materials;  /* 12 floats per row (mpm_material_row) */
  const float* records;    /* count records of 32 floats, reference layout */
  uint32_t count;
} mpm_scene_desc;

/* Device bytes mpm_pool_create allocates for this configuration (for admission against an HBM budget). */
size_t mpm_pool_bytes(const mpm_pool_config* cfg);
mpm_status mpm_pool_create(mpm_ctx* ctx, const mpm_pool_config* cfg, mpm_pool** out);
void mpm_pool_destroy(mpm_pool* pool);
/* Adds a scene with its particles; *scene receives its id. MPM_ERR_OUT_OF_MEMORY when particles,
 * node blocks or scene slots do not fit (nothing is changed then). */
mpm_status mpm_pool_add_scene(mpm_pool* pool, const mpm_scene_desc* desc, uint32_t* scene);
/* Removes a scene; its particles and node blocks are released at the start of the next step. */
mpm_status mpm_pool_remove_scene(mpm_pool* pool, uint32_t scene);
/* substeps substeps of every scene (MPM.step's loop, mpm.js:590-602, with each scene's own gravity). */
mpm_status mpm_pool_step(mpm_pool* pool, uint32_t substeps, float floor_friction, float vmax);
mpm_status mpm_pool_scene_count(mpm_pool* pool, uint32_t scene, uint32_t* n);
/* Blocking read of min(capacity, count) records of one scene (reference layout, scene-local
 * material ids; order is the pool's current particle order). *n receives the scene's count. */
mpm_status mpm_pool_read_scene(mpm_pool* pool, uint32_t scene, float* rec, uint32_t capacity, uint32_t* n);
mpm_status mpm_pool_info(mpm_pool* pool, uint32_t* particles, uint32_t* scenes, uint32_t* free_node_blocks);
mpm_status mpm_pool_sync(mpm_pool* pool);
/* 1: run whole sort periods (sort_every substeps) as one captured HIP graph, replayed while the
 * pool's particle count and vmax do not change. */
mpm_status mpm_pool_use_graphs(mpm_pool* pool, int32_t enable);
mpm_status mpm_pool_set_tuning(mpm_pool* pool, int32_t knob, int32_t value);
/* Diagnostics (no reference counterpart): device counters of the last binning and P2G. Blocking. */
typedef struct mpm_diagnostics {
  uint32_t work_items;          /* occupied 4x4x4-cell particle blocks of the last binning */
  uint32_t s_blocks;            /* node blocks the tiles of those items can reach */
  uint32_t d_blocks;            /* further node blocks written by the fallback path in the last P2G */
  uint32_t tile_nodes_used;     /* tile arena nodes taken in the last P2G: a slot of (6 + 2 drift)^3
                                 * per work item when they all fit, else what the tiles took (may
                                 * exceed capacity: overflow) */
  uint32_t tile_nodes_capacity;
  uint32_t particles;
} mpm_diagnostics;
mpm_status mpm_pool_diagnostics(mpm_pool* pool, mpm_diagnostics* out);
mpm_status mpm_diagnostics_of(mpm_sim* sim, mpm_diagnostics* out);

/* Momentum bookkeeping of the grid update (no counterpart in the reference: the mass and momentum
 * law checkMPM states for P2G, src/verify/index.js:175-193, carried through the grid update,
 * mpm.js:428-470). While enabled, every grid update adds, over the nodes with mass (p = the node's
 * P2G momentum words, m and v = its mass and velocity after the update, vf = p/m + g = its velocity
 * before the walls, floor and colliders act, all f32 as the update computes them), the sums below
 * as wrapping 64-bit integers, so they do not depend on the order the GPU takes them in. Per
 * substep: sum(m v) - sum(p) = sum(m g) + boundary impulse, up to f32 rounding. */
typedef struct mpm_grid_law {
  int64_t mass_words;  /* sum of P2G mass words (i32 fixed point, 2^22 per unit) */
  int64_t p2g[3];      /* sum of P2G momentum words (2^22) */
  int64_t momentum[3]; /* sum of m v after the update, each f32 product as rint(x * 2^32) */
  int64_t gravity[3];  /* sum of m g (2^32) */
  int64_t boundary[3]; /* sum of m (v - vf): the impulse of walls, floor friction and colliders (2^32) */
  uint64_t nodes;      /* node updates with mass */
  uint64_t updates;    /* grid updates summed */
} mpm_grid_law;
/* enable = 1 starts the sums from zero (and restarts them when already on); 0 stops them. */
mpm_status mpm_set_grid_law(mpm_sim* sim, int32_t enable);
/* Blocking read of the sums since they were started or last reset (reset = 1 zeroes them). */
mpm_status mpm_read_grid_law(mpm_sim* sim, mpm_grid_law* out, int32_t reset);
/* The same over all scenes of a pool. */
mpm_status mpm_pool_set_grid_law(mpm_pool* pool, int32_t enable);
mpm_status mpm_pool_read_grid_law(mpm_pool* pool, mpm_grid_law* out, int32_t reset);

/* Times the kernel shapes for trial_substeps substeps each (rounded up to whole sort periods) with HIP
 * events and keeps the fastest: first the fused G2P + P2G's workgroup shape (MPM_TUNE_FUSE_SHAPE, 2
 * trials), then P2G's tile shape (MPM_TUNE_DRIFT, 2 trials), then the occupancy of the G2P without a
 * fused P2G and of the binned P2G (12 trials). *g2p_waves and *p2g_waves receive the chosen
 * occupancies (mpm_pool_tuning reads every knob). The trials are real steps of the pool with the vmax
 * of its last step (every variant gives identical results), so the pool advances by 16 trials. */
mpm_status mpm_pool_autotune(mpm_pool* pool, uint32_t trial_substeps, int32_t* g2p_waves, int32_t* p2g_waves);
/* The current value of a tuning knob. */
mpm_status mpm_pool_tuning(mpm_pool* pool, int32_t knob, int32_t* value);

/* Scene templates: a scene kept on the device for repeated admission (serving). The records are
 * validated and copied to device memory once; mpm_pool_add_scene_from_template then admits a copy
 * with a device-to-device copy. Template memory is not part of mpm_pool_bytes. */
mpm_status mpm_pool_add_template(mpm_pool* pool, const mpm_scene_desc* desc, uint32_t* tmpl);
mpm_status mpm_pool_remove_template(mpm_pool* pool, uint32_t tmpl);
mpm_status mpm_pool_add_scene_from_template(mpm_pool* pool, uint32_t tmpl, uint32_t* scene);

/* ---- serving loop (no reference counterpart; ShaderIntel steps one MPM object per scene) ----
 * A native thread owns the pool from mpm_server_create until mpm_server_stop: the caller must not
 * call any other function on that pool meanwhile. Requests (a template and a lifetime in steps) are
 * admitted in arrival order while the pool has room (its capacity is the HBM budget, see
 * mpm_pool_bytes); a request that does not fit waits and blocks the ones behind it; one that does not
 * fit an empty pool is rejected. Each step advances every resident scene by `substeps` substeps in
 * one launch per phase; scenes join and leave between steps (continuous batching). The loop keeps one
 * step in flight: what follows step k (retirements, admissions, step k + 1) is enqueued before it
 * waits for step k. */
typedef struct mpm_server mpm_server;
typedef struct mpm_server_config {
  uint32_t substeps;    /* substeps per step, e.g. 4 */
  float floor_friction; /* 0.4 in the reference */
  float vmax;           /* 0.9 in the reference */
  uint32_t graph_after; /* 0: eager launches; n > 0: replay captured sort periods (mpm_pool_use_graphs)
                         * once the resident scenes have not changed for n steps */
} mpm_server_config;
typedef struct mpm_server_counters {
  uint64_t submitted, completed, rejected, waiting, running, steps, particle_substeps;
  uint32_t resident_particles;
  int32_t error;        /* status of a failed step (the loop stops), else 0 */
  uint64_t scene_steps; /* scene-step requests served: the sessions of every finished step */
} mpm_server_counters;
#define MPM_LAT_STEP_MS 0       /* device time of each step (HIP events around it), ms */
#define MPM_LAT_REQUEST_S 1     /* per completed request: submit -> its last step done, s */
#define MPM_LAT_WAIT_S 2        /* per completed request: submit -> admission, s */
#define MPM_LAT_REQUEST_AGE_S 3 /* per request not rejected, censored: completed ones by their latency,
                                 * waiting and running ones by their age so far (a lower bound), s */
/* Scene-step requests: each step of a session is one request, issued when the session's previous
 * step is done (its first on arrival); latency = issue -> end of the step that serves it. */
#define MPM_LAT_SCENE_STEP_S 4     /* per served scene-step request, s */
#define MPM_LAT_SCENE_STEP_AGE_S 5 /* censored: served ones by their latency, each resident and waiting
                                    * session's pending request by its age so far, s */
mpm_status mpm_server_create(mpm_pool* pool, const mpm_server_config* cfg, mpm_server** out);
mpm_status mpm_server_submit(mpm_server* srv, uint32_t tmpl, uint32_t lifetime_steps, uint64_t* ticket);
/* mpm_server_submit, and when the request's last step is done the scene's records (reference layout,
 * scene-local material ids, pool order; min(count, capacity) of them) are copied into `result`
 * before the scene leaves the pool. `result` must stay valid until mpm_server_wait reports the
 * ticket done. */
mpm_status mpm_server_submit_result(mpm_server* srv, uint32_t tmpl, uint32_t lifetime_steps, float* result,
                                    uint32_t capacity, uint64_t* ticket);
/* Waits up to timeout_ms (negative: without limit) until a ticket is done (its last step finished on
 * the device) or rejected, or the loop has stopped. *state: 0 waiting, 1 running, 2 done, 3 rejected;
 * *count (may be NULL): records copied for mpm_server_submit_result. */
mpm_status mpm_server_wait(mpm_server* srv, uint64_t ticket, int32_t timeout_ms, int32_t* state, uint32_t* count);
mpm_status mpm_server_stats(mpm_server* srv, mpm_server_counters* stats);
/* Copies up to capacity samples of one latency series; *n receives the series length. */
mpm_status mpm_server_latencies(mpm_server* srv, int32_t kind, double* out, uint64_t capacity, uint64_t* n);
/* Stops the loop (drain != 0: first finish every queued and resident request) and joins the thread. */
mpm_status mpm_server_stop(mpm_server* srv, int32_t drain);
void mpm_server_destroy(mpm_server* srv);

/* ---- primitives ---- */
/* Scan.run(src, dst, n) — src/prims/scan.js:292-306: exclusive prefix sum of n u32 (wrapping).
 * Host arrays in and out (may alias). */
mpm_status mpm_scan_u32(mpm_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n);
/* No counterpart in the reference: mpm_scan_u32 for callers that cannot allocate device memory
 * (work under HIP graph capture, tight serving loops, harnesses that forbid allocation). in and out
 * are DEVICE arrays of n u32 (out may equal in); the scan is enqueued on `stream` (a hipStream_t;
 * NULL is the null stream) and the call returns without waiting. It calls no allocator: the tile
 * partial sums live in a fixed array in each device's memory (257 KiB, held until the process
 * ends), so n is at most MPM_SCAN_NOALLOC_MAX (larger n: MPM_ERR_INVALID_ARGUMENT). Every call on
 * one device shares that array: calls that overlap on one device (other streams or threads) race,
 * and must be ordered by the caller. */
#define MPM_SCAN_NOALLOC_MAX (1ull << 28)
mpm_status mpm_scan_u32_noalloc(mpm_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, void* stream);
/* SpatialGrid.build — src/prims/grid.js:70-79: counting sort of n keys in [0, num_buckets):
 * start[k] = first slot of bucket k (num_buckets + 1 entries, start[num_buckets] = n) and
 * order[start[k] + r] = index of the r-th element (in index order) with key k. */
mpm_status mpm_counting_sort(mpm_ctx* ctx, const uint32_t* keys, uint64_t n, uint32_t num_buckets, uint32_t* order,
                             uint32_t* start);

/* SpatialGrid (src/prims/grid.js:24-117) in one call: counting sort of n points (float4, xyz used)
 * into the uniform grid over [lo, hi) with cells of cell_size (dims = max(1, ceil((hi - lo) /
 * cell_size)), cell = clamp(floor((p - lo) * f32(1 / cell_size)), 0, dims - 1), x-fastest), then
 * for each of m query points the sorted-order indices j of the 27 cells around its cell, in the
 * order SpatialGrid.neighbors visits them (z, then y, then the row's contiguous x range).
 * Outputs (host): order[n] (sorted position -> point; stable), start[numCells + 1], offsets[m + 1]
 * and neighbors[offsets[m]] (CSR). *total receives offsets[m]; if neighbors is NULL or capacity is
 * smaller, only the counts are produced (call again with room). */
mpm_status mpm_spatial_grid(mpm_ctx* ctx, const float* points, uint32_t n, const double lo[3], const double hi[3],
                            double cell_size, const float* queries, uint32_t m, uint32_t* order, uint32_t* start,
                            uint32_t* offsets, uint32_t* neighbors, uint64_t capacity, uint64_t* total);

/* ---- host helpers (no device work) ---- */
/* _uploadMaterials row for one material — mpm.js:170-196 (float64 math, float32 out). */
void mpm_material_row(int32_t kind, double density, double stiffness, double poisson, double viscosity, double hardening,
                      double compression, double stretch, double friction_deg, int32_t ppc_axis, float out12[12]);
/* sampleShape — mpm.js:683-714. shape 0 box {lo[3], hi[3]}, 1 sphere {c[3], r}, 2 cylinder
 * {c[3], r, h}. Writes min(count, capacity) float32 xyz points; returns count (-1: bad shape). */
int64_t mpm_sample_shape(int32_t shape, const double* params, double spacing, double jitter, uint32_t seed, float* out_xyz,
                         int64_t capacity);
/* MPM.add packing — mpm.js:213-227: n world-space points into n records (reference layout).
 * vel in world units per second, s = substepDt / dx. noise_sigma > 0 adds Gaussian velocity noise
 * (world units per second, Box-Muller on an xorshift32 stream seeded with noise_seed). */
void mpm_pack_particles(const float* pts, int64_t n, const double origin[3], double dx, const int32_t dims[3], double bound,
                        uint32_t material, const double vel[3], double s, double noise_sigma, uint32_t noise_seed,
                        float* rec);

/* ---- errors ---- */
const char* mpm_status_string(mpm_status s);
const char* mpm_last_error_message(void);

#ifdef __cplusplus
}
#endif
#endif /* MPM_H_ */
