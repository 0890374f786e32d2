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

/* Geometry and capacity of one simulation (mpm.js:89-127). */
typedef struct mpm_config {
  int32_t dims[3];        /* grid cells per axis, 8..1024 (mpm.js:95) */
  uint32_t max_particles; /* capacity (mpm.js:91) */
  int32_t ppc_axis;       /* particles per cell per axis, default 2 (mpm.js:102) */
  float bound;            /* boundary cells, 3..16; 3 in the reference (mpm.js:103) */
  uint32_t sort_every;    /* sort every this many substeps, default 4 (mpm.js:101) */
  uint32_t flags;         /* MPM_FLAG_* */
} mpm_config;

/* flags */
#define MPM_FLAG_NO_FUSE 1u /* run G2P and the next P2G as separate kernels (reference shape) */

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
/* MPM.step's substep loop — mpm.js:590-602. g_lattice = gravity*dt^2/dx as float (mpm.js:588);
 * the reference passes floor_friction 0.4 and vmax 0.9 (mpm.js:594, 599). */
mpm_status mpm_step(mpm_sim* sim, uint32_t substeps, const float g_lattice[3], float floor_friction, float vmax);
/* MPM.p2g(P, forceGlobal) — mpm.js:607-612: clear the grid and scatter; the grid is left in i32
 * fixed point for mpm_read_grid. force_global skips the LDS tiles (the reference's baseline). */
mpm_status mpm_p2g(mpm_sim* sim, int32_t force_global);
/* MPM._sort — mpm.js:614-621: stable counting sort of the particles by stencil cell. */
mpm_status mpm_sort(mpm_sim* sim);
/* rescaleK — mpm.js:405-413: v and C times k (used when substeps / timeScale change). */
mpm_status mpm_rescale(mpm_sim* sim, float k);
/* P.read(n*8) — mpm.js:625. Blocking; reference layout out. */
mpm_status mpm_read_particles(mpm_sim* sim, float* rec, uint32_t offset, uint32_t n);
/* grid.read() — src/verify/index.js:178. Blocking; 4 words per node, reference layout. */
mpm_status mpm_read_grid(mpm_sim* sim, int32_t* words);
/* ctx.idle() — src/core/context.js:198-204. */
mpm_status mpm_sync(mpm_sim* sim);
/* ctx.region timings (src/core/profiler.js): milliseconds of the last mpm_step per phase
 * {sort, p2g, grid, g2p}, measured with HIP events when profiling is enabled. */
mpm_status mpm_set_profiling(mpm_sim* sim, int32_t enable);
mpm_status mpm_phase_times(mpm_sim* sim, float ms[4]);

/* ---- primitives ---- */
/* Scan.run(src, dst, n) — src/prims/scan.js:292-306: exclusive prefix sum of n u32 (wrapping).
 * Host arrays in and out (may alias). */
mpm_status mpm_scan_u32(mpm_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n);
/* SpatialGrid.build — src/prims/grid.js:70-79: counting sort of n keys in [0, num_buckets):
 * start[k] = first slot of bucket k (num_buckets + 1 entries, start[num_buckets] = n) and
 * order[start[k] + r] = index of the r-th element (in index order) with key k. */
mpm_status mpm_counting_sort(mpm_ctx* ctx, const uint32_t* keys, uint64_t n, uint32_t num_buckets, uint32_t* order,
                             uint32_t* start);

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
