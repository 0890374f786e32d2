// MLS-MPM substep kernels for gfx950 (P2G, grid update, G2P, sort keys, layout conversion).
//
// Device data layout
//  - Particles: structure of float4 streams. Stream k (k = 0..7) of a buffer holds vec4 k of every
//    particle record (x|mat, v|J, c0|t00, c1|t11, c2|t22, f0|t01, f1|t02, f2|t12), so a wavefront
//    reading one stream touches 64 consecutive 16-byte elements (1 KiB, fully coalesced).
//  - Grid: block-major. The domain is cut into 4x4x4-node blocks; block b = bx + NBX*(by + NBY*bz)
//    holds its 64 nodes contiguously (local = lx + 4*(ly + 4*lz)), so one block is 1 KiB of int4 (the
//    accumulation grid A, i32 fixed point) or float4 (the velocity grid V, f32 velocity and mass).
//  - Active blocks: stamp[b] holds the number of the last P2G that touched block b. The first
//    touch in a P2G appends b to that P2G's list; the grid update visits only listed blocks,
//    writes V and returns A to zero, so A is all-zero between substeps without a full clear.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

struct Geo {
  int32_t X, Y, Z;     // cells (= nodes) per axis
  int32_t nbx, nby, nbz;  // 4^3 blocks per axis
};

struct GridArgs {
  float g[3];
  float bound;
  float floor_friction;
  uint32_t ncol;
};

// Launch helpers (all enqueue on `stream`; none synchronizes).
hipError_t launch_records_in(const float4* aos, float4* P, uint32_t cap, uint32_t offset, uint32_t n, hipStream_t s);
hipError_t launch_records_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, float4* aos, hipStream_t s);
hipError_t launch_rescale(float4* P, uint32_t cap, uint32_t n, float k, hipStream_t s);
hipError_t launch_keys(const float4* P, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s);
hipError_t launch_permute(const float4* src, float4* dst, uint32_t cap, uint32_t n, const uint32_t* perm, hipStream_t s);

struct P2GArgs {
  const float4* P;
  uint32_t cap, n;
  float vol4;
  int4* A;
  uint32_t* stamp;
  uint32_t* list;
  uint32_t* list_count;
  uint32_t epoch;
};
// P2G into A. mats: 96 floats of material table in device memory. tile_cap: LDS tile capacity in
// nodes (0 = global atomics only, the reference's forceGlobal baseline). ppg: particles handled by
// one 256-thread workgroup.
hipError_t launch_p2g_dev(const P2GArgs& a, const Geo& geo, const float* mats, int tile_cap, int ppg, hipStream_t s);
// Grid update over the blocks listed by the last P2G (cols: 192 floats in device memory).
hipError_t launch_grid_update_dev(const uint32_t* list, const uint32_t* list_count, int4* A, float4* V, const Geo& geo,
                                  const GridArgs& ga, const float* cols, uint32_t max_blocks, hipStream_t s);
hipError_t launch_g2p_dev(float4* P, uint32_t cap, uint32_t n, const float4* V, const Geo& geo, const float* mats,
                          float bound, float vmax, hipStream_t s);
hipError_t launch_clear_listed(const uint32_t* list, const uint32_t* list_count, int4* A, uint32_t max_blocks,
                               hipStream_t s);
// Reference-layout export: mode 0 = A (i32 words), mode 1 = V (f32 bits, 0 outside stamped blocks).
hipError_t launch_grid_export(const int4* A, const float4* V, const uint32_t* stamp, uint32_t epoch, int mode,
                              const Geo& geo, int4* out, hipStream_t s);

}  // namespace mpm
