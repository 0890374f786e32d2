// MLS-MPM substep kernels for gfx950 (wave64, 160 KiB LDS per CU).
//
// Restates src/physics/mpm.js:306-577 (P2G, grid update, G2P, sort key, reorder) for the layouts
// described in mpm_kernels.hpp. Each float expression follows the op order fixed in mpm_math.hpp,
// so results are bit-for-bit reproducible; P2G accumulates in 32-bit fixed point with wrapping
// integer atomics, so its sums do not depend on thread order (mpm.js:64-66, 340-345).
#include <climits>

#include "mpm_kernels.hpp"
#include "mpm_math.hpp"

namespace mpm {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t node_addr(int gx, int gy, int gz, const Geo& g) {
  uint32_t b = (uint32_t)(gx >> 2) + (uint32_t)g.nbx * ((uint32_t)(gy >> 2) + (uint32_t)g.nby * (uint32_t)(gz >> 2));
  return (b << 6) | (uint32_t)((gx & 3) | ((gy & 3) << 2) | ((gz & 3) << 4));
}

__device__ __forceinline__ void load_record(const float4* __restrict__ P, uint32_t cap, uint32_t i, float4 r[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) r[k] = P[(size_t)k * cap + i];
}

// The 27 fixed-point contributions of one particle (mpm.js:318-349). add(gx, gy, gz, q0..q3).
template <class Add>
__device__ __forceinline__ void p2g_particle(const float4 r[8], const float* __restrict__ mats, float vol4, Add&& add) {
  const f3 xp = mk(r[0].x, r[0].y, r[0].z);
  const int bx = (int)floorf(xp.x - 0.5f), by = (int)floorf(xp.y - 0.5f), bz = (int)floorf(xp.z - 0.5f);
  const f3 fx = xp - mk((float)bx, (float)by, (float)bz);
  f3 w[3];
  bspline(fx, w);
  const uint32_t mid = __float_as_uint(r[0].w);
  const float mass = mats[mid * 12 + 1];
  const m3 tau = mat(mk(r[2].w, r[5].w, r[6].w), mk(r[5].w, r[3].w, r[7].w), mk(r[6].w, r[7].w, r[4].w));
  const m3 C = mat(mk(r[2].x, r[2].y, r[2].z), mk(r[3].x, r[3].y, r[3].z), mk(r[4].x, r[4].y, r[4].z));
  const m3 affine = (-vol4) * tau + mass * C;
  const f3 mv = mass * mk(r[1].x, r[1].y, r[1].z);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const f3 dpos = mk((float)a, (float)b, (float)c) - fx;
        const float wx = a == 0 ? w[0].x : (a == 1 ? w[1].x : w[2].x);
        const float wy = b == 0 ? w[0].y : (b == 1 ? w[1].y : w[2].y);
        const float wz = c == 0 ? w[0].z : (c == 1 ? w[1].z : w[2].z);
        const float wt = (wx * wy) * wz;
        const f3 mom = wt * (mv + mul(affine, dpos));
        add(bx + a, by + b, bz + c, to_fixed(mom.x), to_fixed(mom.y), to_fixed(mom.z), to_fixed(wt * mass));
      }
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
  return v;
}

__device__ __forceinline__ float collider_sdf(uint32_t kind, f3 d, const float* s) {  // mpm.js:419-426
  if (kind == 0u) return len3(d) - s[0];
  if (kind == 1u) {
    const f3 q = mk(fabsf(d.x), fabsf(d.y), fabsf(d.z)) - mk(s[0], s[1], s[2]);
    const f3 qp = mk(fmaxf(q.x, 0.0f), fmaxf(q.y, 0.0f), fmaxf(q.z, 0.0f));
    return len3(qp) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
  }
  return dot3(d, mk(s[0], s[1], s[2]));
}

// One node of the grid update (mpm.js:428-470): fixed point -> velocity, gravity, slip walls,
// Coulomb floor, SDF colliders.
__device__ __forceinline__ float4 grid_node(int4 q, int ix, int iy, int iz, const Geo& geo, const GridArgs& ga,
                                            const float* __restrict__ cols) {
  if (!(q.w > 0)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float m = (float)q.w * kInvFP;
  f3 v = mk(((float)q.x * kInvFP) / m + ga.g[0], ((float)q.y * kInvFP) / m + ga.g[1], ((float)q.z * kInvFP) / m + ga.g[2]);
  const f3 c = mk((float)ix, (float)iy, (float)iz);
  const float bound = ga.bound;
  const f3 hi = mk((float)geo.X - bound, (float)geo.Y - bound, (float)geo.Z - bound);
  if (c.y < bound && v.y < 0.0f) {
    const float vn = -v.y;
    v.y = 0.0f;
    const float vt = sqrtf(v.x * v.x + v.z * v.z);
    if (vt > 1e-9f) {
      const float f = fmaxf(0.0f, 1.0f - (ga.floor_friction * vn) / vt);
      v = mk(v.x * f, 0.0f, v.z * f);
    }
  }
  if (c.x < bound && v.x < 0.0f) v.x = 0.0f;
  if (c.z < bound && v.z < 0.0f) v.z = 0.0f;
  if (c.x > hi.x && v.x > 0.0f) v.x = 0.0f;
  if (c.y > hi.y && v.y > 0.0f) v.y = 0.0f;
  if (c.z > hi.z && v.z > 0.0f) v.z = 0.0f;
  for (uint32_t k = 0; k < ga.ncol; k++) {
    const float* h = cols + 12 * k;
    const float* s = h + 4;
    const f3 cv = mk(h[8], h[9], h[10]);
    const uint32_t kind = __float_as_uint(h[0]);
    const f3 d = c - mk(h[1], h[2], h[3]);
    const float dist = collider_sdf(kind, d, s);
    if (dist < 1.0f) {
      const float e = 0.5f;
      const f3 n = norm3(mk(collider_sdf(kind, d + mk(e, 0, 0), s) - collider_sdf(kind, d - mk(e, 0, 0), s),
                            collider_sdf(kind, d + mk(0, e, 0), s) - collider_sdf(kind, d - mk(0, e, 0), s),
                            collider_sdf(kind, d + mk(0, 0, e), s) - collider_sdf(kind, d - mk(0, 0, e), s)) +
                         mk(0.0f, 1e-7f, 0.0f));
      f3 vr = v - cv;
      const float vn = dot3(vr, n);
      if (vn < 0.0f) {
        vr = vr - vn * n;
        const float vt = len3(vr);
        if (vt > 1e-9f) vr = vr * fmaxf(0.0f, 1.0f - (s[3] * -vn) / vt);
        v = vr + cv;
      }
    }
  }
  return make_float4(v.x, v.y, v.z, m);
}

// G2P for particle i (mpm.js:479-559): gather velocity and APIC C, advect, update F, plasticity,
// stress. Writes the new record to the 8 streams and returns it in r.
__device__ __forceinline__ void g2p_particle(float4* __restrict__ P, uint32_t cap, uint32_t i,
                                             const float4* __restrict__ V, const Geo& geo,
                                             const float* __restrict__ mats, float bound, float vmax, float4 r[8]) {
  const float4 x4 = P[i];
  const float vw0 = P[(size_t)cap + i].w;
  const f3 xp = mk(x4.x, x4.y, x4.z);
  const int bx = (int)floorf(xp.x - 0.5f), by = (int)floorf(xp.y - 0.5f), bz = (int)floorf(xp.z - 0.5f);
  const f3 fx = xp - mk((float)bx, (float)by, (float)bz);
  f3 w[3];
  bspline(fx, w);
  f3 v = mk(0, 0, 0);
  m3 B = zero3();
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const f3 dpos = mk((float)a, (float)b, (float)c) - fx;
        const float wx = a == 0 ? w[0].x : (a == 1 ? w[1].x : w[2].x);
        const float wy = b == 0 ? w[0].y : (b == 1 ? w[1].y : w[2].y);
        const float wz = c == 0 ? w[0].z : (c == 1 ? w[1].z : w[2].z);
        const float wt = (wx * wy) * wz;
        const float4 g4 = V[node_addr(bx + a, by + b, bz + c, geo)];
        const f3 gv = mk(g4.x, g4.y, g4.z);
        v = v + wt * gv;
        B = B + outer(wt * gv, dpos);
      }
  const m3 C = 4.0f * B;
  const float sp = len3(v);
  if (sp > vmax) v = v * (vmax / sp);
  const float lo = bound - 1.0f;
  const f3 x = mk(clampf(xp.x + v.x, lo, ((float)geo.X - bound) + 1.0f), clampf(xp.y + v.y, lo, ((float)geo.Y - bound) + 1.0f),
                  clampf(xp.z + v.z, lo, ((float)geo.Z - bound) + 1.0f));
  const uint32_t mid = __float_as_uint(x4.w);
  const float* m0 = mats + 12 * mid;
  const uint32_t kind = __float_as_uint(m0[0]);
  m3 F, tau;
  float vw = vw0;
  if (kind == 0u) {
    const float J = clampf(vw * (1.0f + trace(C)), 0.6f, 1.4f);
    vw = J;
    tau = (m0[2] * fminf(J - 1.0f, 0.02f)) * ident() + (m0[4] * J) * (C + tr(C));
    F = ident();
  } else {
    const float4 f0 = P[(size_t)5 * cap + i], f1 = P[(size_t)6 * cap + i], f2 = P[(size_t)7 * cap + i];
    F = mul(ident() + C, mat(mk(f0.x, f0.y, f0.z), mk(f1.x, f1.y, f1.z), mk(f2.x, f2.y, f2.z)));
    const float mu0 = m0[2], la0 = m0[3];
    if (kind == 1u) {
      const float J = fmaxf(det3(F), 0.05f);
      tau = mu0 * (mul(F, tr(F)) - ident()) + (la0 * log_cw(J)) * ident();
    } else if (kind == 2u) {
      m3 U, Vm;
      f3 sv;
      svd3(F, U, sv, Vm);
      const float clo = 1.0f - m0[6], chi = 1.0f + m0[7];
      const f3 s = mk(clampf(sv.x, clo, chi), clampf(sv.y, clo, chi), clampf(sv.z, clo, chi));
      const float Jold = (sv.x * sv.y) * sv.z;
      const float Jnew = (s.x * s.y) * s.z;
      vw = clampf((vw * Jold) / Jnew, 0.6f, 20.0f);
      F = mul(mul(U, diag(s)), tr(Vm));
      const float h = exp_cw(m0[5] * (1.0f - vw));
      const float mu = mu0 * h, la = la0 * h;
      const m3 R = mul(U, tr(Vm));
      const float J = Jnew;
      tau = mul((2.0f * mu) * (F - R), tr(F)) + ((la * J) * (J - 1.0f)) * ident();
    } else {
      m3 U, Vm;
      f3 sv;
      svd3(F, U, sv, Vm);
      f3 e = mk(log_cw(fmaxf(fabsf(sv.x), 1e-4f)), log_cw(fmaxf(fabsf(sv.y), 1e-4f)), log_cw(fmaxf(fabsf(sv.z), 1e-4f)));
      const float trc = (e.x + e.y) + e.z;
      const f3 dev = e - mk(trc / 3.0f, trc / 3.0f, trc / 3.0f);
      const float devn = len3(dev);
      if (trc >= 0.0f) {
        e = mk(0, 0, 0);
      } else {
        const float dg = devn + (((3.0f * la0 + 2.0f * mu0) / (2.0f * mu0)) * trc) * m0[8];
        if (dg > 0.0f && devn > 1e-12f) e = e - (dg * dev) / devn;
      }
      const f3 s = mk(exp_cw(e.x), exp_cw(e.y), exp_cw(e.z));
      F = mul(mul(U, diag(s)), tr(Vm));
      const float te = (e.x + e.y) + e.z;
      const f3 dd = (2.0f * mu0) * e + mk(la0 * te, la0 * te, la0 * te);
      tau = mul(mul(U, diag(dd)), tr(U));
    }
  }
  r[0] = make_float4(x.x, x.y, x.z, x4.w);
  r[1] = make_float4(v.x, v.y, v.z, vw);
  r[2] = make_float4(C.c[0].x, C.c[0].y, C.c[0].z, tau.c[0].x);
  r[3] = make_float4(C.c[1].x, C.c[1].y, C.c[1].z, tau.c[1].y);
  r[4] = make_float4(C.c[2].x, C.c[2].y, C.c[2].z, tau.c[2].z);
  r[5] = make_float4(F.c[0].x, F.c[0].y, F.c[0].z, tau.c[1].x);
  r[6] = make_float4(F.c[1].x, F.c[1].y, F.c[1].z, tau.c[2].x);
  r[7] = make_float4(F.c[2].x, F.c[2].y, F.c[2].z, tau.c[2].y);
#pragma unroll
  for (int k = 0; k < 8; k++) P[(size_t)k * cap + i] = r[k];
}

__global__ __launch_bounds__(kThreads) void k_g2p(float4* __restrict__ P, uint32_t cap, uint32_t n,
                                                  const float4* __restrict__ V, Geo geo,
                                                  const float* __restrict__ mats, float bound, float vmax) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float4 r[8];
  g2p_particle(P, cap, i, V, geo, mats, bound, vmax, r);
}

__global__ void k_records_in(const float4* __restrict__ aos, float4* __restrict__ P, uint32_t cap, uint32_t offset,
                             uint32_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // one float4 per thread
  if (t >= (size_t)n * 8) return;
  const uint32_t i = (uint32_t)(t >> 3), k = (uint32_t)(t & 7);
  P[(size_t)k * cap + offset + i] = aos[t];
}

__global__ void k_records_out(const float4* __restrict__ P, uint32_t cap, uint32_t offset, uint32_t n,
                              float4* __restrict__ aos) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * 8) return;
  const uint32_t i = (uint32_t)(t >> 3), k = (uint32_t)(t & 7);
  aos[t] = P[(size_t)k * cap + offset + i];
}

__global__ void k_rescale(float4* __restrict__ P, uint32_t cap, uint32_t n, float k) {  // mpm.js:405-413
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int s = 1; s <= 4; s++) {
    float4 q = P[(size_t)s * cap + i];
    q.x = q.x * k; q.y = q.y * k; q.z = q.z * k;
    P[(size_t)s * cap + i] = q;
  }
}

// Sort key: Morton code of the stencil base cell (a refinement of the reference's 4^3-block key,
// mpm.js:567-568); vals = current index, so the radix sort's stable order breaks ties by it.
__global__ void k_keys(const float4* __restrict__ P, uint32_t n, uint32_t* __restrict__ keys,
                       uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = P[i];
  keys[i] = morton3((uint32_t)fmaxf(floorf(x.x - 0.5f), 0.0f), (uint32_t)fmaxf(floorf(x.y - 0.5f), 0.0f),
                    (uint32_t)fmaxf(floorf(x.z - 0.5f), 0.0f));
  vals[i] = i;
}

__global__ void k_permute(const float4* __restrict__ src, float4* __restrict__ dst, uint32_t cap, uint32_t n,
                          const uint32_t* __restrict__ perm) {  // mpm.js:572-577 as a gather
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t p = perm[j];
#pragma unroll
  for (int k = 0; k < 8; k++) dst[(size_t)k * cap + j] = src[(size_t)k * cap + p];
}

// ---- binning ----
__device__ __forceinline__ uint32_t compact3(uint32_t x) {  // inverse of spread3
  x &= 0x09249249u;
  x = (x ^ (x >> 2)) & 0x030c30c3u;
  x = (x ^ (x >> 4)) & 0x0300f00fu;
  x = (x ^ (x >> 8)) & 0xff0000ffu;
  x = (x ^ (x >> 16)) & 0x000003ffu;
  return x;
}

__global__ void k_bin_flags(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  flags[i] = (i < n && (i == 0 || (keys[i] >> 6) != (keys[i - 1] >> 6))) ? 1u : 0u;
}

__global__ void k_bin_emit(const uint32_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ scanned,
                           uint32_t* __restrict__ wl_start, uint32_t* __restrict__ wl_block, uint32_t* __restrict__ map,
                           uint32_t* __restrict__ item_count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t blk = keys[i] >> 6;
  if (i == 0 || (keys[i - 1] >> 6) != blk) {
    const uint32_t k = scanned[i];
    wl_start[k] = i;
    wl_block[k] = blk;
    map[blk] = k;
  }
  if (i == n - 1) {
    wl_start[scanned[n]] = n;
    *item_count = scanned[n];
  }
}


// Within-block transposition of the sorted order (run once per binning): the particles of block k,
// sorted by cell, are reordered by (rank within cell, cell). pos(r, c) = sum_c' min(n_c', r) +
// #{c' < c : n_c' > r} over the 64 cells c' of the block with counts n_c'. Writes out[start + pos] =
// in[start + j].
__global__ __launch_bounds__(kThreads) void k_bin_transpose(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ in,
                                                            const uint32_t* __restrict__ wl_start,
                                                            const uint32_t* __restrict__ item_count,
                                                            uint32_t* __restrict__ out) {
  __shared__ uint32_t ncell[64];
  __shared__ uint32_t cstart[64];
  const int tid = threadIdx.x;
  const uint32_t items = *item_count;
  for (uint32_t k = blockIdx.x; k < items; k += gridDim.x) {
    const uint32_t start = wl_start[k], cnt = wl_start[k + 1] - start;
    __syncthreads();
    if (tid < 64) ncell[tid] = 0;
    __syncthreads();
    for (uint32_t j = tid; j < cnt; j += kThreads) atomicAdd(&ncell[keys[start + j] & 63u], 1u);
    __syncthreads();
    if (tid < 64) {  // exclusive scan of the cell counts (one wave)
      const uint32_t c = ncell[tid];
      uint32_t inc = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (tid >= d) inc += o;
      }
      cstart[tid] = inc - c;
    }
    __syncthreads();
    for (uint32_t j = tid; j < cnt; j += kThreads) {
      const uint32_t c = keys[start + j] & 63u;
      const uint32_t r = j - cstart[c];
      uint32_t pos = 0;
      for (uint32_t q = 0; q < 64; q++) {
        const uint32_t nq = ncell[q];
        pos += min(nq, r) + ((q < c && nq > r) ? 1u : 0u);
      }
      out[start + pos] = in[start + j];
    }
  }
}

// First touch of node block b in this binning appends it to S.
__device__ __forceinline__ void add_to_set(uint32_t b, uint32_t epoch, uint32_t* stamp, uint32_t* list, uint32_t* count) {
  if (__hip_atomic_load(&stamp[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch)
    if (atomicExch(&stamp[b], epoch) != epoch) list[atomicAdd(count, 1u)] = b;
}

__global__ __launch_bounds__(kThreads) void k_bin_dilate(const uint32_t* __restrict__ wl_block,
                                                         const uint32_t* __restrict__ item_count, Geo geo,
                                                         uint32_t* __restrict__ sstamp, uint32_t* __restrict__ slist,
                                                         uint32_t* __restrict__ scount, uint32_t bin_epoch) {
  const uint32_t cnt = *item_count;
  // 64 threads per item: one per node block in [B-1, B+2]^3
  for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < cnt * 64; t += gridDim.x * kThreads) {
    const uint32_t k = t >> 6, d = t & 63;
    const uint32_t mb = wl_block[k];
    const int gx = (int)compact3(mb) - 1 + (int)(d & 3), gy = (int)compact3(mb >> 1) - 1 + (int)((d >> 2) & 3),
              gz = (int)compact3(mb >> 2) - 1 + (int)(d >> 4);
    if (gx < 0 || gy < 0 || gz < 0 || gx >= geo.nbx || gy >= geo.nby || gz >= geo.nbz) continue;
    add_to_set((uint32_t)gx + (uint32_t)geo.nbx * ((uint32_t)gy + (uint32_t)geo.nby * (uint32_t)gz), bin_epoch, sstamp,
               slist, scount);
  }
}

// A fallback write to node block b: mark it written in this P2G (so the grid update reads and clears
// A there) and list it in D unless S already covers it.
__device__ __forceinline__ void stamp_fallback(uint32_t b, const FallbackView& fb, const uint32_t* sstamp,
                                               uint32_t bin_epoch) {
  if (__hip_atomic_load(&fb.stamp[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != fb.epoch)
    if (atomicExch(&fb.stamp[b], fb.epoch) != fb.epoch)
      if (sstamp == nullptr || sstamp[b] != bin_epoch) fb.list[atomicAdd(fb.count, 1u)] = b;
}

__device__ __forceinline__ void fallback_particle(const float4 r[8], const float* __restrict__ mats, float vol4, int4* A,
                                                  const FallbackView& fb, const uint32_t* sstamp, uint32_t bin_epoch,
                                                  const Geo& geo) {
  const int bx = (int)floorf(r[0].x - 0.5f), by = (int)floorf(r[0].y - 0.5f), bz = (int)floorf(r[0].z - 0.5f);
  for (int kz = bz >> 2; kz <= (bz + 2) >> 2; kz++)
    for (int ky = by >> 2; ky <= (by + 2) >> 2; ky++)
      for (int kx = bx >> 2; kx <= (bx + 2) >> 2; kx++)
        stamp_fallback((uint32_t)kx + (uint32_t)geo.nbx * ((uint32_t)ky + (uint32_t)geo.nby * (uint32_t)kz), fb, sstamp,
                       bin_epoch);
  p2g_particle(r, mats, vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
    int* g = reinterpret_cast<int*>(A + node_addr(gx, gy, gz, geo));
    atomicAdd(g, q0);
    atomicAdd(g + 1, q1);
    atomicAdd(g + 2, q2);
    atomicAdd(g + 3, q3);
  });
}

// P2G over the binning's work items (mpm.js:306-401 restated for aligned particle blocks). One
// 256-thread workgroup per item at a time (persistent loop). The LDS tile covers the bounding box of
// the stencils of the item's particles whose base lies within kDrift cells of the block. Inside a
// block the sort orders particles by (rank in cell, cell), so the 64 lanes of one instruction read
// consecutive records (coalesced) that sit in different cells (their LDS atomics hit different nodes).
__global__ __launch_bounds__(kThreads) void k_p2g_binned(const float4* __restrict__ P, uint32_t cap,
                                                         const float* __restrict__ mats, float vol4, BinView bv, int4* A,
                                                         FallbackView fb, Geo geo) {
  __shared__ int tile[4 * kTileNodes];
  __shared__ int box[6];
  const int tid = threadIdx.x;
  const uint32_t items = *bv.item_count;
  for (uint32_t k = blockIdx.x; k < items; k += gridDim.x) {
    const uint32_t start = bv.wl_start[k], cnt = bv.wl_start[k + 1] - start;
    const uint32_t mb = bv.wl_block[k];
    const int wx = 4 * (int)compact3(mb) - kDrift, wy = 4 * (int)compact3(mb >> 1) - kDrift,
              wz = 4 * (int)compact3(mb >> 2) - kDrift;  // lowest in-window base
    const bool has_slot = k < bv.tile_slots;
    __syncthreads();  // the previous item's write-out has finished reading the tile
    if (tid < 3) { box[tid] = INT_MAX; box[3 + tid] = INT_MIN; }
    __syncthreads();
    int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
    for (uint32_t v = tid; v < cnt; v += kThreads) {
      const uint32_t j = start + v;
      const float4 x = P[bv.perm ? bv.perm[j] : j];
      const int b0 = (int)floorf(x.x - 0.5f), b1 = (int)floorf(x.y - 0.5f), b2 = (int)floorf(x.z - 0.5f);
      if ((unsigned)(b0 - wx) < (unsigned)(4 + 2 * kDrift) && (unsigned)(b1 - wy) < (unsigned)(4 + 2 * kDrift) &&
          (unsigned)(b2 - wz) < (unsigned)(4 + 2 * kDrift)) {
        lo0 = min(lo0, b0); lo1 = min(lo1, b1); lo2 = min(lo2, b2);
        hi0 = max(hi0, b0); hi1 = max(hi1, b1); hi2 = max(hi2, b2);
      }
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    if ((tid & 63) == 0) {
      atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
      atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
    }
    __syncthreads();
    const bool any = box[0] != INT_MAX;
    const int ox = box[0], oy = box[1], oz = box[2];
    const int ex = any ? box[3] - ox + 3 : 0, ey = any ? box[4] - oy + 3 : 0, ez = any ? box[5] - oz + 3 : 0;
    const int nodes = ex * ey * ez;  // <= kTileNodes by the window test
    if (has_slot)
      for (int t = tid; t < nodes; t += kThreads) {
        tile[t] = 0;
        tile[kTileNodes + t] = 0;
        tile[2 * kTileNodes + t] = 0;
        tile[3 * kTileNodes + t] = 0;
      }
    __syncthreads();
    for (uint32_t li = tid; li < cnt; li += kThreads) {
      const uint32_t j = start + li;
      float4 r[8];
      load_record(P, cap, bv.perm ? bv.perm[j] : j, r);
      const int b0 = (int)floorf(r[0].x - 0.5f), b1 = (int)floorf(r[0].y - 0.5f), b2 = (int)floorf(r[0].z - 0.5f);
      const bool inwin = (unsigned)(b0 - wx) < (unsigned)(4 + 2 * kDrift) &&
                         (unsigned)(b1 - wy) < (unsigned)(4 + 2 * kDrift) && (unsigned)(b2 - wz) < (unsigned)(4 + 2 * kDrift);
      if (inwin && has_slot) {
        const int tb = ((b2 - oz) * ey + (b1 - oy)) * ex + (b0 - ox);
        p2g_particle(r, mats, vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
          const int t = tb + ((gz - b2) * ey + (gy - b1)) * ex + (gx - b0);
          atomicAdd(&tile[t], q0);
          atomicAdd(&tile[kTileNodes + t], q1);
          atomicAdd(&tile[2 * kTileNodes + t], q2);
          atomicAdd(&tile[3 * kTileNodes + t], q3);
        });
      } else {
        fallback_particle(r, mats, vol4, A, fb, bv.sstamp, bv.bin_epoch, geo);
      }
    }
    __syncthreads();
    if (has_slot) {
      if (tid == 0) bv.tile_hdr[k] = make_int4(ox, oy, oz, ex | (ey << 8) | (ez << 16));
      int4* out = bv.tiles + (size_t)k * kTileNodes;
      for (int t = tid; t < nodes; t += kThreads)
        out[t] = make_int4(tile[t], tile[kTileNodes + t], tile[2 * kTileNodes + t], tile[3 * kTileNodes + t]);
    }
  }
}


// G2P of substep s fused with P2G of substep s+1 over the binning's work items. Each particle's new
// record is scattered from registers right after it is written, so P2G never re-reads it. The LDS
// tile spans the item's previous tile box grown by one cell (a particle's base moves at most one cell
// per substep while |v| <= vmax < 1), clipped to the item's window; particles outside it take the
// global-atomic fallback. The stored tile is trimmed to the box of the bases actually scattered.
__global__ __launch_bounds__(kThreads) void k_g2p_p2g(float4* __restrict__ P, uint32_t cap,
                                                      const float4* __restrict__ V, Geo geo,
                                                      const float* __restrict__ mats, float bound, float vmax,
                                                      float vol4, BinView bv, int4* A, FallbackView fb) {
  __shared__ int tile[4 * kTileNodes];
  __shared__ int box[6];
  const int tid = threadIdx.x;
  const uint32_t items = *bv.item_count;
  for (uint32_t k = blockIdx.x; k < items; k += gridDim.x) {
    const uint32_t start = bv.wl_start[k], cnt = bv.wl_start[k + 1] - start;
    const uint32_t mb = bv.wl_block[k];
    const int wx = 4 * (int)compact3(mb) - kDrift, wy = 4 * (int)compact3(mb >> 1) - kDrift,
              wz = 4 * (int)compact3(mb >> 2) - kDrift;
    const bool has_slot = k < bv.tile_slots;
    // node box D of the LDS tile: previous box +- 1, clipped to the window's nodes [w, w + kTileSpan)
    int dx0 = wx, dy0 = wy, dz0 = wz, dx1 = wx + kTileSpan, dy1 = wy + kTileSpan, dz1 = wz + kTileSpan;
    if (has_slot) {
      const int4 h = bv.tile_hdr[k];
      const int ex = h.w & 255, ey = (h.w >> 8) & 255, ez = (h.w >> 16) & 255;
      if (ex > 0) {
        dx0 = max(wx, h.x - 1); dy0 = max(wy, h.y - 1); dz0 = max(wz, h.z - 1);
        dx1 = min(wx + kTileSpan, h.x + ex + 1); dy1 = min(wy + kTileSpan, h.y + ey + 1);
        dz1 = min(wz + kTileSpan, h.z + ez + 1);
      }
    }
    const int ex = dx1 - dx0, ey = dy1 - dy0, ez = dz1 - dz0;
    const int nodes = ex * ey * ez;
    __syncthreads();  // the previous item's write-out has finished reading the tile
    if (tid < 3) { box[tid] = INT_MAX; box[3 + tid] = INT_MIN; }
    if (has_slot)
      for (int t = tid; t < nodes; t += kThreads) {
        tile[t] = 0;
        tile[kTileNodes + t] = 0;
        tile[2 * kTileNodes + t] = 0;
        tile[3 * kTileNodes + t] = 0;
      }
    __syncthreads();
    int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
    for (uint32_t li = tid; li < cnt; li += kThreads) {
      const uint32_t j = start + li;
      float4 r[8];
      g2p_particle(P, cap, bv.perm ? bv.perm[j] : j, V, geo, mats, bound, vmax, r);
      const int b0 = (int)floorf(r[0].x - 0.5f), b1 = (int)floorf(r[0].y - 0.5f), b2 = (int)floorf(r[0].z - 0.5f);
      const bool inwin = (unsigned)(b0 - wx) < (unsigned)(4 + 2 * kDrift) &&
                         (unsigned)(b1 - wy) < (unsigned)(4 + 2 * kDrift) && (unsigned)(b2 - wz) < (unsigned)(4 + 2 * kDrift);
      const bool inbox = b0 >= dx0 && b1 >= dy0 && b2 >= dz0 && b0 + 3 <= dx1 && b1 + 3 <= dy1 && b2 + 3 <= dz1;
      if (inwin && inbox && has_slot) {
        lo0 = min(lo0, b0); lo1 = min(lo1, b1); lo2 = min(lo2, b2);
        hi0 = max(hi0, b0); hi1 = max(hi1, b1); hi2 = max(hi2, b2);
        const int tb = ((b2 - dz0) * ey + (b1 - dy0)) * ex + (b0 - dx0);
        p2g_particle(r, mats, vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
          const int t = tb + ((gz - b2) * ey + (gy - b1)) * ex + (gx - b0);
          atomicAdd(&tile[t], q0);
          atomicAdd(&tile[kTileNodes + t], q1);
          atomicAdd(&tile[2 * kTileNodes + t], q2);
          atomicAdd(&tile[3 * kTileNodes + t], q3);
        });
      } else {
        fallback_particle(r, mats, vol4, A, fb, bv.sstamp, bv.bin_epoch, geo);
      }
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    if ((tid & 63) == 0) {
      atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
      atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
    }
    __syncthreads();
    if (has_slot) {
      const bool any = box[0] != INT_MAX;
      const int ox = box[0], oy = box[1], oz = box[2];
      const int fx = any ? box[3] - ox + 3 : 0, fy = any ? box[4] - oy + 3 : 0, fz = any ? box[5] - oz + 3 : 0;
      if (tid == 0) bv.tile_hdr[k] = make_int4(ox, oy, oz, fx | (fy << 8) | (fz << 16));
      int4* out = bv.tiles + (size_t)k * kTileNodes;
      for (int t = tid; t < fx * fy * fz; t += kThreads) {
        const int lx = t % fx, ly = (t / fx) % fy, lz = t / (fx * fy);
        const int s = ((oz + lz - dz0) * ey + (oy + ly - dy0)) * ex + (ox + lx - dx0);
        out[t] = make_int4(tile[s], tile[kTileNodes + s], tile[2 * kTileNodes + s], tile[3 * kTileNodes + s]);
      }
    }
  }
}

// Global-atomic P2G (the reference's forceGlobal path, mpm.js:342-345), any particle order.
__global__ __launch_bounds__(kThreads) void k_p2g_global(const float4* __restrict__ P, uint32_t cap, uint32_t n,
                                                         const float* __restrict__ mats, float vol4, int4* A,
                                                         FallbackView fb, const uint32_t* sstamp, uint32_t bin_epoch,
                                                         Geo geo) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float4 r[8];
  load_record(P, cap, i, r);
  fallback_particle(r, mats, vol4, A, fb, sstamp, bin_epoch, geo);
}

// Sum of the P2G result at node (n0, n1, n2) of node block (g0, g1, g2): the tiles of particle blocks
// G-2..G+1 per axis that cover the node. All lanes of the wave share the node block, so the map and
// header reads are uniform.
__device__ __forceinline__ int4 pull_tiles(const BinView& bv, const Geo& geo, int g0, int g1, int g2, int n0, int n1,
                                           int n2, int4 acc) {
  const int pb = 1 << geo.mbits;  // particle blocks per axis covered by the Morton code
  for (int dz = -2; dz <= 1; dz++) {
    const int bz = g2 + dz;
    if (bz < 0 || bz >= pb) continue;
    for (int dy = -2; dy <= 1; dy++) {
      const int by = g1 + dy;
      if (by < 0 || by >= pb) continue;
      for (int dx = -2; dx <= 1; dx++) {
        const int bx = g0 + dx;
        if (bx < 0 || bx >= pb) continue;
        const uint32_t k = bv.map[morton3((uint32_t)bx, (uint32_t)by, (uint32_t)bz)];
        if (k == 0xffffffffu || k >= bv.tile_slots) continue;
        const int4 h = bv.tile_hdr[k];
        const int ex = h.w & 255, ey = (h.w >> 8) & 255, ez = (h.w >> 16) & 255;
        const int lx = n0 - h.x, ly = n1 - h.y, lz = n2 - h.z;
        if ((unsigned)lx < (unsigned)ex && (unsigned)ly < (unsigned)ey && (unsigned)lz < (unsigned)ez) {
          const int4 q = bv.tiles[(size_t)k * kTileNodes + (lz * ey + ly) * ex + lx];
          acc.x = (int)((uint32_t)acc.x + (uint32_t)q.x);
          acc.y = (int)((uint32_t)acc.y + (uint32_t)q.y);
          acc.z = (int)((uint32_t)acc.z + (uint32_t)q.z);
          acc.w = (int)((uint32_t)acc.w + (uint32_t)q.w);
        }
      }
    }
  }
  return acc;
}

// Grid update (VELOCITY = true, mpm.js:428-470) or tile resolve into A (VELOCITY = false) over S then D.
template <bool VELOCITY>
__global__ __launch_bounds__(kThreads) void k_grid_pull(BinView bv, FallbackView fb, int4* __restrict__ A,
                                                        float4* __restrict__ V, Geo geo, GridArgs ga,
                                                        const float* __restrict__ cols) {
  const uint32_t ns = bv.scount ? *bv.scount : 0u, nd = *fb.count;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t j = blockIdx.x * 4 + w; j < ns + nd; j += gridDim.x * 4) {
    const uint32_t b = __builtin_amdgcn_readfirstlane(j < ns ? bv.slist[j] : fb.list[j - ns]);
    const int g0 = (int)(b % (uint32_t)geo.nbx), g1 = (int)((b / (uint32_t)geo.nbx) % (uint32_t)geo.nby),
              g2 = (int)(b / ((uint32_t)geo.nbx * (uint32_t)geo.nby));
    const int n0 = 4 * g0 + (lane & 3), n1 = 4 * g1 + ((lane >> 2) & 3), n2 = 4 * g2 + (lane >> 4);
    const uint32_t addr = (b << 6) | (uint32_t)lane;
    int4 acc = make_int4(0, 0, 0, 0);
    const bool written = fb.stamp[b] == fb.epoch;
    if (written) acc = A[addr];
    if (j < ns) acc = pull_tiles(bv, geo, g0, g1, g2, n0, n1, n2, acc);
    if (VELOCITY) {
      V[addr] = grid_node(acc, n0, n1, n2, geo, ga, cols);
      if (written) A[addr] = make_int4(0, 0, 0, 0);
    } else {
      A[addr] = acc;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_clear_blocks(const uint32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ count, int4* __restrict__ A) {
  const uint32_t cnt = *count;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t j = blockIdx.x * 4 + w; j < cnt; j += gridDim.x * 4) A[(list[j] << 6) | (uint32_t)lane] = make_int4(0, 0, 0, 0);
}

__global__ void k_grid_export(const int4* __restrict__ A, const float4* __restrict__ V, const uint32_t* __restrict__ sstamp,
                              uint32_t bin_epoch, const uint32_t* __restrict__ stamp, uint32_t epoch, int mode, Geo geo,
                              int4* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  if (i >= nodes) return;
  const int gx = (int)(i % (size_t)geo.X), gy = (int)((i / (size_t)geo.X) % (size_t)geo.Y),
            gz = (int)(i / ((size_t)geo.X * geo.Y));
  const uint32_t addr = node_addr(gx, gy, gz, geo);
  const uint32_t b = addr >> 6;
  if (mode == 0) {
    out[i] = A[addr];
  } else if ((sstamp && sstamp[b] == bin_epoch) || stamp[b] == epoch) {
    const float4 v = V[addr];
    out[i] = make_int4(__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w));
  } else {
    out[i] = make_int4(0, 0, 0, 0);
  }
}

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t launch_records_in(const float4* aos, float4* P, uint32_t cap, uint32_t offset, uint32_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_records_in, dim3(blocks_for((size_t)n * 8, 256)), dim3(256), 0, s, aos, P, cap, offset, n);
  return hipGetLastError();
}
hipError_t launch_records_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, float4* aos, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_records_out, dim3(blocks_for((size_t)n * 8, 256)), dim3(256), 0, s, P, cap, offset, n, aos);
  return hipGetLastError();
}
hipError_t launch_rescale(float4* P, uint32_t cap, uint32_t n, float k, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_rescale, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, cap, n, k);
  return hipGetLastError();
}
hipError_t launch_keys(const float4* P, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, n, keys, vals);
  return hipGetLastError();
}
hipError_t launch_permute(const float4* src, float4* dst, uint32_t cap, uint32_t n, const uint32_t* perm, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_permute, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, cap, n, perm);
  return hipGetLastError();
}

namespace {
unsigned persistent_grid(uint32_t max_work, unsigned per_block, unsigned cap) {
  unsigned g = (unsigned)((max_work + per_block - 1) / per_block);
  if (g > cap) g = cap;
  return g ? g : 1u;
}
}  // namespace

hipError_t launch_bin_flags(const uint32_t* keys, uint32_t n, uint32_t* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_flags, dim3(blocks_for((size_t)n + 1, 256)), dim3(256), 0, s, keys, n, flags);
  return hipGetLastError();
}
hipError_t launch_bin_emit(const uint32_t* keys, uint32_t n, const uint32_t* scanned, uint32_t* wl_start,
                           uint32_t* wl_block, uint32_t* map, uint32_t* item_count, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_bin_emit, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, scanned, wl_start, wl_block, map,
                     item_count);
  return hipGetLastError();
}
hipError_t launch_bin_transpose(const uint32_t* keys_sorted, const uint32_t* in, const uint32_t* wl_start,
                                const uint32_t* item_count, uint32_t* out, uint32_t max_items, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_transpose, dim3(persistent_grid(max_items, 1, 4096)), dim3(kThreads), 0, s, keys_sorted, in,
                     wl_start, item_count, out);
  return hipGetLastError();
}
hipError_t launch_bin_dilate(const uint32_t* wl_block, const uint32_t* item_count, const Geo& geo, uint32_t* sstamp,
                             uint32_t* slist, uint32_t* scount, uint32_t bin_epoch, uint32_t max_items, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_dilate, dim3(persistent_grid(max_items, 4, 4096)), dim3(kThreads), 0, s, wl_block, item_count,
                     geo, sstamp, slist, scount, bin_epoch);
  return hipGetLastError();
}
hipError_t launch_p2g_binned(const float4* P, uint32_t cap, const float* mats, float vol4, const BinView& bv, int4* A,
                             const FallbackView& fb, const Geo& geo, uint32_t max_items, hipStream_t s) {
  hipLaunchKernelGGL(k_p2g_binned, dim3(persistent_grid(max_items, 1, 2048)), dim3(kThreads), 0, s, P, cap, mats, vol4, bv,
                     A, fb, geo);
  return hipGetLastError();
}
hipError_t launch_p2g_global(const float4* P, uint32_t cap, uint32_t n, const float* mats, float vol4, int4* A,
                             const FallbackView& fb, const uint32_t* sstamp, uint32_t bin_epoch, const Geo& geo,
                             hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_p2g_global, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P, cap, n, mats, vol4, A, fb,
                     sstamp, bin_epoch, geo);
  return hipGetLastError();
}
hipError_t launch_grid_update(const BinView& bv, const FallbackView& fb, int4* A, float4* V, const Geo& geo,
                              const GridArgs& ga, const float* cols, uint32_t max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_grid_pull<true>, dim3(persistent_grid(max_blocks, 4, 8192)), dim3(kThreads), 0, s, bv, fb, A, V,
                     geo, ga, cols);
  return hipGetLastError();
}
hipError_t launch_resolve_tiles(const BinView& bv, const FallbackView& fb, int4* A, const Geo& geo, uint32_t max_blocks,
                                hipStream_t s) {
  GridArgs ga{};
  hipLaunchKernelGGL(k_grid_pull<false>, dim3(persistent_grid(max_blocks, 4, 8192)), dim3(kThreads), 0, s, bv, fb, A,
                     (float4*)nullptr, geo, ga, (const float*)nullptr);
  return hipGetLastError();
}
hipError_t launch_clear_blocks(const uint32_t* list, const uint32_t* count, int4* A, uint32_t max_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_blocks, dim3(persistent_grid(max_blocks, 4, 8192)), dim3(kThreads), 0, s, list, count, A);
  return hipGetLastError();
}
hipError_t launch_g2p_p2g(float4* P, uint32_t cap, const float4* V, const Geo& geo, const float* mats, float bound,
                          float vmax, float vol4, const BinView& bv, int4* A, const FallbackView& fb, uint32_t max_items,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_g2p_p2g, dim3(persistent_grid(max_items, 1, 2048)), dim3(kThreads), 0, s, P, cap, V, geo, mats,
                     bound, vmax, vol4, bv, A, fb);
  return hipGetLastError();
}
hipError_t launch_g2p(float4* P, uint32_t cap, uint32_t n, const float4* V, const Geo& geo, const float* mats, float bound,
                      float vmax, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_g2p, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P, cap, n, V, geo, mats, bound, vmax);
  return hipGetLastError();
}
hipError_t launch_grid_export(const int4* A, const float4* V, const uint32_t* sstamp, uint32_t bin_epoch,
                              const uint32_t* stamp, uint32_t epoch, int mode, const Geo& geo, int4* out, hipStream_t s) {
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  hipLaunchKernelGGL(k_grid_export, dim3(blocks_for(nodes, 256)), dim3(256), 0, s, A, V, sstamp, bin_epoch, stamp, epoch,
                     mode, geo, out);
  return hipGetLastError();
}

}  // namespace mpm
