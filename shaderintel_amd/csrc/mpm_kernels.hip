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

// First touch of block b in this P2G appends it to the active list.
__device__ __forceinline__ void stamp_block(uint32_t b, uint32_t epoch, uint32_t* stamp, uint32_t* list,
                                            uint32_t* count) {
  if (__hip_atomic_load(&stamp[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
    if (atomicExch(&stamp[b], epoch) != epoch) list[atomicAdd(count, 1u)] = b;
  }
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

// P2G: one workgroup scatters `ppg` consecutive particles. When the AABB of their stencils fits
// tile_cap nodes, contributions are summed in an LDS tile with LDS integer atomics and only
// non-zero nodes are flushed with global atomics (mpm.js:351-400); otherwise every contribution is
// a global atomic (the reference's forceGlobal path, mpm.js:342-345).
__global__ __launch_bounds__(kThreads) void k_p2g(P2GArgs a, Geo geo, const float* __restrict__ mats, int tile_cap,
                                                  int ppg) {
  extern __shared__ int tile[];
  __shared__ int box[6];
  const int tid = threadIdx.x;
  const uint32_t first = blockIdx.x * (uint32_t)ppg;
  const uint32_t last = min(first + (uint32_t)ppg, a.n);
  if (first >= last) return;  // uniform across the workgroup

  if (tid < 3) { box[tid] = INT_MAX; box[3 + tid] = INT_MIN; }
  __syncthreads();
  int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
  for (uint32_t i = first + tid; i < last; i += kThreads) {
    const float4 x = a.P[i];
    const int b0 = (int)floorf(x.x - 0.5f), b1 = (int)floorf(x.y - 0.5f), b2 = (int)floorf(x.z - 0.5f);
    lo0 = min(lo0, b0); lo1 = min(lo1, b1); lo2 = min(lo2, b2);
    hi0 = max(hi0, b0); hi1 = max(hi1, b1); hi2 = max(hi2, b2);
  }
  lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
  hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
  if ((tid & 63) == 0) {
    atomicMin(&box[0], lo0); atomicMin(&box[1], lo1); atomicMin(&box[2], lo2);
    atomicMax(&box[3], hi0); atomicMax(&box[4], hi1); atomicMax(&box[5], hi2);
  }
  __syncthreads();
  const int ox = box[0], oy = box[1], oz = box[2];
  const int ex = box[3] - ox + 3, ey = box[4] - oy + 3, ez = box[5] - oz + 3;
  const long long nodes = (long long)ex * ey * ez;
  const bool use_tile = tile_cap > 0 && nodes <= (long long)tile_cap;

  if (use_tile) {
    // stamp every block the tile overlaps
    const int b0x = ox >> 2, b0y = oy >> 2, b0z = oz >> 2;
    const int nbxk = ((ox + ex - 1) >> 2) - b0x + 1, nbyk = ((oy + ey - 1) >> 2) - b0y + 1,
              nbzk = ((oz + ez - 1) >> 2) - b0z + 1;
    const int nbk = nbxk * nbyk * nbzk;
    for (int t = tid; t < nbk; t += kThreads) {
      const int kx = t % nbxk, ky = (t / nbxk) % nbyk, kz = t / (nbxk * nbyk);
      const uint32_t b = (uint32_t)(b0x + kx) + (uint32_t)geo.nbx * ((uint32_t)(b0y + ky) + (uint32_t)geo.nby * (uint32_t)(b0z + kz));
      stamp_block(b, a.epoch, a.stamp, a.list, a.list_count);
    }
    for (int t = tid; t < (int)nodes * 4; t += kThreads) tile[t] = 0;
    __syncthreads();
    for (uint32_t i = first + tid; i < last; i += kThreads) {
      float4 r[8];
      load_record(a.P, a.cap, i, r);
      p2g_particle(r, mats, a.vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
        const int ti = (((gz - oz) * ey + (gy - oy)) * ex + (gx - ox)) * 4;
        atomicAdd(&tile[ti], q0);
        atomicAdd(&tile[ti + 1], q1);
        atomicAdd(&tile[ti + 2], q2);
        atomicAdd(&tile[ti + 3], q3);
      });
    }
    __syncthreads();
    for (int t = tid; t < (int)nodes; t += kThreads) {
      const int q0 = tile[4 * t], q1 = tile[4 * t + 1], q2 = tile[4 * t + 2], q3 = tile[4 * t + 3];
      if ((q0 | q1 | q2 | q3) != 0) {
        const int lx = t % ex, ly = (t / ex) % ey, lz = t / (ex * ey);
        int* g = reinterpret_cast<int*>(a.A + node_addr(ox + lx, oy + ly, oz + lz, geo));
        atomicAdd(g, q0);
        atomicAdd(g + 1, q1);
        atomicAdd(g + 2, q2);
        atomicAdd(g + 3, q3);
      }
    }
  } else {
    for (uint32_t i = first + tid; i < last; i += kThreads) {
      float4 r[8];
      load_record(a.P, a.cap, i, r);
      const int bx = (int)floorf(r[0].x - 0.5f), by = (int)floorf(r[0].y - 0.5f), bz = (int)floorf(r[0].z - 0.5f);
      for (int kz = bz >> 2; kz <= (bz + 2) >> 2; kz++)
        for (int ky = by >> 2; ky <= (by + 2) >> 2; ky++)
          for (int kx = bx >> 2; kx <= (bx + 2) >> 2; kx++)
            stamp_block((uint32_t)kx + (uint32_t)geo.nbx * ((uint32_t)ky + (uint32_t)geo.nby * (uint32_t)kz), a.epoch,
                        a.stamp, a.list, a.list_count);
      p2g_particle(r, mats, a.vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
        int* g = reinterpret_cast<int*>(a.A + node_addr(gx, gy, gz, geo));
        atomicAdd(g, q0);
        atomicAdd(g + 1, q1);
        atomicAdd(g + 2, q2);
        atomicAdd(g + 3, q3);
      });
    }
  }
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

__global__ __launch_bounds__(kThreads) void k_grid_update(const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ count, int4* __restrict__ A,
                                                          float4* __restrict__ V, Geo geo, GridArgs ga,
                                                          const float* __restrict__ cols) {
  const uint32_t cnt = *count;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t j = blockIdx.x * 4 + w; j < cnt; j += gridDim.x * 4) {
    const uint32_t b = list[j];
    const int bx = (int)(b % (uint32_t)geo.nbx), by = (int)((b / (uint32_t)geo.nbx) % (uint32_t)geo.nby),
              bz = (int)(b / ((uint32_t)geo.nbx * (uint32_t)geo.nby));
    const uint32_t addr = (b << 6) | (uint32_t)lane;
    const int4 q = A[addr];
    V[addr] = grid_node(q, bx * 4 + (lane & 3), by * 4 + ((lane >> 2) & 3), bz * 4 + (lane >> 4), geo, ga, cols);
    A[addr] = make_int4(0, 0, 0, 0);
  }
}

__global__ __launch_bounds__(kThreads) void k_clear_listed(const uint32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ count, int4* __restrict__ A) {
  const uint32_t cnt = *count;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t j = blockIdx.x * 4 + w; j < cnt; j += gridDim.x * 4) A[(list[j] << 6) | (uint32_t)lane] = make_int4(0, 0, 0, 0);
}

// G2P (mpm.js:479-559): gather velocity and APIC C, advect, update F, plasticity, stress.
__global__ __launch_bounds__(kThreads) void k_g2p(float4* __restrict__ P, uint32_t cap, uint32_t n,
                                                  const float4* __restrict__ V, Geo geo,
                                                  const float* __restrict__ mats, float bound, float vmax) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
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
  P[i] = make_float4(x.x, x.y, x.z, x4.w);
  P[(size_t)cap + i] = make_float4(v.x, v.y, v.z, vw);
  P[(size_t)2 * cap + i] = make_float4(C.c[0].x, C.c[0].y, C.c[0].z, tau.c[0].x);
  P[(size_t)3 * cap + i] = make_float4(C.c[1].x, C.c[1].y, C.c[1].z, tau.c[1].y);
  P[(size_t)4 * cap + i] = make_float4(C.c[2].x, C.c[2].y, C.c[2].z, tau.c[2].z);
  P[(size_t)5 * cap + i] = make_float4(F.c[0].x, F.c[0].y, F.c[0].z, tau.c[1].x);
  P[(size_t)6 * cap + i] = make_float4(F.c[1].x, F.c[1].y, F.c[1].z, tau.c[2].x);
  P[(size_t)7 * cap + i] = make_float4(F.c[2].x, F.c[2].y, F.c[2].z, tau.c[2].y);
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

__global__ void k_grid_export(const int4* __restrict__ A, const float4* __restrict__ V, const uint32_t* __restrict__ stamp,
                              uint32_t epoch, int mode, Geo geo, int4* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  if (i >= nodes) return;
  const int gx = (int)(i % (size_t)geo.X), gy = (int)((i / (size_t)geo.X) % (size_t)geo.Y),
            gz = (int)(i / ((size_t)geo.X * geo.Y));
  const uint32_t addr = node_addr(gx, gy, gz, geo);
  if (mode == 0) {
    out[i] = A[addr];
  } else {
    if (stamp[addr >> 6] == epoch) {
      const float4 v = V[addr];
      out[i] = make_int4(__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w));
    } else {
      out[i] = make_int4(0, 0, 0, 0);
    }
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
}  // namespace mpm

// Device-pointer variants used by the runtime (mats / cols in device memory).
namespace mpm {
hipError_t launch_p2g_dev(const P2GArgs& a, const Geo& geo, const float* mats, int tile_cap, int ppg, hipStream_t s) {
  if (!a.n) return hipSuccess;
  if (ppg <= 0) ppg = 512;
  const size_t lds = tile_cap > 0 ? (size_t)tile_cap * 16 : 0;
  hipLaunchKernelGGL(k_p2g, dim3(blocks_for(a.n, (unsigned)ppg)), dim3(kThreads), lds, s, a, geo, mats, tile_cap, ppg);
  return hipGetLastError();
}
hipError_t launch_grid_update_dev(const uint32_t* list, const uint32_t* list_count, int4* A, float4* V, const Geo& geo,
                                  const GridArgs& ga, const float* cols, uint32_t max_blocks, hipStream_t s) {
  unsigned g = blocks_for(max_blocks, 4);
  if (g > 8192) g = 8192;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_grid_update, dim3(g), dim3(kThreads), 0, s, list, list_count, A, V, geo, ga, cols);
  return hipGetLastError();
}
hipError_t launch_g2p_dev(float4* P, uint32_t cap, uint32_t n, const float4* V, const Geo& geo, const float* mats,
                          float bound, float vmax, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_g2p, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, P, cap, n, V, geo, mats, bound, vmax);
  return hipGetLastError();
}
hipError_t launch_clear_listed(const uint32_t* list, const uint32_t* list_count, int4* A, uint32_t max_blocks,
                               hipStream_t s) {
  unsigned g = blocks_for(max_blocks, 4);
  if (g > 8192) g = 8192;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_clear_listed, dim3(g), dim3(kThreads), 0, s, list, list_count, A);
  return hipGetLastError();
}
hipError_t launch_grid_export(const int4* A, const float4* V, const uint32_t* stamp, uint32_t epoch, int mode,
                              const Geo& geo, int4* out, hipStream_t s) {
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  hipLaunchKernelGGL(k_grid_export, dim3(blocks_for(nodes, 256)), dim3(256), 0, s, A, V, stamp, epoch, mode, geo, out);
  return hipGetLastError();
}
}  // namespace mpm
