// MLS-MPM substep kernels for gfx950 (wave64, 160 KiB LDS per CU).
//
// Restates src/physics/mpm.js:306-577 (P2G, grid update, G2P, sort key, reorder) for the pooled
// layouts described in mpm_kernels.hpp. Each float expression follows the op order fixed in
// mpm_math.hpp, so results are bit-for-bit reproducible; P2G accumulates in 32-bit fixed point with
// wrapping integer additions, so its sums do not depend on thread order (mpm.js:64-66, 340-345).
#include <climits>

#include <hip/hip_ext.h>

#include "item_sched.hpp"
#include "mpm_kernels.hpp"
#include "mpm_math.hpp"

namespace mpm {

namespace {

constexpr int kThreads = 256;
// Shapes of the fused G2P + P2G (28 KB of LDS per workgroup with the drift-2 window, 39 KB with
// drift 3): 256 threads at the 4-waves-per-SIMD register budget (128 VGPRs; every SIMD holds 4
// waves, 4 workgroups per CU), or 192 threads at 3 waves per SIMD (room for 168 VGPRs).
constexpr int kFuseShapes = 2;
constexpr int kFuseThreadsOf[kFuseShapes] = {192, 256};
constexpr int kFuseWavesOf[kFuseShapes] = {3, 4};

__device__ __forceinline__ uint32_t node_addr(int gx, int gy, int gz, const SceneGeo& g) {
  const uint32_t b = g.block_base + (uint32_t)(gx >> 2) +
                     (uint32_t)g.nbx * ((uint32_t)(gy >> 2) + (uint32_t)g.nby * (uint32_t)(gz >> 2));
  return (b << 6) | (uint32_t)((gx & 3) | ((gy & 3) << 2) | ((gz & 3) << 4));
}

static_assert(kMatRows == 8, "row >> 3 is the scene of a material row");

// Wave-uniform read of a scene's geometry through the constant address space, so it is fetched with
// scalar loads into scalar registers (the table is never written by a kernel).
typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
__device__ __forceinline__ SceneGeo uniform_geo(const Pool& pool, uint32_t scene) {
  static_assert(sizeof(SceneGeo) % 4 == 0, "SceneGeo is whole words");
  const ConstWords src = (ConstWords)(pool.scenes + scene);
  SceneGeo g;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&g);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(SceneGeo) / 4); k++) dst[k] = src[k];
  return g;
}

__device__ __forceinline__ const SceneGeo& scene_of_row(const Pool& pool, uint32_t row) { return pool.scenes[row >> 3]; }

// The streams P2G reads: S0-S5 (x, row, v, C, tau; S5 also carries J and F22, which P2G ignores).
constexpr int kP2GStreams = 6;
__device__ __forceinline__ void load_p2g_record(const float4* __restrict__ P, uint32_t cap, uint32_t i,
                                                float4 r[kP2GStreams]) {
#pragma unroll
  for (int k = 0; k < kP2GStreams; k++) r[k] = P[(size_t)k * cap + i];
}

// Sort key of a particle at x (x.xyz in its scene's lattice): the scene in the top bits, then the
// Morton code of the stencil base cell (a refinement of the reference's 4^3-block key,
// mpm.js:567-568); particles of an inactive scene get 0xffffffff (sorted behind every live one).
__device__ __forceinline__ uint32_t sort_key(const Pool& pool, uint32_t scene, uint32_t active, float4 x) {
  if (!active) return 0xffffffffu;
  return (scene << pool.key_bits) |
         morton3((uint32_t)fmaxf(floorf(x.x - 0.5f), 0.0f), (uint32_t)fmaxf(floorf(x.y - 0.5f), 0.0f),
                 (uint32_t)fmaxf(floorf(x.z - 0.5f), 0.0f));
}

// toFixed's clamp to +-2e9 (mpm.js:64-66) can only act on a particle whose per-node quantity
// q = mv + affine*dpos exceeds ~1130 in some component (wts <= 0.421875 * 2^22 after rounding) or
// whose mass does. p2g_unclamped(r) bounds |q| over all 27 nodes from the record (|dpos| <= 1.5 per
// axis) and is true when the bound stays under 1024, with every comparison false on NaN or inf; for
// such a particle rint(clamp(x)) == rint(x) for every contribution, so the clamp can be skipped.
__device__ __forceinline__ bool p2g_unclamped(const float4 r[kP2GStreams], const float* __restrict__ mats, float vol4) {
  const float mass = fabsf(mats[12 * __float_as_uint(r[0].w) + 1]);
  const float t00 = fabsf(r[2].w), t11 = fabsf(r[3].w), t22 = fabsf(r[4].w);
  const float t01 = fabsf(r[1].w), t02 = fabsf(r[5].x), t12 = fabsf(r[5].y);
  const float vs = fabsf(vol4);
  // row i of |affine| <= vol4 |tau row i| + mass |C row i| (C row i = component i of the columns)
  const float a0 = ((t00 + t01) + t02) * vs + ((fabsf(r[2].x) + fabsf(r[3].x)) + fabsf(r[4].x)) * mass;
  const float a1 = ((t01 + t11) + t12) * vs + ((fabsf(r[2].y) + fabsf(r[3].y)) + fabsf(r[4].y)) * mass;
  const float a2 = ((t02 + t12) + t22) * vs + ((fabsf(r[2].z) + fabsf(r[3].z)) + fabsf(r[4].z)) * mass;
  return mass <= 1024.0f && fabsf(r[1].x) * mass + 1.5f * a0 <= 1024.0f && fabsf(r[1].y) * mass + 1.5f * a1 <= 1024.0f &&
         fabsf(r[1].z) * mass + 1.5f * a2 <= 1024.0f;
}

// The 27 fixed-point contributions of one particle (mpm.js:318-349). add(gx, gy, gz, q0..q3).
// CLAMP = false skips toFixed's clamp: only for particles p2g_unclamped accepts.
template <bool CLAMP = true, class Add>
__device__ __forceinline__ void p2g_particle(const float4 r[kP2GStreams], const float* __restrict__ mats, float vol4,
                                             Add&& add) {
  const f3 xp = mk(r[0].x, r[0].y, r[0].z);
  const int bx = (int)floorf(xp.x - 0.5f), by = (int)floorf(xp.y - 0.5f), bz = (int)floorf(xp.z - 0.5f);
  const f3 fx = xp - mk((float)bx, (float)by, (float)bz);
  f3 w[3];
  bspline(fx, w);
  const float mass = mats[12 * __float_as_uint(r[0].w) + 1];
  const m3 tau = mat(mk(r[2].w, r[1].w, r[5].x), mk(r[1].w, r[3].w, r[5].y), mk(r[5].x, r[5].y, r[4].w));
  const m3 C = mat(mk(r[2].x, r[2].y, r[2].z), mk(r[3].x, r[3].y, r[3].z), mk(r[4].x, r[4].y, r[4].z));
  const m3 affine = (-vol4) * tau + mass * C;
  const f3 mv = mass * mk(r[1].x, r[1].y, r[1].z);
  // q(a, b, c) = mv + affine * dpos with dpos = vec3f(a, b, c) - fx (mpm.js:330-334), the mat*vec
  // evaluated as fma(A2, dz_c, fma(A1, dy_b, A0 * dx_a)) (mul's order) with its partial products
  // shared between nodes: A0 * dx_a once per a, the first fma once per (a, b). Node by node the
  // same operations on the same operands as the per-node expression.
  const f3 A0 = affine.c[0], A1 = affine.c[1], A2 = affine.c[2];
  const float dx[3] = {0.0f - fx.x, 1.0f - fx.x, 2.0f - fx.x};
  const float dy[3] = {0.0f - fx.y, 1.0f - fx.y, 2.0f - fx.y};
  const float dz[3] = {0.0f - fx.z, 1.0f - fx.z, 2.0f - fx.z};
  const float wzs[3] = {w[0].z * kFP, w[1].z * kFP, w[2].z * kFP};  // 2^22 folded in (exact)
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const f3 ta = mk(A0.x * dx[a], A0.y * dx[a], A0.z * dx[a]);
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const float wx = a == 0 ? w[0].x : (a == 1 ? w[1].x : w[2].x);
      const float wy = b == 0 ? w[0].y : (b == 1 ? w[1].y : w[2].y);
      const float wxy = wx * wy;
      const f3 tab = mk(fmaf(A1.x, dy[b], ta.x), fmaf(A1.y, dy[b], ta.y), fmaf(A1.z, dy[b], ta.z));
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const f3 U = mv + mk(fmaf(A2.x, dz[c], tab.x), fmaf(A2.y, dz[c], tab.y), fmaf(A2.z, dz[c], tab.z));
        // wts = wt * 2^22 with wt = (wx wy) wz: toFixed(wt q) = rint(clamp(wts q)) since the power-of-
        // two scaling is exact (a subnormal wt q rounds to 0 either way; overflow clamps alike)
        const float wts = wxy * wzs[c];
        const f3 mom = wts * U;
        if (CLAMP)
          add(bx + a, by + b, bz + c, fixed_of_scaled(mom.x), fixed_of_scaled(mom.y), fixed_of_scaled(mom.z),
              fixed_of_scaled(wts * mass));
        else
          add(bx + a, by + b, bz + c, (int32_t)rintf(mom.x), (int32_t)rintf(mom.y), (int32_t)rintf(mom.z),
              (int32_t)rintf(wts * mass));
      }
    }
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

// Exact n / d for 1 <= d, n < 2^20, d <= 4096, by a multiply-high with m = ceil(2^32 / d): the excess
// e = m d - 2^32 lies in [0, d - 1], so n m / 2^32 = n / d + n e / (d 2^32) with n e < 2^32, which
// stays below floor(n / d) + 1 (the box loops index at most 2^11 nodes with extents <= 12^2).
struct BoxDiv {
  uint32_t m, d;
};
__device__ __forceinline__ BoxDiv box_div(uint32_t d) { return BoxDiv{d > 1u ? 0xffffffffu / d + 1u : 0u, d}; }
__device__ __forceinline__ uint32_t box_quot(uint32_t n, BoxDiv q) { return q.d > 1u ? __umulhi(n, q.m) : n; }
// Node t of an ex x ey x ez box (x fastest) as (lx, ly, lz); dx = box_div(ex), dxy = box_div(ex * ey).
__device__ __forceinline__ void box_coords(int t, int ex, int ey, BoxDiv dx, BoxDiv dxy, int& lx, int& ly, int& lz) {
  const uint32_t r = box_quot((uint32_t)t, dx);   // t / ex
  lz = (int)box_quot((uint32_t)t, dxy);           // t / (ex ey)
  lx = t - (int)r * ex;
  ly = (int)r - lz * ey;
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
// Coulomb floor, SDF colliders of the node's scene. vfree receives p/m + g (the velocity before the
// walls, floor and colliders act; zero for a node without mass).
__device__ __forceinline__ float4 grid_node(int4 q, int ix, int iy, int iz, const SceneGeo& geo,
                                            const float* __restrict__ cols, f3& vfree) {
  vfree = mk(0.0f, 0.0f, 0.0f);
  if (!(q.w > 0)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float m = (float)q.w * kInvFP;
  f3 v = mk(((float)q.x * kInvFP) / m + geo.g[0], ((float)q.y * kInvFP) / m + geo.g[1],
            ((float)q.z * kInvFP) / m + geo.g[2]);
  vfree = v;
  const f3 c = mk((float)ix, (float)iy, (float)iz);
  const float bound = geo.bound;
  const f3 hi = mk((float)geo.X - bound, (float)geo.Y - bound, (float)geo.Z - bound);
  if (c.y < bound && v.y < 0.0f) {
    const float vn = -v.y;
    v.y = 0.0f;
    const float vt = sqrtf(v.x * v.x + v.z * v.z);
    if (vt > 1e-9f) {
      const float f = fmaxf(0.0f, 1.0f - (geo.floor_friction * vn) / vt);
      v = mk(v.x * f, 0.0f, v.z * f);
    }
  }
  if (c.x < bound && v.x < 0.0f) v.x = 0.0f;
  if (c.z < bound && v.z < 0.0f) v.z = 0.0f;
  if (c.x > hi.x && v.x > 0.0f) v.x = 0.0f;
  if (c.y > hi.y && v.y > 0.0f) v.y = 0.0f;
  if (c.z > hi.z && v.z > 0.0f) v.z = 0.0f;
  for (uint32_t k = 0; k < geo.ncol; k++) {
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

// G2P for one particle (mpm.js:479-559): gather velocity and APIC C, advect, update F, plasticity,
// stress. Split in three: g2p_load (the record's x, J and F22 and the stencil base), g2p_gather
// (the 27-node sums; gather(bx, by, bz, a, b, c) returns the grid velocity of node
// (bx + a, by + b, bz + c)) and g2p_finish. g2p_finish reads record si of src and writes the new
// record to slot di of dst (the 8 streams). In place: src == dst and si == di; every stream a
// slot's new value depends on is read before that slot is written.
struct G2PIn {
  float4 x4, s5;  // S0 (x, row) and S5 (t02, t12, J, F22) of the record
  f3 xp, fx;
  int bx, by, bz;
};

// Particle records stream through G2P: each is read once and written once per substep, so their
// loads and stores are non-temporal (they do not displace the velocity boxes and tiles in the caches).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float4* p) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_stream(float4* p, float4 v) {
  nt_f4 w;
  w.x = v.x; w.y = v.y; w.z = v.z; w.w = v.w;
  __builtin_nontemporal_store(w, reinterpret_cast<nt_f4*>(p));
}

// g2p_load = g2p_fetch (the two loads) + g2p_base (what follows from them).
__device__ __forceinline__ void g2p_fetch(const float4* src, uint32_t si, uint32_t cap, G2PIn& in) {
  in.x4 = load_stream(src + si);
  in.s5 = load_stream(src + (size_t)5 * cap + si);
}
__device__ __forceinline__ void g2p_base(G2PIn& in) {
  in.xp = mk(in.x4.x, in.x4.y, in.x4.z);
  in.bx = (int)floorf(in.xp.x - 0.5f);
  in.by = (int)floorf(in.xp.y - 0.5f);
  in.bz = (int)floorf(in.xp.z - 0.5f);
  in.fx = in.xp - mk((float)in.bx, (float)in.by, (float)in.bz);
}
__device__ __forceinline__ void g2p_load(const float4* src, uint32_t si, uint32_t cap, G2PIn& in) {
  g2p_fetch(src, si, cap, in);
  g2p_base(in);
}

// v = sum wt gv and B = sum outer(wt gv, dpos) over the 27 nodes (mpm.js:487-496) with
// wt = wx_a wy_b wz_c and dpos = (a, b, c) - fx, summed per (a, b) row of three nodes:
// S = sum_c wz_c gv and T = sum_c (wz_c dz_c) gv, then v += wxy S, B col x += (wxy dx_a) S,
// B col y += (wxy dy_b) S, B col z += wxy T (wxy = wx_a wy_b); every multiply-add contracted. The
// same sums as node by node, with the rounding fixed this way.
template <class Gather>
__device__ __forceinline__ void g2p_gather(const G2PIn& in, f3& v, m3& B, Gather&& gather) {
  const f3 fx = in.fx;
  f3 w[3];
  bspline(fx, w);
  const f3 d[3] = {mk(0.0f, 0.0f, 0.0f) - fx, mk(1.0f, 1.0f, 1.0f) - fx, mk(2.0f, 2.0f, 2.0f) - fx};
  const float wzd[3] = {w[0].z * d[0].z, w[1].z * d[1].z, w[2].z * d[2].z};
  v = mk(0, 0, 0);
  B = zero3();
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const float wx = a == 0 ? w[0].x : (a == 1 ? w[1].x : w[2].x);
      const float wy = b == 0 ? w[0].y : (b == 1 ? w[1].y : w[2].y);
      const float dxa = a == 0 ? d[0].x : (a == 1 ? d[1].x : d[2].x);
      const float dyb = b == 0 ? d[0].y : (b == 1 ? d[1].y : d[2].y);
      const float4 g0 = gather(in.bx, in.by, in.bz, a, b, 0);
      const float4 g1 = gather(in.bx, in.by, in.bz, a, b, 1);
      const float4 g2 = gather(in.bx, in.by, in.bz, a, b, 2);
      f3 S = mk(w[0].z * g0.x, w[0].z * g0.y, w[0].z * g0.z);
      S = mk(fmaf(w[1].z, g1.x, S.x), fmaf(w[1].z, g1.y, S.y), fmaf(w[1].z, g1.z, S.z));
      S = mk(fmaf(w[2].z, g2.x, S.x), fmaf(w[2].z, g2.y, S.y), fmaf(w[2].z, g2.z, S.z));
      f3 T = mk(wzd[0] * g0.x, wzd[0] * g0.y, wzd[0] * g0.z);
      T = mk(fmaf(wzd[1], g1.x, T.x), fmaf(wzd[1], g1.y, T.y), fmaf(wzd[1], g1.z, T.z));
      T = mk(fmaf(wzd[2], g2.x, T.x), fmaf(wzd[2], g2.y, T.y), fmaf(wzd[2], g2.z, T.z));
      const float wxy = wx * wy;
      const float bxs = wxy * dxa, bys = wxy * dyb;
      v = mk(fmaf(wxy, S.x, v.x), fmaf(wxy, S.y, v.y), fmaf(wxy, S.z, v.z));
      B.c[0] = mk(fmaf(bxs, S.x, B.c[0].x), fmaf(bxs, S.y, B.c[0].y), fmaf(bxs, S.z, B.c[0].z));
      B.c[1] = mk(fmaf(bys, S.x, B.c[1].x), fmaf(bys, S.y, B.c[1].y), fmaf(bys, S.z, B.c[1].z));
      B.c[2] = mk(fmaf(wxy, T.x, B.c[2].x), fmaf(wxy, T.y, B.c[2].y), fmaf(wxy, T.z, B.c[2].z));
    }
}

__device__ __forceinline__ void g2p_finish(const float* __restrict__ mats, const SceneGeo& geo, const float4* src,
                                           uint32_t si, float4* dst, uint32_t di, uint32_t cap, float vmax,
                                           const G2PIn& in, f3 v, const m3& B, float4 (&r)[8]) {
  const float4 x4 = in.x4, s5 = in.s5;
  const float* m0 = mats + 12 * __float_as_uint(x4.w);
  const float bound = geo.bound;
  const float vw0 = s5.z;
  const f3 xp = in.xp;
  const m3 C = 4.0f * B;
  const float sp = len3(v);
  if (sp > vmax) v = v * (vmax / sp);
  const float lo = bound - 1.0f;
  const f3 x = mk(clampf(xp.x + v.x, lo, ((float)geo.X - bound) + 1.0f), clampf(xp.y + v.y, lo, ((float)geo.Y - bound) + 1.0f),
                  clampf(xp.z + v.z, lo, ((float)geo.Z - bound) + 1.0f));
  const uint32_t kind = __float_as_uint(m0[0]);
  m3 F, tau;
  float vw = vw0;
  if (kind == 0u) {
    const float J = clampf(vw * (1.0f + trace(C)), 0.6f, 1.4f);
    vw = J;
    tau = (m0[2] * fminf(J - 1.0f, 0.02f)) * ident() + (m0[4] * J) * (C + tr(C));
    F = ident();
  } else {
    const float4 f6 = load_stream(src + (size_t)6 * cap + si), f7 = load_stream(src + (size_t)7 * cap + si);
    F = mul(ident() + C, mat(mk(f6.x, f6.y, f6.z), mk(f6.w, f7.x, f7.y), mk(f7.z, f7.w, s5.w)));
    const float mu0 = m0[2], la0 = m0[3];
    // snow and sand share one SVD of F (one copy of its code; a wavefront holding both runs it once)
    m3 U, Vm;
    f3 sv;
    if (kind != 1u) svd3(F, U, sv, Vm);
    if (kind == 1u) {
      const float J = fmaxf(det3(F), 0.05f);
      tau = mu0 * (mul(F, tr(F)) - ident()) + (la0 * log_cw(J)) * ident();
    } else if (kind == 2u) {
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
  r[1] = make_float4(v.x, v.y, v.z, tau.c[1].x);
  r[2] = make_float4(C.c[0].x, C.c[0].y, C.c[0].z, tau.c[0].x);
  r[3] = make_float4(C.c[1].x, C.c[1].y, C.c[1].z, tau.c[1].y);
  r[4] = make_float4(C.c[2].x, C.c[2].y, C.c[2].z, tau.c[2].z);
  r[5] = make_float4(tau.c[2].x, tau.c[2].y, vw, F.c[2].z);
  r[6] = make_float4(F.c[0].x, F.c[0].y, F.c[0].z, F.c[1].x);
  r[7] = make_float4(F.c[1].y, F.c[1].z, F.c[2].x, F.c[2].y);
#pragma unroll
  for (int k = 0; k < 8; k++) store_stream(dst + (size_t)k * cap + di, r[k]);
}

template <class Gather>
__device__ __forceinline__ void g2p_particle(const float* __restrict__ mats, const SceneGeo& geo, const float4* src,
                                             uint32_t si, float4* dst, uint32_t di, uint32_t cap, float vmax,
                                             Gather&& gather) {
  G2PIn in;
  g2p_load(src, si, cap, in);
  f3 v;
  m3 B;
  g2p_gather(in, v, B, gather);
  float4 r[8];
  g2p_finish(mats, geo, src, si, dst, di, cap, vmax, in, v, B, r);
}

template <int WAVES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_g2p(
    Pool pool, float4* __restrict__ P, uint32_t n, float vmax) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const SceneGeo& geo = pool.scenes[__float_as_uint(P[i].w) >> 3];
  const float4* __restrict__ V = pool.V;
  g2p_particle(pool.mats, geo, P, i, P, i, pool.cap, vmax,
               [&](int bx, int by, int bz, int a, int b, int c) { return V[node_addr(bx + a, by + b, bz + c, geo)]; });
}

// G2P of the substep that sorted: slot i of the new order takes particle perm[i] of the old one, so
// the sort's reorder rides on the record writes G2P makes anyway (src and dst are the two buffers).
template <int WAVES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_g2p_perm(
    Pool pool, const float4* __restrict__ src, float4* __restrict__ dst, const uint32_t* __restrict__ perm, uint32_t n,
    float vmax) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t si = perm[i];
  const SceneGeo& geo = pool.scenes[__float_as_uint(src[si].w) >> 3];
  const float4* __restrict__ V = pool.V;
  g2p_particle(pool.mats, geo, src, si, dst, i, pool.cap, vmax,
               [&](int bx, int by, int bz, int a, int b, int c) { return V[node_addr(bx + a, by + b, bz + c, geo)]; });
}

// Work items (item_sched.hpp) with this launch's grid and the device work counters at wctr.
__device__ __forceinline__ uint32_t first_item(uint32_t items) { return first_item_of(items, gridDim.x, blockIdx.x); }
__device__ __forceinline__ uint32_t next_item(uint32_t* wctr, uint32_t items, uint32_t& range, uint32_t& spent) {
  return next_item_of([&](uint32_t r) { return atomicAdd(wctr + kWorkStride * r, 1u); }, items, gridDim.x, range,
                      spent);
}

// G2P over the work items of this substep's P2G. The item's tile box (tile_hdr: the nodes its
// in-window particles' stencils cover, unchanged since P2G because no particle has moved) is copied
// from V into LDS, and the 27-node gathers of the item's particles read it there; a wavefront with
// a particle outside the box (it took P2G's fallback path) gathers from V for all its lanes (the
// same values). perm: sorted position -> record (nullptr: the identity). sorted_out: record perm[j]
// of src goes to slot j of dst (the sort's reorder applied by the writes); otherwise the update is
// in place (src == dst, each record back to its own slot).
template <int WAVES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_g2p_items(
    Pool pool, const float4* src, float4* dst, const uint32_t* __restrict__ perm, int sorted_out, float vmax,
    uint32_t* __restrict__ keys_out) {
  __shared__ float4 box[kTileNodes];
  __shared__ uint32_t s_item;
  const int tid = threadIdx.x;
  const uint32_t items = pool.ctr[kCtrItems];
  const uint32_t mshift = pool.key_bits - 6;
  uint32_t* const wctr = pool.arena_ctr + kArenaRegions * kArenaStride + kWorkSplits * kWorkStride;
  uint32_t range = blockIdx.x % kWorkSplits, spent = 0;
  const float4* __restrict__ V = pool.V;
  if (tid == 0) s_item = first_item(items);
  __syncthreads();
  for (;;) {
    const uint32_t k = s_item;
    if (k >= items) break;
    const uint32_t start = pool.wl_start[k], cnt = pool.wl_start[k + 1] - start;
    const uint32_t scene = pool.wl_block[k] >> mshift;
    const bool has_box = pool.tile_off[k] != 0xffffffffu;
    const int4 h = pool.tile_hdr[k];  // meaningful when has_box; loaded beside tile_off
    const SceneGeo geo = uniform_geo(pool, scene);
    int ox = 0, oy = 0, oz = 0, ex = 0, ey = 0, ez = 0;
    if (has_box) {
      ox = h.x; oy = h.y; oz = h.z;
      ex = h.w & 255; ey = (h.w >> 8) & 255; ez = (h.w >> 16) & 255;
    }
    const int nodes = ex * ey * ez;  // <= kTileNodes
    const BoxDiv dx = box_div((uint32_t)ex), dxy = box_div((uint32_t)(ex * ey));
    for (int t = tid; t < nodes; t += kThreads) {
      int lx, ly, lz;
      box_coords(t, ex, ey, dx, dxy, lx, ly, lz);
      box[t] = V[node_addr(ox + lx, oy + ly, oz + lz, geo)];
    }
    __syncthreads();
    for (uint32_t li = tid; li < cnt; li += kThreads) {
      const uint32_t j = start + li;
      const uint32_t si = perm ? perm[j] : j;
      G2PIn in;
      g2p_load(src, si, pool.cap, in);
      const bool inbox = (unsigned)(in.bx - ox) <= (unsigned)(ex - 3) && (unsigned)(in.by - oy) <= (unsigned)(ey - 3) &&
                         (unsigned)(in.bz - oz) <= (unsigned)(ez - 3) && nodes > 0;
      f3 v;
      m3 B;
      if (__builtin_amdgcn_ballot_w64(!inbox) == 0) {
        const int tb = ((in.bz - oz) * ey + (in.by - oy)) * ex + (in.bx - ox);
        g2p_gather(in, v, B, [&](int, int, int, int a, int b, int c) { return box[tb + (c * ey + b) * ex + a]; });
      } else {
        g2p_gather(in, v, B,
                   [&](int bx, int by, int bz, int a, int b, int c) { return V[node_addr(bx + a, by + b, bz + c, geo)]; });
      }
      float4 r[8];
      const uint32_t di = sorted_out ? j : si;
      g2p_finish(pool.mats, geo, src, si, dst, di, pool.cap, vmax, in, v, B, r);
      if (keys_out) keys_out[di] = sort_key(pool, scene, geo.active, r[0]);  // for the next substep's sort
    }
    // the next item, fetched while the other waves finish (every thread read k before the barrier above)
    if (tid == 0) s_item = next_item(wctr, items, range, spent);
    __syncthreads();  // the item is finished with the box; s_item holds the next item
  }
}

__device__ __forceinline__ float& comp4(float4& q, int c) { return c == 0 ? q.x : (c == 1 ? q.y : (c == 2 ? q.z : q.w)); }

// Reference records (32 floats each) <-> the 8 streams, one particle per thread (kRecSlot).
__global__ void k_records_in(const float4* __restrict__ aos, float4* __restrict__ P, uint32_t cap, uint32_t offset,
                             uint32_t n, uint32_t row_base) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 in[8], out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) in[k] = aos[(size_t)i * 8 + k];
#pragma unroll
  for (int f = 0; f < 32; f++) comp4(out[kRecSlot[f] >> 2], kRecSlot[f] & 3) = comp4(in[f >> 2], f & 3);
  out[0].w = __uint_as_float(__float_as_uint(out[0].w) + row_base);  // scene-local -> pool material row
#pragma unroll
  for (int k = 0; k < 8; k++) P[(size_t)k * cap + offset + i] = out[k];
}

__global__ void k_records_out(const float4* __restrict__ P, uint32_t cap, uint32_t offset, uint32_t n,
                              uint32_t row_base, float4* __restrict__ aos) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 in[8], out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) in[k] = P[(size_t)k * cap + offset + i];
#pragma unroll
  for (int f = 0; f < 32; f++) comp4(out[f >> 2], f & 3) = comp4(in[kRecSlot[f] >> 2], kRecSlot[f] & 3);
  out[0].w = __uint_as_float(__float_as_uint(out[0].w) - row_base);
#pragma unroll
  for (int k = 0; k < 8; k++) aos[(size_t)i * 8 + k] = out[k];
}

__global__ void k_rescale(float4* __restrict__ P, uint32_t cap, uint32_t offset, uint32_t n, float k) {  // mpm.js:405-413
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int s = 1; s <= 4; s++) {
    float4 q = P[(size_t)s * cap + offset + i];
    q.x = q.x * k; q.y = q.y * k; q.z = q.z * k;
    P[(size_t)s * cap + offset + i] = q;
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void k_block_info(uint2* __restrict__ info, uint32_t base, uint32_t n, uint32_t scene, uint32_t nbx,
                             uint32_t nby) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t bx = i % nbx, by = (i / nbx) % nby, bz = i / (nbx * nby);
  info[base + i] = make_uint2(scene, bx | (by << 10) | (bz << 20));
}

__global__ void k_begin_p2g(uint32_t* ctr, uint32_t* arena_ctr) {
  const int t = threadIdx.x;  // kArenaRegions threads
  arena_ctr[t * kArenaStride] = 0u;
  if (t < 2 * kWorkSplits) arena_ctr[kArenaRegions * kArenaStride + kWorkStride * t] = 0u;  // P2G's, G2P's
  if (t == 0) {
    ctr[kCtrEpoch] += 1u;
    ctr[kCtrD] = 0u;
  }
}

__global__ void k_begin_bin(uint32_t* ctr) {
  ctr[kCtrBinEpoch] += 1u;
  ctr[kCtrS] = 0u;
  ctr[kCtrItems] = 0u;
}

// k_begin_bin and the occupancy bitmap cleared (occ_n 32-bit words), in one launch.
__global__ void k_bin_reset(Pool pool, uint32_t occ_n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < occ_n) reinterpret_cast<uint32_t*>(pool.occ)[i] = 0u;
  if (i == 0) {
    pool.ctr[kCtrBinEpoch] += 1u;
    pool.ctr[kCtrS] = 0u;
    pool.ctr[kCtrItems] = 0u;
  }
}


// Sort keys of the particles in their current order (sort_key); the radix sort's values are the
// indices, so its stable order breaks ties by them.
__global__ void k_keys(Pool pool, const float4* __restrict__ P, uint32_t n, uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = P[i];
  const uint32_t scene = __float_as_uint(x.w) >> 3;
  keys[i] = sort_key(pool, scene, pool.scenes[scene].active, x);
}

__global__ void k_live_keys(Pool pool, const float4* __restrict__ P, uint32_t n, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t scene = __float_as_uint(P[i].w) >> 3;
  keys[i] = pool.scenes[scene].active ? 0u : 1u;
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
// Work items from the block-sorted keys: a head is a particle whose block (key >> 6) differs from
// its predecessor's. Tiles of kItemTile keys, 16 consecutive keys per thread: k_bin_count stores
// the heads of each tile, and after an exclusive scan of those counts k_bin_emit numbers the heads
// in order (item k starts at the k-th head) and writes wl_start and wl_block.
constexpr int kItemPer = 16;
constexpr uint32_t kItemTile = kThreads * kItemPer;

__device__ __forceinline__ uint32_t head_mask(const uint32_t* __restrict__ keys, uint32_t n, uint32_t first,
                                              uint32_t k[kItemPer]) {
  uint32_t prev = first > 0 && first <= n ? keys[first - 1] >> 6 : 0xffffffffu;
  uint32_t m = 0;
  if (first + kItemPer <= n) {  // keys is 16-byte aligned and first a multiple of 16
    const uint4* q = reinterpret_cast<const uint4*>(keys + first);
#pragma unroll
    for (int j = 0; j < kItemPer / 4; j++) {
      const uint4 v = q[j];
      k[4 * j] = v.x; k[4 * j + 1] = v.y; k[4 * j + 2] = v.z; k[4 * j + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kItemPer; j++) k[j] = first + j < n ? keys[first + j] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kItemPer; j++) {
    const uint32_t b = k[j] >> 6;
    if (first + j < n && (first + j == 0 || b != prev)) m |= 1u << j;
    prev = b;
  }
  return m;
}

__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < kThreads / 64; q++) {
    const uint32_t x = wsum[q];
    if (q < w) off += x;
    tot += x;
  }
  *total = tot;
  return off + inc - v;
}

__global__ __launch_bounds__(kThreads) void k_bin_count(const uint32_t* __restrict__ keys, uint32_t n,
                                                        uint32_t* __restrict__ tile_heads) {
  uint32_t k[kItemPer];
  const uint32_t m = head_mask(keys, n, blockIdx.x * kItemTile + threadIdx.x * kItemPer, k);
  uint32_t tot;
  block_exclusive_sum((uint32_t)__builtin_popcount(m), &tot);
  if (threadIdx.x == 0) tile_heads[blockIdx.x] = tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) tile_heads[gridDim.x] = 0;  // becomes the total after the scan
}

// scanned: tile_first[t] is the exclusive scan of the tile head counts (tiles + 1 entries; the last
// is the total). Otherwise it holds the counts themselves, and each workgroup adds up those of the
// tiles before it (workgroup 0: of all tiles, the total); for at most kItemScanFree tiles.
constexpr uint32_t kItemScanFree = 4096;
__global__ __launch_bounds__(kThreads) void k_bin_emit(Pool pool, const uint32_t* __restrict__ keys, uint32_t n,
                                                       const uint32_t* __restrict__ tile_first, uint32_t tiles,
                                                       int scanned) {
  uint32_t k[kItemPer];
  const uint32_t first = blockIdx.x * kItemTile + threadIdx.x * kItemPer;
  uint32_t m = head_mask(keys, n, first, k);
  uint32_t base = 0, total = 0;
  if (scanned) {
    base = tile_first[blockIdx.x];
  } else {
    const uint32_t lim = blockIdx.x == 0 ? tiles : blockIdx.x;
    uint32_t ps = 0;
    for (uint32_t t = threadIdx.x; t < lim; t += kThreads) ps += tile_first[t];
    uint32_t sum;
    block_exclusive_sum(ps, &sum);
    __syncthreads();  // the block sum's LDS is used again below
    base = blockIdx.x == 0 ? 0u : sum;
    total = sum;
  }
  uint32_t tot;
  uint32_t item = base + block_exclusive_sum((uint32_t)__builtin_popcount(m), &tot);
  const uint32_t mmask = (1u << (pool.key_bits - 6)) - 1u;
  while (m) {
    const int j = __builtin_ctz(m);
    m &= m - 1;
    const uint32_t blk = k[j] >> 6, scene = k[j] >> pool.key_bits, mb = blk & mmask;
    pool.wl_start[item] = first + j;
    pool.wl_block[item] = blk;
    {  // the item's velocity box until its first P2G: the stencils of bases in the block, in the grid
      const SceneGeo& g = pool.scenes[scene];
      const int ox = 4 * (int)compact3(mb), oy = 4 * (int)compact3(mb >> 1), oz = 4 * (int)compact3(mb >> 2);
      const int ex = max(min(6, g.X - ox), 0), ey = max(min(6, g.Y - oy), 0), ez = max(min(6, g.Z - oz), 0);
      pool.tile_off[item] = 0u;
      pool.tile_hdr[item] = make_int4(ox, oy, oz, ex | (ey << 8) | (ez << 16));
    }
    const uint32_t pb = 1u << (pool.key_bits / 3 - 2), W = occ_row_words(pb);
    const uint32_t bx = compact3(mb), by = compact3(mb >> 1), bz = compact3(mb >> 2);
    atomicOr(pool.occ + ((size_t)scene * pb * pb + (size_t)bz * pb + by) * W + (bx >> 6), 1ull << (bx & 63u));
    item++;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint32_t items = scanned ? tile_first[tiles] : total;
    pool.wl_start[items] = n;
    pool.ctr[kCtrItems] = items;
  }
}

// Within-block transposition: the particles of item k arrive grouped by block and otherwise in
// their previous order; they leave ordered by (rank within cell, cell_in_block), where a particle's
// rank within its cell counts the earlier particles of the same cell (a stable order). Output
// position pos(r, c) = S(r) + #{c' < c : n_c' > r} with S(r) = sum_c' min(n_c', r) over the 64 cells
// c' of the block with counts n_c'. One wavefront per item, lane c owning cell c: it counts the
// cells, tabulates S(r) and the mask of cells with n_c' > r for r < kRankTable (a cell holds ~8
// particles, up to ~200 where a material lands), then ranks the item in index order (per-cell lane
// masks in LDS find the lanes of equal cell) so a position is one table read and a popcount.
// out[start + pos] = in[start + j].
// A particle's cell inside its 4^3-cell block, numbered x fastest (cx + 4 cy + 16 cz) from the low 6
// bits of its key (the cell's Morton code, bits x0 y0 z0 x1 y1 z1). Ordering a block's rank layers by
// this number puts the 16 cells of one z layer in each quarter of a wavefront: the quarter that one
// LDS cycle of a 96-bit velocity-box read serves then reads 16 distinct banks (fixed box strides 12
// and 96), where Morton order put a 4 x 2 x 2 group there and read each bank twice.
__device__ __forceinline__ uint32_t cell_in_block(uint32_t key) {
  const uint32_t m = key & 63u;
  return (m & 1u) | ((m >> 2) & 2u) | (((m >> 1) & 1u) << 2) | (((m >> 4) & 1u) << 3) | (((m >> 2) & 1u) << 4) |
         (((m >> 5) & 1u) << 5);
}
constexpr int kRankTable = 256;
constexpr int kBinWaves = kThreads / 64;
constexpr int kCellCache = 1024;  // cells of an item's first 1024 particles kept in LDS by the count pass
constexpr int kTransposeBatch = 8;  // steps of 64 particles whose loads are issued together
__global__ __launch_bounds__(kThreads) void k_bin_transpose(Pool pool, const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ in,
                                                            uint32_t* __restrict__ out) {
  __shared__ uint32_t wrun[kBinWaves][64];  // per wave: cell counts, then the next rank per cell
  __shared__ uint32_t wcnt[kBinWaves][64];
  __shared__ uint32_t srank[kBinWaves][kRankTable];
  __shared__ uint64_t mrank[kBinWaves][kRankTable];
  __shared__ uint8_t wcell[kBinWaves][kCellCache];
  __shared__ unsigned long long wmask[kBinWaves][64];  // lanes of one step per cell; zero between steps
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const uint32_t items = pool.ctr[kCtrItems];
  // workgroups are dealt to the 8 XCDs in turn (the grid is a multiple of 8): XCD x takes the
  // contiguous items [x * per, (x + 1) * per), so the particle ranges it writes share no cache line
  // with another XCD's but at the ends
  const uint32_t per = (items + 7u) / 8u, xcd = blockIdx.x % 8u, wgs = gridDim.x / 8u;
  const uint32_t kend = min(items, (xcd + 1u) * per);
  wmask[w][lane] = 0ull;
  for (uint32_t k = xcd * per + (blockIdx.x / 8u) * kBinWaves + w; k < kend; k += wgs * kBinWaves) {
    const uint32_t start = __builtin_amdgcn_readfirstlane(pool.wl_start[k]);
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(pool.wl_start[k + 1]) - start;
    wrun[w][lane] = 0;
    // the cell counts (LDS atomics; their order does not matter), kTransposeBatch steps of loads at once
    for (uint32_t j0 = 0; j0 < cnt; j0 += 64 * kTransposeBatch) {
      uint32_t cb[kTransposeBatch];
#pragma unroll
      for (int u = 0; u < kTransposeBatch; u++) {
        const uint32_t j = j0 + 64 * u + lane;
        cb[u] = j < cnt ? cell_in_block(keys[start + j]) : 64u;
      }
#pragma unroll
      for (int u = 0; u < kTransposeBatch; u++) {
        const uint32_t j = j0 + 64 * u + lane;
        if (cb[u] < 64u) {
          atomicAdd(&wrun[w][cb[u]], 1u);
          if (j < (uint32_t)kCellCache) wcell[w][j] = (uint8_t)cb[u];
        }
      }
    }
    const uint32_t nc = wrun[w][lane];  // particles in cell `lane`
    wcnt[w][lane] = nc;
    wrun[w][lane] = 0;
    uint32_t rmax = nc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) rmax = max(rmax, (uint32_t)__shfl_xor((int)rmax, d, 64));
    const uint32_t rtab = min(rmax, (uint32_t)kRankTable);
    uint32_t sr = 0;
    for (uint32_t r = 0; r < rtab; r++) {
      const uint64_t m = __ballot(nc > r);
      if (lane == 0) {
        srank[w][r] = sr;
        mrank[w][r] = m;
      }
      sr += (uint32_t)__popcll(m);
    }
    for (uint32_t j0 = 0; j0 < cnt; j0 += 64 * kTransposeBatch) {
      uint32_t vb[kTransposeBatch], kb[kTransposeBatch];  // this batch's values (and keys past the cache)
#pragma unroll
      for (int u = 0; u < kTransposeBatch; u++) {
        const uint32_t j = j0 + 64 * u + lane;
        vb[u] = j < cnt ? in[start + j] : 0u;
        kb[u] = j < cnt && j >= (uint32_t)kCellCache ? cell_in_block(keys[start + j]) : 0u;
      }
#pragma unroll
      for (int u = 0; u < kTransposeBatch; u++) {
        if (j0 + 64 * u >= cnt) break;  // uniform
        const uint32_t j = j0 + 64 * u + lane;
        const bool act = j < cnt;
        const uint32_t v = vb[u];
        const uint32_t c = !act ? 0u : (j < (uint32_t)kCellCache ? (uint32_t)wcell[w][j] : kb[u]);
        // rank within the cell: the count of its earlier steps (wrun) plus the lanes of this step
        // below this one with the same cell. The lanes of a cell find each other by OR-ing their bits
        // into the cell's mask and reading it back (ORs commute: no atomic decides an order; the
        // wavefront's LDS operations complete in program order, and the fences keep the compiler to it)
        if (act) __hip_atomic_fetch_or(&wmask[w][c], 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint64_t peers = act ? __hip_atomic_load(&wmask[w][c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0ull;
        const uint32_t r = act ? wrun[w][c] + (uint32_t)__popcll(peers & below) : 0u;
        if (act) {
          uint32_t pos;
          if (r < (uint32_t)kRankTable) {
            pos = srank[w][r] + (uint32_t)__builtin_popcountll(mrank[w][r] & ((1ull << c) - 1ull));
          } else {
            pos = 0;
            for (uint32_t q = 0; q < 64; q++) {
              const uint32_t nq = wcnt[w][q];
              pos += min(nq, r) + ((q < c && nq > r) ? 1u : 0u);
            }
          }
          out[start + pos] = v;
        }
        // every lane has read wrun[w][c] and the mask; the first lane of each cell advances the one and
        // clears the other
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (act && (peers & below) == 0) {
          wrun[w][c] += (uint32_t)__popcll(peers);
          __hip_atomic_store(&wmask[w][c], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
      }
    }
  }
}

// S: the node blocks [B-1, B+D-1]^3 of every work item B (everything its drift-D tile can reach:
// nodes [4B-D, 4B+6+D)), i.e. node block b when a particle block in [b-D+1, b+1]^3 is occupied (D =
// 2 or 3). One thread per 64-block word of a row
// (by, bz) of a scene's node-block grid: the row's word ORs the x-dilated occupancy words
// (O | O<<1 | O<<2 | O>>1) of the 16 rows (by-2..by+1, bz-2..bz+1); its set bits inside the grid are
// appended to slist (one atomic per wave) and stamped.
__device__ __forceinline__ unsigned long long occ_word(const Pool& pool, uint32_t scene, uint32_t pb, uint32_t W, int by,
                                                       int bz, int w) {
  if (by < 0 || bz < 0 || by >= (int)pb || bz >= (int)pb || w < 0 || w >= (int)W) return 0ull;
  return pool.occ[((size_t)scene * pb * pb + (size_t)bz * pb + (size_t)by) * W + (size_t)w];
}
// kDilateSplit threads per word: each computes the word (the same loads) and writes the S entries
// of its 64 / kDilateSplit bits, so a dense row's 64 entries are written by several threads.
constexpr int kDilateSplit = 8;
__global__ __launch_bounds__(kThreads) void k_bin_dilate(Pool pool, uint32_t max_scenes, int drift) {
  const uint32_t pb = 1u << (pool.key_bits / 3 - 2), W = occ_row_words(pb);
  const uint32_t epoch = pool.ctr[kCtrBinEpoch];
  const int lane = threadIdx.x & 63;
  const uint64_t total = (uint64_t)max_scenes * pb * pb * W * kDilateSplit;
  for (uint64_t t0 = (uint64_t)blockIdx.x * kThreads; t0 < total; t0 += (uint64_t)gridDim.x * kThreads) {
    const uint64_t ts = t0 + threadIdx.x;  // the loop bound is uniform over the block
    const uint64_t t = ts / kDilateSplit;  // the word
    const int part = (int)(ts % kDilateSplit);
    unsigned long long d = 0;
    uint32_t scene = 0, base = 0;
    int w = 0, by = 0, bz = 0, nbx = 0, nby = 0;
    if (ts < total) {
      w = (int)(t % W);
      by = (int)((t / W) % pb);
      bz = (int)((t / ((uint64_t)W * pb)) % pb);
      scene = (uint32_t)(t / ((uint64_t)W * pb * pb));
      const SceneGeo& geo = pool.scenes[scene];
      nbx = geo.nbx; nby = geo.nby;
      base = geo.block_base;
      if (geo.active && by < geo.nby && bz < geo.nbz && 64 * w < geo.nbx) {
        const int r = drift - 1;  // particle blocks b-r .. b+1 reach node block b
        for (int dz = -r; dz <= 1; dz++)
          for (int dy = -r; dy <= 1; dy++) {
            const unsigned long long o = occ_word(pool, scene, pb, W, by + dy, bz + dz, w);
            const unsigned long long lo = occ_word(pool, scene, pb, W, by + dy, bz + dz, w - 1);
            const unsigned long long hi = occ_word(pool, scene, pb, W, by + dy, bz + dz, w + 1);
            d |= o | (o << 1) | (lo >> 63) | (o >> 1) | (hi << 63);
            if (r == 2) d |= (o << 2) | (lo >> 62);
          }
        const int valid = geo.nbx - 64 * w;  // bits of this word inside the grid
        if (valid < 64) d &= (1ull << valid) - 1ull;
        constexpr int kBits = 64 / kDilateSplit;  // this thread's bits of the word
        d &= ((1ull << kBits) - 1ull) << (kBits * part);
      }
    }
    // wave prefix of the counts, one atomic per wave
    const uint32_t c = (uint32_t)__popcll(d);
    uint32_t inc = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t x = __shfl_up(inc, o, 64);
      if (lane >= o) inc += x;
    }
    uint32_t wbase = 0;
    if (lane == 63 && inc) wbase = atomicAdd(pool.ctr + kCtrS, inc);
    wbase = (uint32_t)__shfl((int)wbase, 63, 64);
    uint32_t at = wbase + inc - c;
    while (d) {
      const int i = __builtin_ctzll(d);
      d &= d - 1;
      const uint32_t b = base + (uint32_t)(64 * w + i) + (uint32_t)nbx * ((uint32_t)by + (uint32_t)nby * (uint32_t)bz);
      pool.sinfo[at] = make_uint2(scene, (uint32_t)(64 * w + i) | ((uint32_t)by << 10) | ((uint32_t)bz << 20));
      pool.slist[at++] = b;
      pool.sstamp[b] = epoch;
    }
  }
}

// A fallback write to node block b: mark it written in this P2G (the grid update reads and clears A
// there) and list it in D unless S already covers it (use_s).
__device__ __forceinline__ void stamp_fallback(const Pool& pool, uint32_t b, uint32_t epoch, bool use_s,
                                               uint32_t bin_epoch) {
  if (__hip_atomic_load(&pool.stamp[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch)
    if (atomicExch(&pool.stamp[b], epoch) != epoch)
      if (!use_s || pool.sstamp[b] != bin_epoch) pool.dlist[atomicAdd(pool.ctr + kCtrD, 1u)] = b;
}

__device__ __forceinline__ void fallback_particle(const Pool& pool, const float4 r[kP2GStreams], uint32_t epoch, bool use_s,
                                                  uint32_t bin_epoch) {
  const SceneGeo& geo = scene_of_row(pool, __float_as_uint(r[0].w));
  const float vol4 = geo.vol4;
  const int bx = (int)floorf(r[0].x - 0.5f), by = (int)floorf(r[0].y - 0.5f), bz = (int)floorf(r[0].z - 0.5f);
  for (int kz = bz >> 2; kz <= (bz + 2) >> 2; kz++)
    for (int ky = by >> 2; ky <= (by + 2) >> 2; ky++)
      for (int kx = bx >> 2; kx <= (bx + 2) >> 2; kx++)
        stamp_fallback(pool, geo.block_base + (uint32_t)kx + (uint32_t)geo.nbx * ((uint32_t)ky + (uint32_t)geo.nby * (uint32_t)kz),
                       epoch, use_s, bin_epoch);
  p2g_particle(r, pool.mats, vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
    int* g = reinterpret_cast<int*>(pool.A + node_addr(gx, gy, gz, geo));
    atomicAdd(g, q0);
    atomicAdd(g + 1, q1);
    atomicAdd(g + 2, q2);
    atomicAdd(g + 3, q3);
  });
}

// Work-item loop state of a workgroup (kept by thread 0): the work counter range it takes items
// from, the ranges found used up, and whether its arena region was found full in this P2G.
struct ItemCursor {
  uint32_t* wctr;
  uint32_t items, range, spent;
  bool full;
};

// The end of an item's P2G (after a barrier that follows the bbox atomics): the LDS window tile
// (S^3 nodes from the lowest in-window base (wx, wy, wz), 4 planes) is stored over the bounding box
// bbox of the in-window stencil bases (+ 2 nodes) in arena space, recorded in tile_off / tile_hdr;
// when the arena has no room, the box's non-zero nodes are added to A with atomics instead (and
// their node blocks stamped), which gives the same sums. The nodes read are cleared, which leaves
// the whole tile zero for the next item (the stencils of the in-window particles lie in the box).
// slots (when the launch's items times N nodes fit the arena): item k's tile goes to its own slot
// at k N, which every thread knows, so the store starts at once (no barrier after an arena atomic;
// the caller took the workgroup's next item before the bounding box). Otherwise thread 0 takes the
// arena space and the next item with two atomics issued together (one round trip) before the store. Once
// an add finds the region full, the workgroup stops adding to it (an add past the end takes nothing,
// so a region's counter ends at most one tile per workgroup past its size). Every thread of the
// workgroup calls it; s_item is read after the caller's next barrier.
template <int THREADS, int D>
__device__ __forceinline__ void store_item_tile(const Pool& pool, uint32_t k, uint32_t wb, const SceneGeo& geo, int* tile,
                                                const int* bbox, int wx, int wy, int wz, uint32_t epoch,
                                                uint32_t bin_epoch, uint32_t& s_off, ItemCursor& cur,
                                                uint32_t& s_item, bool slots) {
  constexpr int S = Window<D>::kSpan, N = Window<D>::kNodes;
  const int tid = threadIdx.x;
  const bool any = bbox[0] != INT_MAX;
  const int px = bbox[0], py = bbox[1], pz = bbox[2];
  const int fx = any ? bbox[3] - px + 3 : 0, fy = any ? bbox[4] - py + 3 : 0, fz = any ? bbox[5] - pz + 3 : 0;
  const int pnodes = fx * fy * fz;  // <= N: inside the window tile
  if (slots) {
    const uint32_t toff = pnodes > 0 ? k * (uint32_t)N : 0xffffffffu;
    if (tid == 0) {
      pool.tile_off[k] = toff;
      const int ext = fx | (fy << 8) | (fz << 16);
      if (pnodes > 0) pool.tile_hdr[k] = make_int4(px, py, pz, ext);
      pool.mtile[wb] = make_uint4(toff, (uint32_t)px | ((uint32_t)py << 10) | ((uint32_t)pz << 20), (uint32_t)ext, epoch);
    }
    const BoxDiv dx = box_div((uint32_t)fx), dxy = box_div((uint32_t)(fx * fy));
    for (int t = tid; t < pnodes; t += THREADS) {
      int lx, ly, lz;
      box_coords(t, fx, fy, dx, dxy, lx, ly, lz);
      const int w = ((pz + lz - wz) * S + (py + ly - wy)) * S + (px + lx - wx);
      const int4 q = make_int4(tile[w], tile[N + w], tile[2 * N + w], tile[3 * N + w]);
      tile[w] = 0;
      tile[N + w] = 0;
      tile[2 * N + w] = 0;
      tile[3 * N + w] = 0;
      pool.tiles[toff + tile_index(px, py, pz, fx, fy, fz, lx, ly, lz)] = q;
    }
    return;
  }
  if (tid == 0) {  // arena space for the tile, from this workgroup's region
    const uint32_t region = blockIdx.x % kArenaRegions, rsize = pool.tile_nodes / kArenaRegions;
    const bool add = pnodes > 0 && !cur.full;
    uint32_t o = 0;
    if (add) o = atomicAdd(pool.arena_ctr + region * kArenaStride, (uint32_t)pnodes);
    s_item = next_item(cur.wctr, cur.items, cur.range, cur.spent);
    uint32_t off = 0xffffffffu;
    if (add) {
      off = arena_offset(region, rsize, o, (uint32_t)pnodes);
      if (off == 0xffffffffu) cur.full = true;
    }
    s_off = off;
    pool.tile_off[k] = off;
    const int ext = fx | (fy << 8) | (fz << 16);
    if (off != 0xffffffffu) pool.tile_hdr[k] = make_int4(px, py, pz, ext);
    // the grid update finds the tile by the block's slot (wb: scene and Morton block)
    pool.mtile[wb] = make_uint4(off, (uint32_t)px | ((uint32_t)py << 10) | ((uint32_t)pz << 20), (uint32_t)ext, epoch);
  }
  __syncthreads();
  const uint32_t toff = s_off;
  const BoxDiv dx = box_div((uint32_t)fx), dxy = box_div((uint32_t)(fx * fy));
  for (int t = tid; t < pnodes; t += THREADS) {
    int lx, ly, lz;
    box_coords(t, fx, fy, dx, dxy, lx, ly, lz);
    const int w = ((pz + lz - wz) * S + (py + ly - wy)) * S + (px + lx - wx);
    const int4 q = make_int4(tile[w], tile[N + w], tile[2 * N + w], tile[3 * N + w]);
    tile[w] = 0;
    tile[N + w] = 0;
    tile[2 * N + w] = 0;
    tile[3 * N + w] = 0;
    if (toff != 0xffffffffu) {
      pool.tiles[toff + tile_index(px, py, pz, fx, fy, fz, lx, ly, lz)] = q;
    } else if (q.x | q.y | q.z | q.w) {  // no arena room: add the node to A
      const int gx = px + lx, gy = py + ly, gz = pz + lz;
      stamp_fallback(pool, geo.block_base + (uint32_t)(gx >> 2) + (uint32_t)geo.nbx * ((uint32_t)(gy >> 2) +
                               (uint32_t)geo.nby * (uint32_t)(gz >> 2)), epoch, true, bin_epoch);
      int* g = reinterpret_cast<int*>(pool.A + node_addr(gx, gy, gz, geo));
      atomicAdd(g, q.x);
      atomicAdd(g + 1, q.y);
      atomicAdd(g + 2, q.z);
      atomicAdd(g + 3, q.w);
    }
  }
}

// P2G over the binning's work items (mpm.js:306-401 restated for aligned particle blocks). One
// 256-thread workgroup per item at a time, items from per-XCD counters. Each particle whose base
// lies within kDrift cells of the block scatters its 27x4 fixed-point contributions with LDS
// integer atomics into a tile covering the item's whole window (12^3 nodes); the tile is then
// stored with plain stores over the bounding box of those stencils, in space taken from the
// arena. The others, and the whole tile when the arena has no room, go to A with global atomics.
// Inside a block the sort orders particles by (rank in cell, cell), so the 64 lanes of one
// instruction read consecutive records (coalesced) that sit in different cells (their LDS atomics
// hit different nodes).
template <int WAVES, int D>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_p2g_binned(
    Pool pool, const float4* __restrict__ P, const uint32_t* __restrict__ perm, uint32_t* __restrict__ keys_out) {
  constexpr int S = Window<D>::kSpan, N = Window<D>::kNodes;  // window tile: S^3 nodes from the lowest in-window base
  __shared__ int tile[4 * N];
  __shared__ int bbox[6];
  __shared__ uint32_t s_item, s_off;
  const int tid = threadIdx.x;
  const uint32_t items = pool.ctr[kCtrItems];
  const uint32_t epoch = pool.ctr[kCtrEpoch], bin_epoch = pool.ctr[kCtrBinEpoch];
  const uint32_t mshift = pool.key_bits - 6;
  const bool slots = (uint64_t)items * N <= pool.tile_nodes;  // every item's tile in its own slot
  ItemCursor cur{pool.arena_ctr + kArenaRegions * kArenaStride, items, blockIdx.x % kWorkSplits, 0, false};
  for (int t = tid; t < N; t += kThreads)
    reinterpret_cast<int4*>(tile)[t] = make_int4(0, 0, 0, 0);  // the 4 planes; each item leaves them zero
  if (tid == 0) s_item = first_item(cur.items);
  __syncthreads();
  for (;;) {
    const uint32_t k = s_item;
    if (k >= items) break;
    const uint32_t start = pool.wl_start[k], cnt = pool.wl_start[k + 1] - start;
    const uint32_t wb = pool.wl_block[k];
    // this thread's first record
    uint32_t li = tid, si = 0;
    float4 r[kP2GStreams] = {};
    if (li < cnt) {
      si = perm ? perm[start + li] : start + li;
      load_p2g_record(P, pool.cap, si, r);
    }
    const SceneGeo geo = uniform_geo(pool, wb >> mshift);
    const uint32_t mb = wb & ((1u << mshift) - 1u);
    const int wx = 4 * (int)compact3(mb) - D, wy = 4 * (int)compact3(mb >> 1) - D,
              wz = 4 * (int)compact3(mb >> 2) - D;  // lowest in-window base
    if (tid < 3) { bbox[tid] = INT_MAX; bbox[3 + tid] = INT_MIN; }
    __syncthreads();
    int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
    for (; li < cnt; li += kThreads) {
      if (keys_out) keys_out[si] = sort_key(pool, wb >> mshift, geo.active, r[0]);
      const int b0 = (int)floorf(r[0].x - 0.5f), b1 = (int)floorf(r[0].y - 0.5f), b2 = (int)floorf(r[0].z - 0.5f);
      const bool inwin = in_window<D>(b0, b1, b2, wx, wy, wz);
      if (inwin) {
        lo0 = min(lo0, b0); lo1 = min(lo1, b1); lo2 = min(lo2, b2);
        hi0 = max(hi0, b0); hi1 = max(hi1, b1); hi2 = max(hi2, b2);
        const int tb = ((b2 - wz) * S + (b1 - wy)) * S + (b0 - wx);
        p2g_particle(r, pool.mats, geo.vol4, [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
          const int t = tb + ((gz - b2) * S + (gy - b1)) * S + (gx - b0);
          atomicAdd(&tile[t], q0);
          atomicAdd(&tile[N + t], q1);
          atomicAdd(&tile[2 * N + t], q2);
          atomicAdd(&tile[3 * N + t], q3);
        });
      } else {
        fallback_particle(pool, r, epoch, true, bin_epoch);
      }
      if (li + kThreads < cnt) {  // the next record of this thread
        const uint32_t jn = start + li + kThreads;
        si = perm ? perm[jn] : jn;
        load_p2g_record(P, pool.cap, si, r);
      }
    }
    // with slots, the next item is taken here (every thread read k before the first barrier), its round
    // trip hidden by the bounding box and the tile store
    if (slots && tid == 0) s_item = next_item(cur.wctr, cur.items, cur.range, cur.spent);
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    if ((tid & 63) == 0) {
      atomicMin(&bbox[0], lo0); atomicMin(&bbox[1], lo1); atomicMin(&bbox[2], lo2);
      atomicMax(&bbox[3], hi0); atomicMax(&bbox[4], hi1); atomicMax(&bbox[5], hi2);
    }
    __syncthreads();
    store_item_tile<kThreads, D>(pool, k, wb, geo, tile, bbox, wx, wy, wz, epoch, bin_epoch, s_off, cur, s_item, slots);
    __syncthreads();  // the tile is read out and zero again
  }
}

// G2P of one substep fused with the P2G of the next (when no sort comes between them, the work
// items stay valid): one workgroup per work item. The item's grid-velocity box (its tile box of this
// substep's P2G, when it has at most kFuseBoxNodes nodes) is staged in LDS for the gathers; each
// particle's new record is written out (to slot j of dst when sorted_out, the sort's reorder, else
// in place) and, still in registers, scattered into an LDS tile covering the item's whole window
// (bases within kDrift cells of the block, 12^3 nodes), or through P2G's fallback when its new base
// left the window. The tile is stored over the bounding box of the new in-window stencils in arena
// space, which becomes the item's tile for the next grid update; if the arena has no room, the box
// is added to A with atomics instead (integer sums, the same result). perm: sorted position ->
// record of src (nullptr: identity). keys_out (optional): the sort key of each new record, at its
// slot (the keys of a sort that follows this substep).
//
// The box is stored with fixed strides when it fits (x stride kBoxRow = 12, plane stride kBoxPlane =
// 96, so at most 12 x 8 x 8 nodes): the 27 gather offsets are then immediates, and since 12 = 4 mod 8,
// the 4 x 4 cells of one layer of a block (the lanes that share an LDS cycle) read distinct banks.
// A box that does not fit is stored packed (runtime strides) when it has at most kFuseBoxNodes nodes.
constexpr int kFuseBoxNodes = 768;
constexpr int kBoxRow = 12, kBoxPlane = 96;
static_assert(kBoxPlane * 8 == kFuseBoxNodes, "the fixed-stride box holds 8 planes");
template <int SHAPE, int D, int kFuseThreads = kFuseThreadsOf[SHAPE]>
__global__ __launch_bounds__(kFuseThreadsOf[SHAPE]) __attribute__((amdgpu_waves_per_eu(kFuseWavesOf[SHAPE], 8))) void k_g2p_p2g(
    Pool pool, const float4* src, float4* dst, const uint32_t* __restrict__ perm, int sorted_out, float vmax,
    uint32_t* __restrict__ keys_out) {
  constexpr int S = Window<D>::kSpan, N = Window<D>::kNodes;  // window tile: S^3 nodes from the lowest in-window base
  __shared__ float4 vbox[kFuseBoxNodes];
  __shared__ int tile[4 * N];
  __shared__ int bbox[6];
  __shared__ uint32_t s_item, s_off;
  const int tid = threadIdx.x;
  const uint32_t items = pool.ctr[kCtrItems];
  const uint32_t epoch = pool.ctr[kCtrEpoch], bin_epoch = pool.ctr[kCtrBinEpoch];
  const uint32_t mshift = pool.key_bits - 6;
  const bool slots = (uint64_t)items * N <= pool.tile_nodes;  // every item's tile in its own slot
  // P2G's work counters (the launch_begin_p2g before this kernel cleared them; G2P's stay fresh for
  // the next substep's G2P, which has no launch_begin_p2g of its own before it)
  ItemCursor cur{pool.arena_ctr + kArenaRegions * kArenaStride, items, blockIdx.x % kWorkSplits, 0, false};
  const float4* __restrict__ V = pool.V;
  for (int t = tid; t < N; t += kFuseThreads)
    reinterpret_cast<int4*>(tile)[t] = make_int4(0, 0, 0, 0);  // the 4 planes; each item leaves them zero
  if (tid == 0) s_item = first_item(cur.items);
  __syncthreads();
  for (;;) {
    const uint32_t k = s_item;
    if (k >= items) break;
    const uint32_t start = pool.wl_start[k], cnt = pool.wl_start[k + 1] - start;
    const uint32_t wb = pool.wl_block[k];
    const bool has_box = pool.tile_off[k] != 0xffffffffu;
    const int4 h = pool.tile_hdr[k];  // meaningful when has_box; loaded beside tile_off
    // this thread's first record, loaded while the velocity box is staged (later ones during the
    // previous pass's P2G)
    uint32_t li = tid, si = 0;
    G2PIn in = {};
    if (li < cnt) {
      si = perm ? perm[start + li] : start + li;
      g2p_fetch(src, si, pool.cap, in);
    }
    const SceneGeo geo = uniform_geo(pool, wb >> mshift);
    const uint32_t mb = wb & ((1u << mshift) - 1u);
    const int wx = 4 * (int)compact3(mb) - D, wy = 4 * (int)compact3(mb >> 1) - D,
              wz = 4 * (int)compact3(mb >> 2) - D;  // lowest in-window base
    int ox = 0, oy = 0, oz = 0, ex = 0, ey = 0, ez = 0;
    if (has_box) {
      ox = h.x; oy = h.y; oz = h.z;
      ex = h.w & 255; ey = (h.w >> 8) & 255; ez = (h.w >> 16) & 255;
    }
    int nodes = ex * ey * ez;
    if (nodes > kFuseBoxNodes) nodes = 0;  // gathers from V
    const bool fixed = ex <= kBoxRow && ey * kBoxRow <= kBoxPlane && ez * kBoxPlane <= kFuseBoxNodes;
    const int sx = fixed ? kBoxRow : ex, sp = fixed ? kBoxPlane : ex * ey;  // strides of the stored box
    const BoxDiv dx = box_div((uint32_t)ex), dxy = box_div((uint32_t)(ex * ey));
    for (int t = tid; t < nodes; t += kFuseThreads) {
      int lx, ly, lz;
      box_coords(t, ex, ey, dx, dxy, lx, ly, lz);
      vbox[lx + sx * ly + sp * lz] = V[node_addr(ox + lx, oy + ly, oz + lz, geo)];
    }
    if (tid < 3) { bbox[tid] = INT_MAX; bbox[3 + tid] = INT_MIN; }
    __syncthreads();
    int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
    for (; li < cnt; li += kFuseThreads) {
      const uint32_t j = start + li;
      g2p_base(in);
      const bool inbox = nodes > 0 && (unsigned)(in.bx - ox) <= (unsigned)(ex - 3) &&
                         (unsigned)(in.by - oy) <= (unsigned)(ey - 3) && (unsigned)(in.bz - oz) <= (unsigned)(ez - 3);
      f3 v;
      m3 B;
      if (__builtin_amdgcn_ballot_w64(!inbox) == 0) {
        const int tb = (in.bz - oz) * sp + (in.by - oy) * sx + (in.bx - ox);
        if (fixed)
          g2p_gather(in, v, B, [&](int, int, int, int a, int b, int c) {
            return vbox[tb + c * kBoxPlane + b * kBoxRow + a];
          });
        else
          g2p_gather(in, v, B, [&](int, int, int, int a, int b, int c) { return vbox[tb + c * sp + b * sx + a]; });
      } else {
        g2p_gather(in, v, B,
                   [&](int bx, int by, int bz, int a, int b, int c) { return V[node_addr(bx + a, by + b, bz + c, geo)]; });
      }
      float4 r[8];
      g2p_finish(pool.mats, geo, src, si, dst, sorted_out ? j : si, pool.cap, vmax, in, v, B, r);
      if (keys_out) keys_out[sorted_out ? j : si] = sort_key(pool, wb >> mshift, geo.active, r[0]);
      // this thread's record of the next pass, loaded while this one's P2G runs (in and si are free now;
      // in place, its slot is another particle's, which only this thread reads or writes)
      if (li + kFuseThreads < cnt) {
        si = perm ? perm[j + kFuseThreads] : j + kFuseThreads;
        g2p_fetch(src, si, pool.cap, in);
      }
      // the next substep's P2G of the new record
      const int b0 = (int)floorf(r[0].x - 0.5f), b1 = (int)floorf(r[0].y - 0.5f), b2 = (int)floorf(r[0].z - 0.5f);
      const bool inwin = in_window<D>(b0, b1, b2, wx, wy, wz);
      if (inwin) {
        lo0 = min(lo0, b0); lo1 = min(lo1, b1); lo2 = min(lo2, b2);
        hi0 = max(hi0, b0); hi1 = max(hi1, b1); hi2 = max(hi2, b2);
        const int tb = ((b2 - wz) * S + (b1 - wy)) * S + (b0 - wx);
        auto to_tile = [&](int gx, int gy, int gz, int q0, int q1, int q2, int q3) {
          const int t = tb + ((gz - b2) * S + (gy - b1)) * S + (gx - b0);
          atomicAdd(&tile[t], q0);
          atomicAdd(&tile[N + t], q1);
          atomicAdd(&tile[2 * N + t], q2);
          atomicAdd(&tile[3 * N + t], q3);
        };
        // one wave-uniform choice: the clamp-free form when no active lane's particle can reach the clamp
        if (__builtin_amdgcn_ballot_w64(!p2g_unclamped(r, pool.mats, geo.vol4)) == 0)
          p2g_particle<false>(r, pool.mats, geo.vol4, to_tile);
        else
          p2g_particle<true>(r, pool.mats, geo.vol4, to_tile);
      } else {
        fallback_particle(pool, r, epoch, true, bin_epoch);
      }
    }
    // with slots, the next item is taken here (every thread read k before the box barrier), its round
    // trip hidden by the bounding box and the tile store
    if (slots && tid == 0) s_item = next_item(cur.wctr, cur.items, cur.range, cur.spent);
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    if ((tid & 63) == 0) {
      atomicMin(&bbox[0], lo0); atomicMin(&bbox[1], lo1); atomicMin(&bbox[2], lo2);
      atomicMax(&bbox[3], hi0); atomicMax(&bbox[4], hi1); atomicMax(&bbox[5], hi2);
    }
    __syncthreads();
    store_item_tile<kFuseThreads, D>(pool, k, wb, geo, tile, bbox, wx, wy, wz, epoch, bin_epoch, s_off, cur, s_item,
                                     slots);
    __syncthreads();  // the tile is read out and zero again
  }
}

// Global-atomic P2G (the reference's forceGlobal path, mpm.js:342-345), any particle order; every
// written node block goes to D.
__global__ __launch_bounds__(kThreads) void k_p2g_global(Pool pool, const float4* __restrict__ P, uint32_t n) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float4 r[kP2GStreams];
  load_p2g_record(P, pool.cap, i, r);
  fallback_particle(pool, r, pool.ctr[kCtrEpoch], false, 0);
}

struct Words64 { uint32_t w[64]; };
__global__ void k_write_words(Words64 src, uint32_t* dst, uint32_t n) {
  const uint32_t i = threadIdx.x;
  if (i < n) dst[i] = src.w[i];
}

constexpr int kPullBatch = 8;  // tile loads a grid-update wavefront keeps in flight
// Sum of the P2G result at scene-local node (n0, n1, n2) of node block (g0, g1, g2): the tiles of
// particle blocks G-2..G+1 per axis that cover the node. Lane L looks up candidate particle block L
// (its mtile slot: one load for the whole wave) and the run of its tile that holds the node block;
// the candidates whose tile box overlaps the node block are then visited with wave-uniform tile
// indices, up to kPullBatch tile loads (one contiguous run each) in flight at a time. Integer sums,
// so the visiting order does not matter.
template <int D>
__device__ __forceinline__ int4 pull_tiles(const Pool& pool, uint32_t scene, uint32_t epoch, int g0, int g1, int g2,
                                           int lane, int n0, int n1, int n2, int4 acc) {
  const int pb = 1 << (pool.key_bits / 3 - 2);  // particle blocks per axis a scene's slots address
  // lane L's candidate: the particle blocks whose drift-D tiles can reach node block g, [g-D+1, g+1]
  // per axis (4^3 = 64 candidates for D = 3, 3^3 = 27 for D = 2; the other lanes have none)
  constexpr int kSide = D + 1;
  const int cx = g0 - (D - 1) + lane % kSide, cy = g1 - (D - 1) + (lane / kSide) % kSide,
            cz = g2 - (D - 1) + lane / (kSide * kSide);
  uint4 e = make_uint4(0xffffffffu, 0u, 0u, 0u);
  if (lane < kSide * kSide * kSide && (unsigned)cx < (unsigned)pb && (unsigned)cy < (unsigned)pb &&
      (unsigned)cz < (unsigned)pb)
    e = pool.mtile[(scene << (pool.key_bits - 6)) + morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz)];
  const bool has = e.w == epoch && e.x != 0xffffffffu;  // a tile of this P2G
  // this lane's candidate tile: the chunk it holds for the wave's node block (tile_chunk), as the
  // run's first node in the arena and its corner and widths packed into one word each
  uint32_t run = 0, corner = 0, widths = 0;
  bool hit = false;
  if (has) {
    const int ex = (int)(e.z & 255u), ey = (int)((e.z >> 8) & 255u), ez = (int)((e.z >> 16) & 255u);
    uint32_t co;
    int wx, wy, wz, lo0, lo1, lo2;
    hit = tile_chunk((int)(e.y & 1023u), (int)((e.y >> 10) & 1023u), (int)(e.y >> 20), ex, ey, ez, g0, g1, g2, co, wx,
                     wy, wz, lo0, lo1, lo2);
    run = e.x + co;
    corner = (uint32_t)lo0 | ((uint32_t)lo1 << 10) | ((uint32_t)lo2 << 20);
    widths = (uint32_t)wx | ((uint32_t)wy << 4) | ((uint32_t)wz << 8);
  }
  uint64_t mask = __builtin_amdgcn_ballot_w64(hit);
  while (mask) {
    int4 q[kPullBatch];
#pragma unroll
    for (int u = 0; u < kPullBatch; u++) {
      q[u] = make_int4(0, 0, 0, 0);
      if (mask) {
        const int c = __builtin_ctzll(mask);
        mask &= mask - 1;
        const uint32_t rc = (uint32_t)__builtin_amdgcn_readlane((int)run, c);
        const uint32_t cc = (uint32_t)__builtin_amdgcn_readlane((int)corner, c);
        const uint32_t wc = (uint32_t)__builtin_amdgcn_readlane((int)widths, c);
        const int wx = (int)(wc & 15u), wy = (int)((wc >> 4) & 15u), wz = (int)(wc >> 8);
        const int lx = n0 - (int)(cc & 1023u), ly = n1 - (int)((cc >> 10) & 1023u), lz = n2 - (int)(cc >> 20);
        if ((unsigned)lx < (unsigned)wx && (unsigned)ly < (unsigned)wy && (unsigned)lz < (unsigned)wz)
          q[u] = pool.tiles[(size_t)rc + (uint32_t)(lx + wx * (ly + wy * lz))];
      }
    }
#pragma unroll
    for (int u = 0; u < kPullBatch; u++) {
      acc.x = (int)((uint32_t)acc.x + (uint32_t)q[u].x);
      acc.y = (int)((uint32_t)acc.y + (uint32_t)q[u].y);
      acc.z = (int)((uint32_t)acc.z + (uint32_t)q[u].z);
      acc.w = (int)((uint32_t)acc.w + (uint32_t)q[u].w);
    }
  }
  return acc;
}

// A float quantity as 2^32 fixed point in 64 bits: rint(x * 2^32) (the scaling is exact), clamped to
// +-2^62 first (NaN clamps to -2^62).
__device__ __forceinline__ unsigned long long law_fixed(float x) {
  const float y = fminf(fmaxf(x * 4294967296.0f, -4.611686018427387904e18f), 4.611686018427387904e18f);
  return (unsigned long long)(long long)rintf(y);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// The momentum bookkeeping of one node block (kLawWords, mpm_kernels.hpp): each lane's node terms,
// summed over the wave, added by lane 0 into the workgroup's slot. Called by every lane of the wave.
__device__ __forceinline__ void law_block(unsigned long long* __restrict__ law, int4 q, float4 vo, f3 vf,
                                          const float g[3], int lane) {
  const bool on = q.w > 0;
  unsigned long long t[14];
  t[0] = on ? (unsigned long long)(long long)q.w : 0ull;
  t[1] = on ? (unsigned long long)(long long)q.x : 0ull;
  t[2] = on ? (unsigned long long)(long long)q.y : 0ull;
  t[3] = on ? (unsigned long long)(long long)q.z : 0ull;
  const float m = vo.w;
  const float v[3] = {vo.x, vo.y, vo.z}, f[3] = {vf.x, vf.y, vf.z};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    t[4 + c] = on ? law_fixed(m * v[c]) : 0ull;
    t[7 + c] = on ? law_fixed(m * g[c]) : 0ull;
    t[10 + c] = on ? law_fixed(m * (v[c] - f[c])) : 0ull;
  }
  t[13] = on ? 1ull : 0ull;
  unsigned long long* slot = law + (size_t)(blockIdx.x % kLawSlots) * kLawWords;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const unsigned long long s = wave_sum_u64(t[k]);
    if (lane == 0 && s) atomicAdd(slot + k, s);
  }
}

// Grid update (VELOCITY = true, mpm.js:428-470) or tile resolve into A (VELOCITY = false) over S
// (when use_s) then D, one wavefront per node block.
template <bool VELOCITY, int D>
__global__ __launch_bounds__(kThreads) void k_grid_pull(Pool pool, int use_s) {
  const uint32_t ns = use_s ? pool.ctr[kCtrS] : 0u, nd = pool.ctr[kCtrD];
  const uint32_t epoch = pool.ctr[kCtrEpoch];
  if (VELOCITY && blockIdx.x == 0 && threadIdx.x == 0) {  // label V: the P2G epoch it comes from
    pool.ctr[kCtrVEpoch] = epoch;
    if (pool.law) atomicAdd(pool.law + 14, 1ull);  // one more grid update in the bookkeeping
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t j = blockIdx.x * 4 + w; j < ns + nd; j += gridDim.x * 4) {
    // an S entry carries its block's info; a D entry (a block written through A) looks it up
    const uint32_t b = __builtin_amdgcn_readfirstlane(j < ns ? pool.slist[j] : pool.dlist[j - ns]);
    const uint2 bi = j < ns ? pool.sinfo[j] : pool.block_info[b];
    const uint32_t scene = __builtin_amdgcn_readfirstlane(bi.x), bc = __builtin_amdgcn_readfirstlane(bi.y);
    const SceneGeo geo = uniform_geo(pool, scene);
    const int g0 = (int)(bc & 1023u), g1 = (int)((bc >> 10) & 1023u), g2 = (int)(bc >> 20);
    const int n0 = 4 * g0 + (lane & 3), n1 = 4 * g1 + ((lane >> 2) & 3), n2 = 4 * g2 + (lane >> 4);
    const uint32_t addr = (b << 6) | (uint32_t)lane;
    int4 acc = make_int4(0, 0, 0, 0);
    const bool written = pool.stamp[b] == epoch;
    if (written) acc = pool.A[addr];
    if (j < ns) acc = pull_tiles<D>(pool, scene, epoch, g0, g1, g2, lane, n0, n1, n2, acc);
    if (VELOCITY) {
      f3 vf;
      const float4 vo = grid_node(acc, n0, n1, n2, geo, pool.cols + 12 * kColRows * scene, vf);
      pool.V[addr] = vo;
      if (lane == 0) pool.vstamp[b] = epoch;  // V of this block is this grid update's (read by exports)
      if (written) pool.A[addr] = make_int4(0, 0, 0, 0);
      if (pool.law) law_block(pool.law, acc, vo, vf, geo.g, lane);
    } else {
      pool.A[addr] = acc;
    }
  }
}

// s_stamped: an S block is cleared only if the last P2G wrote it through A (stamp = its epoch): a
// binned P2G's fallback writes to blocks of S go to A without a D entry.
__global__ __launch_bounds__(kThreads) void k_clear_a(Pool pool, int use_s, int s_stamped) {
  const uint32_t ns = use_s ? pool.ctr[kCtrS] : 0u, nd = pool.ctr[kCtrD], epoch = pool.ctr[kCtrEpoch];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t j = blockIdx.x * 4 + w; j < ns + nd; j += gridDim.x * 4) {
    const uint32_t b = j < ns ? pool.slist[j] : pool.dlist[j - ns];
    if (j < ns && s_stamped && pool.stamp[b] != epoch) continue;
    pool.A[(b << 6) | (uint32_t)lane] = make_int4(0, 0, 0, 0);
  }
}

__global__ void k_grid_export(Pool pool, SceneGeo geo, int mode, int4* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  if (i >= nodes) return;
  const int gx = (int)(i % (size_t)geo.X), gy = (int)((i / (size_t)geo.X) % (size_t)geo.Y),
            gz = (int)(i / ((size_t)geo.X * geo.Y));
  const uint32_t addr = node_addr(gx, gy, gz, geo);
  const uint32_t b = addr >> 6;
  if (mode == 0) {
    out[i] = pool.A[addr];
  } else if (pool.vstamp[b] == pool.ctr[kCtrVEpoch]) {  // written by the last grid update
    const float4 v = pool.V[addr];
    out[i] = make_int4(__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w));
  } else {
    out[i] = make_int4(0, 0, 0, 0);
  }
}

__global__ void k_points_out(const float4* __restrict__ P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base,
                             double o0, double o1, double o2, double dx, double vscale, float4* __restrict__ pts,
                             float4* __restrict__ vel) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = P[offset + i];
  pts[i] = make_float4((float)((double)x.x * dx + o0), (float)((double)x.y * dx + o1), (float)((double)x.z * dx + o2),
                       __uint_as_float(__float_as_uint(x.w) - row_base));
  if (vel) {
    const float4 v = P[(size_t)cap + offset + i];
    const float J = P[(size_t)5 * cap + offset + i].z;
    vel[i] = make_float4((float)((double)v.x * vscale), (float)((double)v.y * vscale), (float)((double)v.z * vscale), J);
  }
}

__global__ void k_grid_import(Pool pool, SceneGeo geo, const int4* __restrict__ in) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  if (i >= nodes) return;
  const int gx = (int)(i % (size_t)geo.X), gy = (int)((i / (size_t)geo.X) % (size_t)geo.Y),
            gz = (int)(i / ((size_t)geo.X * geo.Y));
  const int4 w = in[i];
  pool.V[node_addr(gx, gy, gz, geo)] =
      make_float4(__int_as_float(w.x), __int_as_float(w.y), __int_as_float(w.z), __int_as_float(w.w));
}

inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }
inline unsigned persistent_grid(uint32_t max_work, unsigned per_block, unsigned cap) {
  unsigned g = (unsigned)((max_work + per_block - 1) / per_block);
  if (g > cap) g = cap;
  return g ? g : 1u;
}

}  // namespace

hipError_t launch_records_in(const float4* aos, float4* P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base,
                             hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_records_in, dim3(blocks_for(n, 256)), dim3(256), 0, s, aos, P, cap, offset, n, row_base);
  return hipGetLastError();
}
hipError_t launch_records_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base, float4* aos,
                              hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_records_out, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, cap, offset, n, row_base,
                     aos);
  return hipGetLastError();
}
hipError_t launch_rescale(float4* P, uint32_t cap, uint32_t offset, uint32_t n, float k, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_rescale, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, cap, offset, n, k);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}
hipError_t launch_block_info(uint2* info, uint32_t base, uint32_t n, uint32_t scene, uint32_t nbx, uint32_t nby,
                             hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_block_info, dim3(blocks_for(n, 256)), dim3(256), 0, s, info, base, n, scene, nbx, nby);
  return hipGetLastError();
}
hipError_t launch_begin_p2g(const Pool& pool, hipStream_t s) {
  hipLaunchKernelGGL(k_begin_p2g, dim3(1), dim3(kArenaRegions), 0, s, pool.ctr, pool.arena_ctr);
  return hipGetLastError();
}
hipError_t launch_begin_bin(const Pool& pool, hipStream_t s) {
  hipLaunchKernelGGL(k_begin_bin, dim3(1), dim3(1), 0, s, pool.ctr);
  return hipGetLastError();
}
hipError_t launch_bin_reset(const Pool& pool, uint32_t occ_words32, hipStream_t s) {
  if (occ_words32 == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bin_reset, dim3(blocks_for(occ_words32, 256)), dim3(256), 0, s, pool, occ_words32);
  return hipGetLastError();
}
hipError_t launch_keys(const Pool& pool, const float4* P, uint32_t n, uint32_t* keys, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, P, n, keys);
  return hipGetLastError();
}
hipError_t launch_live_keys(const Pool& pool, const float4* P, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_live_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, P, n, keys, vals);
  return hipGetLastError();
}
hipError_t launch_permute(const float4* src, float4* dst, uint32_t cap, uint32_t n, const uint32_t* perm, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_permute, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, dst, cap, n, perm);
  return hipGetLastError();
}
uint32_t bin_tiles(uint32_t n) { return (uint32_t)(((size_t)n + kItemTile - 1) / kItemTile); }
hipError_t launch_bin_count(const uint32_t* keys, uint32_t n, uint32_t* tile_heads, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_bin_count, dim3(bin_tiles(n)), dim3(kThreads), 0, s, keys, n, tile_heads);
  return hipGetLastError();
}
bool bin_emit_scans(uint32_t n) { return bin_tiles(n) <= kItemScanFree; }
hipError_t launch_bin_emit(const Pool& pool, const uint32_t* keys, uint32_t n, const uint32_t* tile_first, bool scanned,
                           hipStream_t s) {
  if (!n) return hipSuccess;
  if (!scanned && !bin_emit_scans(n)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bin_emit, dim3(bin_tiles(n)), dim3(kThreads), 0, s, pool, keys, n, tile_first, bin_tiles(n),
                     scanned ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_bin_transpose(const Pool& pool, const uint32_t* keys, const uint32_t* in, uint32_t* out, hipStream_t s) {
  const unsigned grid = (persistent_grid(pool.max_items, 1, 4096) + 7u) / 8u * 8u;  // a multiple of 8
  hipLaunchKernelGGL(k_bin_transpose, dim3(grid), dim3(kThreads), 0, s, pool, keys, in, out);
  return hipGetLastError();
}
hipError_t launch_bin_dilate(const Pool& pool, uint32_t max_scenes, int drift, hipStream_t s) {
  const uint32_t pb = 1u << (pool.key_bits / 3 - 2);
  const uint64_t total = (uint64_t)max_scenes * pb * pb * occ_row_words(pb) * kDilateSplit;
  hipLaunchKernelGGL(k_bin_dilate, dim3((unsigned)std::min<uint64_t>((total + kThreads - 1) / kThreads, 8192)),
                     dim3(kThreads), 0, s, pool, max_scenes, drift);
  return hipGetLastError();
}
template <int D>
void launch_p2g_binned_d(const Pool& pool, const float4* P, const uint32_t* perm, int waves, hipStream_t s,
                         uint32_t* keys_out) {
  const dim3 grid((persistent_grid(pool.max_items, 1, 2048) + 7u) / 8u * 8u);  // a multiple of 8 (first_item)
  switch (waves) {
    case 5: hipLaunchKernelGGL((k_p2g_binned<5, D>), grid, dim3(kThreads), 0, s, pool, P, perm, keys_out); break;
    case 6: hipLaunchKernelGGL((k_p2g_binned<6, D>), grid, dim3(kThreads), 0, s, pool, P, perm, keys_out); break;
    default: hipLaunchKernelGGL((k_p2g_binned<4, D>), grid, dim3(kThreads), 0, s, pool, P, perm, keys_out); break;
  }
}
hipError_t launch_p2g_binned(const Pool& pool, const float4* P, const uint32_t* perm, int waves, int drift,
                             hipStream_t s, uint32_t* keys_out) {
  if (drift == 2) launch_p2g_binned_d<2>(pool, P, perm, waves, s, keys_out);
  else launch_p2g_binned_d<3>(pool, P, perm, waves, s, keys_out);
  return hipGetLastError();
}
hipError_t launch_p2g_global(const Pool& pool, const float4* P, uint32_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_p2g_global, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, pool, P, n);
  return hipGetLastError();
}
hipError_t launch_write_words(uint32_t* dst, const uint32_t* src, uint32_t n, hipStream_t s) {
  if (n > 64) return hipErrorInvalidValue;
  Words64 w;
  for (uint32_t i = 0; i < 64; i++) w.w[i] = i < n ? src[i] : 0u;
  hipLaunchKernelGGL(k_write_words, dim3(1), dim3(64), 0, s, w, dst, n);
  return hipGetLastError();
}
hipError_t launch_grid_update(const Pool& pool, bool use_s, int drift, hipStream_t s) {
  const dim3 grid(persistent_grid(2 * pool.nblocks, 4, 8192));
  if (drift == 2) hipLaunchKernelGGL((k_grid_pull<true, 2>), grid, dim3(kThreads), 0, s, pool, use_s ? 1 : 0);
  else hipLaunchKernelGGL((k_grid_pull<true, 3>), grid, dim3(kThreads), 0, s, pool, use_s ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_resolve_tiles(const Pool& pool, int drift, hipStream_t s) {
  const dim3 grid(persistent_grid(2 * pool.nblocks, 4, 8192));
  if (drift == 2) hipLaunchKernelGGL((k_grid_pull<false, 2>), grid, dim3(kThreads), 0, s, pool, 1);
  else hipLaunchKernelGGL((k_grid_pull<false, 3>), grid, dim3(kThreads), 0, s, pool, 1);
  return hipGetLastError();
}
hipError_t launch_clear_a(const Pool& pool, bool use_s, bool s_stamped, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_a, dim3(persistent_grid(2 * pool.nblocks, 4, 8192)), dim3(kThreads), 0, s, pool,
                     use_s ? 1 : 0, s_stamped ? 1 : 0);
  return hipGetLastError();
}
// G2P launches; with ev_start / ev_stop (either may be null) the dispatch itself records them
// (hipExtLaunchKernelGGL), so timing a launch adds no packet to the stream.
#define MPM_G2P_LAUNCH_T(KERNEL, GRID, THREADS, S, EV0, EV1, ...)                                   \
  do {                                                                                              \
    if ((EV0) || (EV1))                                                                             \
      hipExtLaunchKernelGGL(KERNEL, GRID, dim3(THREADS), 0, S, EV0, EV1, 0, __VA_ARGS__);          \
    else                                                                                            \
      hipLaunchKernelGGL(KERNEL, GRID, dim3(THREADS), 0, S, __VA_ARGS__);                           \
  } while (0)
#define MPM_G2P_LAUNCH(KERNEL, GRID, S, EV0, EV1, ...) MPM_G2P_LAUNCH_T(KERNEL, GRID, kThreads, S, EV0, EV1, __VA_ARGS__)

hipError_t launch_g2p_perm(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, uint32_t n, float vmax,
                           int waves, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (!n) return hipSuccess;
  const dim3 grid(blocks_for(n, kThreads));
  switch (waves) {
    case 4: MPM_G2P_LAUNCH(k_g2p_perm<4>, grid, s, ev_start, ev_stop, pool, src, dst, perm, n, vmax); break;
    case 5: MPM_G2P_LAUNCH(k_g2p_perm<5>, grid, s, ev_start, ev_stop, pool, src, dst, perm, n, vmax); break;
    case 6: MPM_G2P_LAUNCH(k_g2p_perm<6>, grid, s, ev_start, ev_stop, pool, src, dst, perm, n, vmax); break;
    default: MPM_G2P_LAUNCH(k_g2p_perm<3>, grid, s, ev_start, ev_stop, pool, src, dst, perm, n, vmax); break;
  }
  return hipGetLastError();
}
hipError_t launch_g2p_items(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, bool sorted_out,
                            float vmax, int waves, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop,
                            uint32_t* keys_out) {
  if (sorted_out ? (!perm || src == dst) : src != dst) return hipErrorInvalidValue;
  const dim3 grid((persistent_grid(pool.max_items, 1, 2048) + 7u) / 8u * 8u);  // a multiple of 8 (first_item)
  const int so = sorted_out ? 1 : 0;
  switch (waves) {
    case 4: MPM_G2P_LAUNCH(k_g2p_items<4>, grid, s, ev_start, ev_stop, pool, src, dst, perm, so, vmax, keys_out); break;
    case 5: MPM_G2P_LAUNCH(k_g2p_items<5>, grid, s, ev_start, ev_stop, pool, src, dst, perm, so, vmax, keys_out); break;
    case 6: MPM_G2P_LAUNCH(k_g2p_items<6>, grid, s, ev_start, ev_stop, pool, src, dst, perm, so, vmax, keys_out); break;
    default: MPM_G2P_LAUNCH(k_g2p_items<3>, grid, s, ev_start, ev_stop, pool, src, dst, perm, so, vmax, keys_out); break;
  }
  return hipGetLastError();
}
template <int D>
void launch_g2p_p2g_d(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, int so, float vmax,
                      int shape, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop, uint32_t* keys_out) {
  const dim3 grid((persistent_grid(pool.max_items, 1, 2048) + 7u) / 8u * 8u);  // a multiple of 8 (first_item)
  switch (shape) {
    case 0:
      MPM_G2P_LAUNCH_T((k_g2p_p2g<0, D>), grid, kFuseThreadsOf[0], s, ev_start, ev_stop, pool, src, dst, perm, so, vmax,
                       keys_out);
      break;
    default:
      MPM_G2P_LAUNCH_T((k_g2p_p2g<1, D>), grid, kFuseThreadsOf[1], s, ev_start, ev_stop, pool, src, dst, perm, so, vmax,
                       keys_out);
      break;
  }
}
hipError_t launch_g2p_p2g(const Pool& pool, const float4* src, float4* dst, const uint32_t* perm, bool sorted_out,
                          float vmax, int shape, int drift, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop,
                          uint32_t* keys_out) {
  if (sorted_out ? (!perm || src == dst) : src != dst) return hipErrorInvalidValue;
  const int so = sorted_out ? 1 : 0;
  if (drift == 2) launch_g2p_p2g_d<2>(pool, src, dst, perm, so, vmax, shape, s, ev_start, ev_stop, keys_out);
  else launch_g2p_p2g_d<3>(pool, src, dst, perm, so, vmax, shape, s, ev_start, ev_stop, keys_out);
  return hipGetLastError();
}
hipError_t launch_g2p(const Pool& pool, float4* P, uint32_t n, float vmax, int waves, hipStream_t s, hipEvent_t ev_start,
                      hipEvent_t ev_stop) {
  if (!n) return hipSuccess;
  const dim3 grid(blocks_for(n, kThreads));
  switch (waves) {
    case 4: MPM_G2P_LAUNCH(k_g2p<4>, grid, s, ev_start, ev_stop, pool, P, n, vmax); break;
    case 5: MPM_G2P_LAUNCH(k_g2p<5>, grid, s, ev_start, ev_stop, pool, P, n, vmax); break;
    case 6: MPM_G2P_LAUNCH(k_g2p<6>, grid, s, ev_start, ev_stop, pool, P, n, vmax); break;
    default: MPM_G2P_LAUNCH(k_g2p<3>, grid, s, ev_start, ev_stop, pool, P, n, vmax); break;
  }
  return hipGetLastError();
}
hipError_t launch_points_out(const float4* P, uint32_t cap, uint32_t offset, uint32_t n, uint32_t row_base,
                             const double origin[3], double dx, double vscale, float4* pts, float4* vel, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_points_out, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, cap, offset, n, row_base, origin[0],
                     origin[1], origin[2], dx, vscale, pts, vel);
  return hipGetLastError();
}
hipError_t launch_grid_import(const Pool& pool, const SceneGeo& geo, const int4* in, hipStream_t s) {
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  hipLaunchKernelGGL(k_grid_import, dim3(blocks_for(nodes, 256)), dim3(256), 0, s, pool, geo, in);
  return hipGetLastError();
}
hipError_t launch_grid_export(const Pool& pool, const SceneGeo& geo, int mode, int4* out, hipStream_t s) {
  const size_t nodes = (size_t)geo.X * geo.Y * geo.Z;
  hipLaunchKernelGGL(k_grid_export, dim3(blocks_for(nodes, 256)), dim3(256), 0, s, pool, geo, mode, out);
  return hipGetLastError();
}

}  // namespace mpm
