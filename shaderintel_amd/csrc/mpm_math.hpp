// Device math for the MLS-MPM kernels (gfx950).
//
// Every expression is evaluated in a fixed, documented order so that the GPU result is a pure
// function of the input bits. The compiler contracts nothing (-ffp-contract=off); the fused
// multiply-adds WGSL lets an implementation form are written out as fmaf where this build forms
// them: mat*vec ((m0*v.x + m1*v.y) + m2*v.z) and dot ((a.x*b.x + a.y*b.y) + a.z*b.z) with both
// additions fused, cross products as fma(a, b, -(c*d)), and G2P's v += wt*gv and
// B += outer(wt*gv, dpos) element by element. IEEE division and square root (HIP's default correctly
// rounded fp32 div/sqrt), WGSL clamp(e, lo, hi) = min(max(e, lo), hi), WGSL round =
// round-half-to-even (rint). inverseSqrt (used on [1, 2] only) is rsqrt_1to2 below, and exp and
// log are the self-contained Cody-Waite routines below, so no vendor math library decides any bit
// of the result.
//
// The WGSL being restated is src/physics/mpm.js:64-78 (module `mpm`) and
// src/shaders/linalg.js:8-69 (module `linalg`).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

#define MPM_HD __host__ __device__ __forceinline__

struct f3 { float x, y, z; };
struct m3 { f3 c[3]; };  // column-major like WGSL mat3x3f: c[col], component = row

MPM_HD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
MPM_HD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
MPM_HD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
MPM_HD f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
MPM_HD f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
MPM_HD f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
MPM_HD float dot3(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
MPM_HD float len3(f3 a) { return sqrtf(dot3(a, a)); }
MPM_HD f3 norm3(f3 a) { return a / len3(a); }
MPM_HD f3 cross3(f3 a, f3 b) {
  return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
MPM_HD float clampf(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }

MPM_HD m3 mat(f3 a, f3 b, f3 c) { m3 m; m.c[0] = a; m.c[1] = b; m.c[2] = c; return m; }
MPM_HD m3 ident() { return mat(mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1)); }
MPM_HD m3 zero3() { return mat(mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)); }
MPM_HD f3 mul(const m3& m, f3 v) {
  return mk(fmaf(m.c[2].x, v.z, fmaf(m.c[1].x, v.y, m.c[0].x * v.x)),
            fmaf(m.c[2].y, v.z, fmaf(m.c[1].y, v.y, m.c[0].y * v.x)),
            fmaf(m.c[2].z, v.z, fmaf(m.c[1].z, v.y, m.c[0].z * v.x)));
}
MPM_HD m3 mul(const m3& a, const m3& b) { return mat(mul(a, b.c[0]), mul(a, b.c[1]), mul(a, b.c[2])); }
MPM_HD m3 operator+(const m3& a, const m3& b) { return mat(a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]); }
MPM_HD m3 operator-(const m3& a, const m3& b) { return mat(a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]); }
MPM_HD m3 operator*(float s, const m3& a) { return mat(s * a.c[0], s * a.c[1], s * a.c[2]); }
MPM_HD m3 tr(const m3& m) {
  return mat(mk(m.c[0].x, m.c[1].x, m.c[2].x), mk(m.c[0].y, m.c[1].y, m.c[2].y), mk(m.c[0].z, m.c[1].z, m.c[2].z));
}
MPM_HD float trace(const m3& m) { return (m.c[0].x + m.c[1].y) + m.c[2].z; }
MPM_HD m3 diag(f3 d) { return mat(mk(d.x, 0, 0), mk(0, d.y, 0), mk(0, 0, d.z)); }
MPM_HD m3 outer(f3 a, f3 b) { return mat(a * b.x, a * b.y, a * b.z); }
// B + outer(a, b), each element contracted: B[j][r] = fma(a[r], b[j], B[j][r]).
MPM_HD m3 add_outer(const m3& B, f3 a, f3 b) {
  return mat(mk(fmaf(a.x, b.x, B.c[0].x), fmaf(a.y, b.x, B.c[0].y), fmaf(a.z, b.x, B.c[0].z)),
             mk(fmaf(a.x, b.y, B.c[1].x), fmaf(a.y, b.y, B.c[1].y), fmaf(a.z, b.y, B.c[1].z)),
             mk(fmaf(a.x, b.z, B.c[2].x), fmaf(a.y, b.z, B.c[2].y), fmaf(a.z, b.z, B.c[2].z)));
}
MPM_HD float det3(const m3& m) { return dot3(m.c[0], cross3(m.c[1], m.c[2])); }

// ---- elementary functions (Cody & Waite 1980 reduction + Taylor / atanh series in float) ----
MPM_HD float bits2f(uint32_t u) { return __builtin_bit_cast(float, u); }
MPM_HD uint32_t f2bits(float f) { return __builtin_bit_cast(uint32_t, f); }

MPM_HD float exp_cw(float x) {
  const float kLn2Hi = 0.693145751953125f;         // 16 significant bits: k * kLn2Hi exact
  const float kLn2Lo = 1.428606765330187045e-06f;
  const float kLog2e = 1.44269504088896340736f;
  if (x != x) return x;
  if (x > 88.7f) return bits2f(0x7f800000u);
  if (x < -103.9f) return 0.0f;
  float kf = rintf(x * kLog2e);
  int k = (int)kf;
  float r = (x - kf * kLn2Hi) - kf * kLn2Lo;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int k1 = k / 2, k2 = k - k1;
  return (p * bits2f((uint32_t)(k1 + 127) << 23)) * bits2f((uint32_t)(k2 + 127) << 23);
}

MPM_HD float log_cw(float x) {
  const float kLn2Hi = 0.693145751953125f;
  const float kLn2Lo = 1.428606765330187045e-06f;
  if (x != x) return x;
  if (x < 0.0f) return bits2f(0x7fc00000u);
  if (x == 0.0f) return bits2f(0xff800000u);
  if (x == bits2f(0x7f800000u)) return x;
  int e = 0;
  uint32_t u = f2bits(x);
  if ((u >> 23) == 0) { x = x * 16777216.0f; u = f2bits(x); e = -24; }
  e += (int)(u >> 23) - 127;
  float m = bits2f((u & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float f = m - 1.0f;
  float s = f / (2.0f + f);
  float s2 = s * s;
  float p = 1.0f / 11.0f;
  p = p * s2 + 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  p = p * s2 + 1.0f;
  float lm = (2.0f * s) * p;
  float ef = (float)e;
  return ef * kLn2Hi + (lm + ef * kLn2Lo);
}

// ---- mpm module: fixed point, B-spline, Morton ----
constexpr float kFP = 4194304.0f;           // 2^22
constexpr float kInvFP = 1.0f / 4194304.0f;

MPM_HD int32_t to_fixed(float x) { return (int32_t)rintf(clampf(x * kFP, -2.0e9f, 2.0e9f)); }
// to_fixed of a value already scaled by 2^22
MPM_HD int32_t fixed_of_scaled(float xs) { return (int32_t)rintf(clampf(xs, -2.0e9f, 2.0e9f)); }

MPM_HD void bspline(f3 fx, f3 w[3]) {
  f3 t0 = mk(1.5f, 1.5f, 1.5f) - fx;
  f3 t1 = fx - mk(1.0f, 1.0f, 1.0f);
  f3 t2 = fx - mk(0.5f, 0.5f, 0.5f);
  w[0] = mk((0.5f * t0.x) * t0.x, (0.5f * t0.y) * t0.y, (0.5f * t0.z) * t0.z);
  w[1] = mk(0.75f - t1.x * t1.x, 0.75f - t1.y * t1.y, 0.75f - t1.z * t1.z);
  w[2] = mk((0.5f * t2.x) * t2.x, (0.5f * t2.y) * t2.y, (0.5f * t2.z) * t2.z);
}

MPM_HD uint32_t spread3(uint32_t x) {  // 10-bit interleave helper of morton3
  uint32_t v = x & 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
MPM_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) { return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2); }

// WGSL inverseSqrt(x) for x in [1, 2] (the Jacobi rotation's t^2 + 1 with |t| <= 1): a quadratic
// approximation of x^-1/2 on [1, 2] refined by two Newton steps y += y (1/2 - (x/2) y^2), every
// step a fused multiply-add. Within 1 ulp of 1/sqrt(x) over [1, 2] (WGSL allows 2 ulp); NaN in,
// NaN out.
MPM_HD float rsqrt_1to2(float x) {
  float y = fmaf(fmaf(0.14768398f, x, -0.7304885f), x, 1.5796084f);
  const float hx = 0.5f * x;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const float t = y * y;
    const float e = fmaf(-hx, t, 0.5f);
    y = fmaf(y, e, y);
  }
  return y;
}

// sqrt(x) for x >= 1 (or +inf), correctly rounded as sqrtf: the hardware square root (within 1 ulp)
// moved to the neighbour its residual points to. It is the compiler's own sequence for sqrtf without
// the rescaling of small inputs and the select of +-0 and +inf, which x >= 1 does not need (+inf
// passes through: both residuals are NaN). Checked bit for bit against sqrtf over every float in
// [1, 4), [2^62, 2^64) and [2^126, 2^128) and at +inf (tests/gpu_math_check.hip).
__device__ __forceinline__ float sqrt_ge1(float x) {
  const float y = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(y) - 1u), up = __uint_as_float(__float_as_uint(y) + 1u);
  const float rdn = fmaf(-dn, y, x), rup = fmaf(-up, y, x);
  const float r = rdn <= 0.0f ? dn : y;
  return rup > 0.0f ? up : r;
}

// 1 / d for d >= 1, correctly rounded as 1.0f / d: the hardware reciprocal and one Newton step
// (r1 = fma(fma(-d, r0, 1), r0, r0)) below 2^126, where the result is normal; IEEE division above
// (denormal results, +inf). Checked bit for bit against 1.0f / d over every float in [1, 2), [2^20,
// 2^21), [2^60, 2^61), [2^125, 2^128) and at +inf (tests/gpu_math_check.hip).
__device__ __forceinline__ float rcp_ge1(float d) {
  if (d < 0x1p126f) {
    const float r0 = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, r0, 1.0f), r0, r0);
  }
  return 1.0f / d;
}

// ---- 3x3 symmetric eigen-decomposition and SVD (cyclic Jacobi, 4 sweeps) ----
MPM_HD float& el(m3& m, int c, int r) { return r == 0 ? m.c[c].x : (r == 1 ? m.c[c].y : m.c[c].z); }
MPM_HD float el(const m3& m, int c, int r) { return r == 0 ? m.c[c].x : (r == 1 ? m.c[c].y : m.c[c].z); }

// One cyclic-Jacobi rotation in the (P, Q) plane: A <- J^T A J, V <- V J. Only rows / columns P and
// Q change, each as the two-term sum the full 3x3 products would form (mul above: the first nonzero
// product rounded, the second fused onto it; their other terms are multiples of zero). A lane whose
// |a_pq| < 1e-30 keeps A and V as they are (the rotation is skipped). Every lane computes the
// rotation (a skipping lane's values, possibly inf or NaN, are discarded), and only a wavefront with
// a skipping lane selects the old values back: a branch around the rotation would end in a join
// where the new values are copied into the old ones' registers (10 moves per rotation).
template <int P, int Q>
__device__ __forceinline__ void jacobi(m3& A, m3& V) {
  const m3 A0 = A, V0 = V;
  float apq = el(A, Q, P);
  const bool skip = fabsf(apq) < 1e-30f;
  float theta = (el(A, Q, Q) - el(A, P, P)) / (2.0f * apq);
  float sgn = theta >= 0.0f ? 1.0f : -1.0f;
  // t = sgn / d with d = |theta| + sqrt(theta^2 + 1) >= 1; (+-1) / d = +-(1 / d) exactly
  float t = sgn * rcp_ge1(fabsf(theta) + sqrt_ge1(theta * theta + 1.0f));
  float c = rsqrt_1to2(t * t + 1.0f);  // inverseSqrt (linalg.js:20); |t| <= 1
  float s = t * c;
  float ms = -s;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const float ap = el(A, j, P), aq = el(A, j, Q);
    el(A, j, P) = fmaf(ms, aq, c * ap);
    el(A, j, Q) = fmaf(c, aq, s * ap);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float mp = el(A, P, i), mq = el(A, Q, i);
    el(A, P, i) = fmaf(mq, ms, mp * c);
    el(A, Q, i) = fmaf(mq, c, mp * s);
    const float vp = el(V, P, i), vq = el(V, Q, i);
    el(V, P, i) = fmaf(vq, ms, vp * c);
    el(V, Q, i) = fmaf(vq, c, vp * s);
  }
  if (__builtin_amdgcn_ballot_w64(skip) != 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      A.c[k] = mk(skip ? A0.c[k].x : A.c[k].x, skip ? A0.c[k].y : A.c[k].y, skip ? A0.c[k].z : A.c[k].z);
      V.c[k] = mk(skip ? V0.c[k].x : V.c[k].x, skip ? V0.c[k].y : V.c[k].y, skip ? V0.c[k].z : V.c[k].z);
    }
  }
}

__device__ __forceinline__ void svd3(const m3& F, m3& U, f3& s, m3& V) {
  m3 A = mul(tr(F), F);
  V = ident();
#pragma unroll  // the 12 rotations unrolled: no loop-carried register moves
  for (int sweep = 0; sweep < 4; sweep++) {
    jacobi<0, 1>(A, V);
    jacobi<0, 2>(A, V);
    jacobi<1, 2>(A, V);
  }
  f3 e = mk(A.c[0].x, A.c[1].y, A.c[2].z);
  if (e.x < e.y) { e = mk(e.y, e.x, e.z); V = mat(V.c[1], V.c[0], -1.0f * V.c[2]); }
  if (e.x < e.z) { e = mk(e.z, e.y, e.x); V = mat(V.c[2], -1.0f * V.c[1], V.c[0]); }
  if (e.y < e.z) { e = mk(e.x, e.z, e.y); V = mat(-1.0f * V.c[0], V.c[2], V.c[1]); }
  if (det3(V) < 0.0f) V.c[2] = -1.0f * V.c[2];
  s = mk(sqrtf(fmaxf(e.x, 0.0f)), sqrtf(fmaxf(e.y, 0.0f)), sqrtf(fmaxf(e.z, 0.0f)));
  f3 u0 = mul(F, V.c[0]);
  float l0 = len3(u0);
  u0 = l0 > 1e-12f ? u0 / l0 : mk(1, 0, 0);
  f3 u1 = mul(F, V.c[1]);
  u1 = u1 - dot3(u1, u0) * u0;
  float l1 = len3(u1);
  if (l1 > 1e-12f) {
    u1 = u1 / l1;
  } else {
    u1 = fabsf(u0.x) < 0.9f ? mk(1, 0, 0) : mk(0, 1, 0);
    u1 = norm3(u1 - dot3(u1, u0) * u0);
  }
  f3 u2 = cross3(u0, u1);
  if (dot3(mul(F, V.c[2]), u2) < 0.0f) s.z = -s.z;
  U = mat(u0, u1, u2);
}

}  // namespace mpm
