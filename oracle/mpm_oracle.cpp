// ORACLE — test infrastructure only.
//
// A plain C++17 CPU restatement of ShaderIntel's MLS-MPM substep (src/physics/mpm.js) and of the
// counting-sort / exclusive-scan primitives it depends on (src/prims/scan.js, src/prims/grid.js).
// Only the test suite, the smoke check and the CPU-baseline leg of the benchmark load this file
// (as liboracle_mpm.so through ctypes). The product library never links, includes or calls it.
//
// WGSL leaves FMA contraction and the association of mat*vec / mat*mat to the browser's compiler,
// so "the reference's bits" are adapter-defined. This file fixes ONE faithful reading of each
// expression: left-to-right evaluation, mat*vec as ((m0*v.x + m1*v.y) + m2*v.z) and dot as
// ((a.x*b.x + a.y*b.y) + a.z*b.z) with both additions contracted to fused multiply-adds
// (fma(m2, v.z, fma(m1, v.y, m0*v.x))), cross products as fma(a, b, -(c*d)), G2P's v += wt*gv and
// B += outer(wt*gv, dpos) contracted element by element, and no other contraction (explicit fmaf
// calls; the compiler contracts nothing: -ffp-contract=off), IEEE division and sqrt,
// P2G's per-node momentum exactly as mpm.js:330-335 writes it (tests/test_oracle_literal.py checks
// every i32 word against an independent node-by-node numpy reading), G2P's 27-node sums taken per
// row of three nodes (within the a priori rounding bound of the node-by-node readings; same test),
// WGSL clamp(e,lo,hi) = min(max(e,lo),hi) with NaN-discarding min/max, WGSL round = rint (half to
// even). inverseSqrt is rsqrt_1to2 (it is used on [1, 2] only); log/exp are the self-contained
// routines below.
//
// Every function here has a pinning test under tests/ (closed forms, invariants, brute force or
// numpy references); tests/test_oracle_*.py name the function each test pins.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------------------------------------
// Elementary functions. Written from the textbook method of Cody & Waite, "Software Manual for the
// Elementary Functions" (1980): argument reduction by a power of two and ln 2 split in two parts,
// then a truncated Taylor / atanh series evaluated by Horner's rule in float. All steps are
// +, *, /, rint and exponent-field manipulation, so they give identical bits on any IEEE-754 host.
// Accuracy (checked by tests/test_oracle_math.py against float64 numpy): a few ulp.

inline float f_from_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t bits_from_f(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// ln 2 split: LN2_HI has 16 significant bits, so k * LN2_HI is exact for |k| < 2^8.
constexpr float LN2_HI = 0.693145751953125f;          // 0x3f317200
constexpr float LN2_LO = 1.428606765330187045e-06f;   // ln2 - LN2_HI
constexpr float LOG2E = 1.44269504088896340736f;

float si_expf(float x) {
  if (x != x) return x;
  if (x > 88.7f) return f_from_bits(0x7f800000u);  // overflow -> +inf
  if (x < -103.9f) return 0.0f;
  float kf = rintf(x * LOG2E);
  int k = (int)kf;
  float r = (x - kf * LN2_HI) - kf * LN2_LO;       // |r| <= ~0.347
  // Taylor series of e^r to degree 7 (remainder < 5e-9 relative).
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  // scale by 2^k in two exact steps so that k in [-150, 128] never overflows the exponent field
  int k1 = k / 2, k2 = k - k1;
  float s1 = f_from_bits((uint32_t)(k1 + 127) << 23);
  float s2 = f_from_bits((uint32_t)(k2 + 127) << 23);
  return (p * s1) * s2;
}

float si_logf(float x) {
  if (x != x) return x;
  if (x < 0.0f) return f_from_bits(0x7fc00000u);   // NaN
  if (x == 0.0f) return f_from_bits(0xff800000u);  // -inf
  if (x == f_from_bits(0x7f800000u)) return x;
  int e = 0;
  uint32_t u = bits_from_f(x);
  if ((u >> 23) == 0) { x = x * 16777216.0f; u = bits_from_f(x); e = -24; }  // subnormal
  e += (int)(u >> 23) - 127;
  float m = f_from_bits((u & 0x007fffffu) | 0x3f800000u);   // m in [1, 2)
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }             // m in (0.7071, 1.4142]
  float f = m - 1.0f;
  float s = f / (2.0f + f);                                   // |s| <= 0.1716
  float s2 = s * s;
  // log(m) = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ... + s^11/11)
  float p = 1.0f / 11.0f;
  p = p * s2 + 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  p = p * s2 + 1.0f;
  float lm = (2.0f * s) * p;
  float ef = (float)e;
  return ef * LN2_HI + (lm + ef * LN2_LO);
}

// ------------------------------------------------------------------------------------------------
// vec3 / mat3 with WGSL semantics: mat3x3f is column-major, M[c] is column c, M[c][r] row r.
struct V3 { float x, y, z; };
struct M3 { V3 c[3]; };

inline float comp(const V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline float& comp(V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline float at(const M3& m, int c, int r) { return comp(m.c[c], r); }
inline float& at(M3& m, int c, int r) { return comp(m.c[c], r); }

inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }  // a * s
inline V3 scale(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }  // s * a
inline float dot(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
inline float length(V3 a) { return sqrtf(dot(a, a)); }
inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 normalize(V3 a) { return divs(a, length(a)); }
inline V3 cross(V3 a, V3 b) {
  return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
inline float fminw(float a, float b) { return fminf(a, b); }
inline float fmaxw(float a, float b) { return fmaxf(a, b); }
inline float clampw(float e, float lo, float hi) { return fminf(fmaxf(e, lo), hi); }

inline M3 I3() { return {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }
inline M3 mzero() { return {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}}; }
inline V3 mulmv(const M3& m, V3 v) {  // m * v = (m0 v.x + m1 v.y) + m2 v.z, both additions contracted
  return {fmaf(m.c[2].x, v.z, fmaf(m.c[1].x, v.y, m.c[0].x * v.x)),
          fmaf(m.c[2].y, v.z, fmaf(m.c[1].y, v.y, m.c[0].y * v.x)),
          fmaf(m.c[2].z, v.z, fmaf(m.c[1].z, v.y, m.c[0].z * v.x))};
}
inline M3 mulmm(const M3& a, const M3& b) {  // a * b: column j = a * b[j]
  return {{mulmv(a, b.c[0]), mulmv(a, b.c[1]), mulmv(a, b.c[2])}};
}
inline M3 transpose(const M3& m) {
  M3 t;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) at(t, c, r) = at(m, r, c);
  return t;
}
inline M3 madd(const M3& a, const M3& b) { return {{add(a.c[0], b.c[0]), add(a.c[1], b.c[1]), add(a.c[2], b.c[2])}}; }
inline M3 msub(const M3& a, const M3& b) { return {{sub(a.c[0], b.c[0]), sub(a.c[1], b.c[1]), sub(a.c[2], b.c[2])}}; }
inline M3 mscale(float s, const M3& a) { return {{scale(s, a.c[0]), scale(s, a.c[1]), scale(s, a.c[2])}}; }
inline float trace3(const M3& m) { return (m.c[0].x + m.c[1].y) + m.c[2].z; }
inline M3 diag3(V3 d) { return {{{d.x, 0, 0}, {0, d.y, 0}, {0, 0, d.z}}}; }
inline M3 outer3(V3 a, V3 b) { return {{scale(a, b.x), scale(a, b.y), scale(a, b.z)}}; }
// det(M) = dot(M[0], cross(M[1], M[2]))
inline float determinant(const M3& m) { return dot(m.c[0], cross(m.c[1], m.c[2])); }

// WGSL inverseSqrt(x) for x in [1, 2] (the only place the path uses it: the Jacobi rotation's
// t^2 + 1, |t| <= 1). WGSL gives inverseSqrt an accuracy of 2 ulp and no exact result, so this is
// one faithful reading: a quadratic approximation of x^-1/2 on [1, 2] (coefficients fitted to its
// relative error, 3.2e-3 at most) refined by two Newton steps y += y (1/2 - (x/2) y^2), each a
// fused multiply-add. Within 1 ulp of 1/sqrt(x) for every float in [1, 2]
// (tests/test_oracle_math.py checks all 2^23 + 1 of them).
float rsqrt_1to2(float x) {
  float y = fmaf(fmaf(0.14768398f, x, -0.7304885f), x, 1.5796084f);
  const float hx = 0.5f * x;
  for (int k = 0; k < 2; k++) {
    const float t = y * y;
    const float e = fmaf(-hx, t, 0.5f);
    y = fmaf(y, e, y);
  }
  return y;
}

// ---- linalg (src/shaders/linalg.js:14-69) ----
// One cyclic-Jacobi rotation (linalg.js:14-25): A <- J^T A J, V <- V J with J the Givens rotation
// in the (p, q) plane. Only rows / columns p and q change; each changed entry is the two-term sum
// the full 3x3 products form, in the same order (the remaining terms of those products are
// multiples of zero, which leave a nonzero sum unchanged).
void jacobi_rotate(M3& A, M3& V, int p, int q) {
  float apq = at(A, q, p);
  if (fabsf(apq) < 1e-30f) return;
  float theta = (at(A, q, q) - at(A, p, p)) / (2.0f * apq);
  float sgn = theta >= 0.0f ? 1.0f : -1.0f;
  float t = sgn / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
  float c = rsqrt_1to2(t * t + 1.0f);  // inverseSqrt (linalg.js:20); |t| <= 1, so t^2 + 1 is in [1, 2]
  float s = t * c;
  float ms = -s;
  // M = J^T A: rows p and q of every column
  for (int j = 0; j < 3; j++) {
    float ap = at(A, j, p), aq = at(A, j, q);
    at(A, j, p) = fmaf(ms, aq, c * ap);
    at(A, j, q) = fmaf(c, aq, s * ap);
  }
  // A = M J and V = V J: columns p and q
  for (int i = 0; i < 3; i++) {
    float mp = at(A, p, i), mq = at(A, q, i);
    at(A, p, i) = fmaf(mq, ms, mp * c);
    at(A, q, i) = fmaf(mq, c, mp * s);
    float vp = at(V, p, i), vq = at(V, q, i);
    at(V, p, i) = fmaf(vq, ms, vp * c);
    at(V, q, i) = fmaf(vq, c, vp * s);
  }
}

void eig_sym3(const M3& A0, M3& V, V3& e) {  // linalg.js:29-38
  M3 A = A0;
  V = I3();
  for (int sweep = 0; sweep < 4; sweep++) {
    jacobi_rotate(A, V, 0, 1);
    jacobi_rotate(A, V, 0, 2);
    jacobi_rotate(A, V, 1, 2);
  }
  e = v3(A.c[0].x, A.c[1].y, A.c[2].z);
}

void svd3(const M3& F, M3& U, V3& s, M3& V) {  // linalg.js:46-69
  V3 e;
  eig_sym3(mulmm(transpose(F), F), V, e);
  if (e.x < e.y) { e = v3(e.y, e.x, e.z); V = {{V.c[1], V.c[0], scale(-1.0f, V.c[2])}}; }
  if (e.x < e.z) { e = v3(e.z, e.y, e.x); V = {{V.c[2], scale(-1.0f, V.c[1]), V.c[0]}}; }
  if (e.y < e.z) { e = v3(e.x, e.z, e.y); V = {{scale(-1.0f, V.c[0]), V.c[2], V.c[1]}}; }
  if (determinant(V) < 0.0f) V.c[2] = scale(-1.0f, V.c[2]);
  s = v3(sqrtf(fmaxf(e.x, 0.0f)), sqrtf(fmaxf(e.y, 0.0f)), sqrtf(fmaxf(e.z, 0.0f)));
  V3 u0 = mulmv(F, V.c[0]);
  float l0 = length(u0);
  u0 = l0 > 1e-12f ? divs(u0, l0) : v3(1, 0, 0);
  V3 u1 = mulmv(F, V.c[1]);
  u1 = sub(u1, scale(dot(u1, u0), u0));
  float l1 = length(u1);
  if (l1 > 1e-12f) {
    u1 = divs(u1, l1);
  } else {
    u1 = fabsf(u0.x) < 0.9f ? v3(1, 0, 0) : v3(0, 1, 0);
    u1 = normalize(sub(u1, scale(dot(u1, u0), u0)));
  }
  V3 u2 = cross(u0, u1);
  if (dot(mulmv(F, V.c[2]), u2) < 0.0f) s.z = -s.z;
  U = {{u0, u1, u2}};
}

// ---- mpm module (src/physics/mpm.js:64-78) ----
constexpr float FP = 4194304.0f;
constexpr float INV_FP = 1.0f / 4194304.0f;
inline int32_t to_fixed(float x) { return (int32_t)rintf(clampw(x * FP, -2.0e9f, 2.0e9f)); }
inline void bspline(V3 fx, V3 w[3]) {
  V3 t0 = sub(v3(1.5f, 1.5f, 1.5f), fx);
  V3 t1 = sub(fx, v3(1.0f, 1.0f, 1.0f));
  V3 t2 = sub(fx, v3(0.5f, 0.5f, 0.5f));
  w[0] = v3((0.5f * t0.x) * t0.x, (0.5f * t0.y) * t0.y, (0.5f * t0.z) * t0.z);
  w[1] = v3(0.75f - t1.x * t1.x, 0.75f - t1.y * t1.y, 0.75f - t1.z * t1.z);
  w[2] = v3((0.5f * t2.x) * t2.x, (0.5f * t2.y) * t2.y, (0.5f * t2.z) * t2.z);
}
inline uint32_t part1by2(uint32_t x) {  // mpm.js:70-77
  uint32_t v = x & 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) { return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2); }

// ------------------------------------------------------------------------------------------------
// Scene: the reference's state in the reference layouts (mpm.js:111-124): particles are 8 vec4f per
// particle (AoS, 32 floats), the grid is vec4i per node, node = x + X*(y + Y*z).
struct Scene {
  int dims[3];
  int ppc_axis = 2;
  int sort_every = 4;
  float bound = 3.0f;
  uint32_t count = 0, cap = 0;
  uint64_t step_count = 0;
  float mats[96] = {0};
  float cols[192] = {0};
  uint32_t ncol = 0;
  bool law_on = false;         // accumulate the grid update's momentum bookkeeping (law below)
  uint64_t law[16] = {0};
  std::vector<float> P;        // 32 * cap
  std::vector<int32_t> grid;   // 4 * nodes (i32 after P2G; f32 bits after the grid update)
  size_t nodes() const { return (size_t)dims[0] * dims[1] * dims[2]; }
};

inline uint32_t mat_id(const float* rec) { uint32_t u; std::memcpy(&u, rec + 3, 4); return u; }

// P2G for one particle (mpm.js:318-349): the 27 fixed-point contributions, added with wrapping
// 32-bit integer arithmetic (the WGSL atomicAdd<i32> semantics).
template <class AddFn>
inline void p2g_particle(const float* p, const float* mats, float vol4, AddFn&& addq) {
  V3 xp = v3(p[0], p[1], p[2]);
  int base[3] = {(int)floorf(xp.x - 0.5f), (int)floorf(xp.y - 0.5f), (int)floorf(xp.z - 0.5f)};
  V3 fx = sub(xp, v3((float)base[0], (float)base[1], (float)base[2]));
  V3 w[3];
  bspline(fx, w);
  float mass = mats[mat_id(p) * 12 + 1];
  M3 tau = {{{p[11], p[23], p[27]}, {p[23], p[15], p[31]}, {p[27], p[31], p[19]}}};
  M3 C = {{{p[8], p[9], p[10]}, {p[12], p[13], p[14]}, {p[16], p[17], p[18]}}};
  M3 affine = madd(mscale(-vol4, tau), mscale(mass, C));
  V3 mv = scale(mass, v3(p[4], p[5], p[6]));
  // per node, as mpm.js:330-335 writes it: dpos = vec3f(off) - fx, wt = w[a].x * w[b].y * w[c].z,
  // mom = wt * (mv + affine * dpos), q = (toFixed(mom), toFixed(wt * mass))
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++) {
        V3 dpos = sub(v3((float)a, (float)b, (float)c), fx);
        float wt = (w[a].x * w[b].y) * w[c].z;
        V3 mom = scale(wt, add(mv, mulmv(affine, dpos)));
        int32_t q[4] = {to_fixed(mom.x), to_fixed(mom.y), to_fixed(mom.z), to_fixed(wt * mass)};
        addq(base[0] + a, base[1] + b, base[2] + c, q);
      }
}

void p2g(Scene& S) {  // mpm.js:607-612 (clear + dispatch), scatter mpm.js:318-349
  std::fill(S.grid.begin(), S.grid.end(), 0);
  const float vol4 = (float)(4.0 * (1.0 / std::pow((double)S.ppc_axis, 3)));  // mpm.js:610, f64 -> f32 uniform
  const int X = S.dims[0], Y = S.dims[1];
  uint32_t* g = reinterpret_cast<uint32_t*>(S.grid.data());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)S.count; i++) {
    p2g_particle(&S.P[32 * i], S.mats, vol4, [&](int gx, int gy, int gz, const int32_t* q) {
      size_t gi = 4 * ((size_t)gx + (size_t)X * ((size_t)gy + (size_t)Y * (size_t)gz));
      for (int k = 0; k < 4; k++) {
#pragma omp atomic
        g[gi + k] += (uint32_t)q[k];
      }
    });
  }
}

float collider_sdf(uint32_t kind, V3 d, const float* s) {  // mpm.js:419-426
  if (kind == 0u) return length(d) - s[0];
  if (kind == 1u) {
    V3 q = sub(v3(fabsf(d.x), fabsf(d.y), fabsf(d.z)), v3(s[0], s[1], s[2]));
    V3 qp = v3(fmaxf(q.x, 0.0f), fmaxf(q.y, 0.0f), fmaxf(q.z, 0.0f));
    return length(qp) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
  }
  return dot(d, v3(s[0], s[1], s[2]));
}

// Grid update for one node (mpm.js:428-470). Returns the f32 output (v, m) of the node.
// vfree (optional) receives p/m + g, the velocity before the boundary rules.
void grid_node(const int32_t* gin, float* out, int ix, int iy, int iz, const int dims[3], const float g[3],
               float bound, const float* cols, uint32_t ncol, float floor_friction, V3* vfree = nullptr) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  if (!(gin[3] > 0)) return;
  float m = (float)gin[3] * INV_FP;
  V3 v = v3(((float)gin[0] * INV_FP) / m + g[0], ((float)gin[1] * INV_FP) / m + g[1], ((float)gin[2] * INV_FP) / m + g[2]);
  if (vfree) *vfree = v;
  V3 c = v3((float)ix, (float)iy, (float)iz);
  V3 hi = v3((float)dims[0] - bound, (float)dims[1] - bound, (float)dims[2] - bound);
  if (c.y < bound && v.y < 0.0f) {
    float vn = -v.y;
    v.y = 0.0f;
    float vt = sqrtf(v.x * v.x + v.z * v.z);
    if (vt > 1e-9f) {
      float f = fmaxf(0.0f, 1.0f - (floor_friction * vn) / vt);
      v = v3(v.x * f, 0.0f, v.z * f);
    }
  }
  if (c.x < bound && v.x < 0.0f) v.x = 0.0f;
  if (c.z < bound && v.z < 0.0f) v.z = 0.0f;
  if (c.x > hi.x && v.x > 0.0f) v.x = 0.0f;
  if (c.y > hi.y && v.y > 0.0f) v.y = 0.0f;
  if (c.z > hi.z && v.z > 0.0f) v.z = 0.0f;
  for (uint32_t k = 0; k < ncol; k++) {
    const float* h = cols + 12 * k;
    const float* s = h + 4;
    V3 cv = v3(h[8], h[9], h[10]);
    uint32_t kind;
    std::memcpy(&kind, h, 4);
    V3 d = sub(c, v3(h[1], h[2], h[3]));
    float dist = collider_sdf(kind, d, s);
    if (dist < 1.0f) {
      const float e = 0.5f;
      V3 n = normalize(add(v3(collider_sdf(kind, add(d, v3(e, 0, 0)), s) - collider_sdf(kind, sub(d, v3(e, 0, 0)), s),
                              collider_sdf(kind, add(d, v3(0, e, 0)), s) - collider_sdf(kind, sub(d, v3(0, e, 0)), s),
                              collider_sdf(kind, add(d, v3(0, 0, e)), s) - collider_sdf(kind, sub(d, v3(0, 0, e)), s)),
                           v3(0.0f, 1e-7f, 0.0f)));
      V3 vr = sub(v, cv);
      float vn = dot(vr, n);
      if (vn < 0.0f) {
        vr = sub(vr, scale(vn, n));
        float vt = length(vr);
        if (vt > 1e-9f) vr = scale(vr, fmaxf(0.0f, 1.0f - (s[3] * -vn) / vt));
        v = add(vr, cv);
      }
    }
  }
  out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = m;
}

// A float quantity as 2^32 fixed point in 64 bits: rint(x * 2^32) (the scaling is exact), clamped to
// +-2^62 first (NaN clamps to -2^62).
inline uint64_t law_fixed(float x) {
  const float y = fminf(fmaxf(x * 4294967296.0f, -4.611686018427387904e18f), 4.611686018427387904e18f);
  return (uint64_t)(int64_t)rintf(y);
}

// The grid update's momentum bookkeeping for one node with mass (no counterpart in the reference:
// the momentum law checkMPM states for P2G, verify/index.js:175-193, carried through the grid
// update, mpm.js:428-470). With p the P2G momentum words, m and v the node's mass and velocity after
// the update and vf = p/m + g the velocity before the boundary rules (all f32 as the update computes
// them): [0] mass word, [1..3] momentum words (2^22), [4..6] m v, [7..9] m g, [10..12] m (v - vf),
// the impulse of the walls, the floor friction and the colliders (2^32, f32 products), [13] nodes.
// Wrapping 64-bit sums: independent of the order they are taken in.
inline void law_node(uint64_t* L, const int32_t* q, const float* out, V3 vf, const float g[3]) {
  const float m = out[3];
  const float vo[3] = {out[0], out[1], out[2]}, vfa[3] = {vf.x, vf.y, vf.z};
  L[0] += (uint64_t)(int64_t)q[3];
  for (int c = 0; c < 3; c++) {
    L[1 + c] += (uint64_t)(int64_t)q[c];
    L[4 + c] += law_fixed(m * vo[c]);
    L[7 + c] += law_fixed(m * g[c]);
    L[10 + c] += law_fixed(m * (vo[c] - vfa[c]));
  }
  L[13] += 1;
}

void grid_update(Scene& S, const float g[3], float floor_friction) {  // mpm.js:414-471, 593-595
  const int X = S.dims[0], Y = S.dims[1];
  uint64_t L[14] = {0};
  const bool law = S.law_on;
#pragma omp parallel for schedule(static) reduction(+ : L[:14])
  for (int64_t i = 0; i < (int64_t)S.nodes(); i++) {
    int ix = (int)(i % X), iy = (int)((i / X) % Y), iz = (int)(i / ((int64_t)X * Y));
    float out[4];
    V3 vf = v3(0, 0, 0);
    int32_t q[4];
    std::memcpy(q, &S.grid[4 * i], 16);
    grid_node(q, out, ix, iy, iz, S.dims, g, S.bound, S.cols, S.ncol, floor_friction, &vf);
    if (law && q[3] > 0) law_node(L, q, out, vf, g);
    std::memcpy(&S.grid[4 * i], out, 16);
  }
  if (law) {
    for (int k = 0; k < 14; k++) S.law[k] += L[k];
    S.law[14] += 1;  // grid updates summed
  }
}

// G2P for one particle (mpm.js:479-559). gv(x,y,z) returns the node's f32 velocity.
template <class GV>
inline void g2p_particle(float* p, const float* mats, const int dims[3], float bound, float vmax, GV&& gvat) {
  V3 xp = v3(p[0], p[1], p[2]);
  int base[3] = {(int)floorf(xp.x - 0.5f), (int)floorf(xp.y - 0.5f), (int)floorf(xp.z - 0.5f)};
  V3 fx = sub(xp, v3((float)base[0], (float)base[1], (float)base[2]));
  V3 w[3];
  bspline(fx, w);
  // v = sum wt gv and B = sum outer(wt gv, dpos) (mpm.js:487-496), wt = wx_a wy_b wz_c,
  // dpos = (a, b, c) - fx, summed per (a, b) row of three nodes: S = sum_c wz_c gv,
  // T = sum_c (wz_c dz_c) gv, then v += wxy S, B col x += (wxy dx_a) S, B col y += (wxy dy_b) S,
  // B col z += wxy T; every multiply-add contracted (the same sums, the rounding fixed this way)
  V3 d[3] = {sub(v3(0, 0, 0), fx), sub(v3(1, 1, 1), fx), sub(v3(2, 2, 2), fx)};
  float wzd[3] = {w[0].z * d[0].z, w[1].z * d[1].z, w[2].z * d[2].z};
  V3 v = v3(0, 0, 0);
  M3 B = mzero();
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      V3 g[3];
      for (int c = 0; c < 3; c++) g[c] = gvat(base[0] + a, base[1] + b, base[2] + c);
      V3 S = scale(w[0].z, g[0]), T = scale(wzd[0], g[0]);
      for (int c = 1; c < 3; c++) {
        S = v3(fmaf(w[c].z, g[c].x, S.x), fmaf(w[c].z, g[c].y, S.y), fmaf(w[c].z, g[c].z, S.z));
        T = v3(fmaf(wzd[c], g[c].x, T.x), fmaf(wzd[c], g[c].y, T.y), fmaf(wzd[c], g[c].z, T.z));
      }
      float wxy = w[a].x * w[b].y;
      float bxs = wxy * d[a].x, bys = wxy * d[b].y;
      v = v3(fmaf(wxy, S.x, v.x), fmaf(wxy, S.y, v.y), fmaf(wxy, S.z, v.z));
      B.c[0] = v3(fmaf(bxs, S.x, B.c[0].x), fmaf(bxs, S.y, B.c[0].y), fmaf(bxs, S.z, B.c[0].z));
      B.c[1] = v3(fmaf(bys, S.x, B.c[1].x), fmaf(bys, S.y, B.c[1].y), fmaf(bys, S.z, B.c[1].z));
      B.c[2] = v3(fmaf(wxy, T.x, B.c[2].x), fmaf(wxy, T.y, B.c[2].y), fmaf(wxy, T.z, B.c[2].z));
    }
  M3 C = mscale(4.0f, B);
  float sp = length(v);
  if (sp > vmax) v = scale(v, vmax / sp);
  float lo = bound - 1.0f;
  V3 x = v3(clampw(xp.x + v.x, lo, ((float)dims[0] - bound) + 1.0f),
            clampw(xp.y + v.y, lo, ((float)dims[1] - bound) + 1.0f),
            clampw(xp.z + v.z, lo, ((float)dims[2] - bound) + 1.0f));
  uint32_t mid = mat_id(p);
  const float* m0 = mats + 12 * mid;
  const float* m1 = m0 + 4;
  const float* m2 = m0 + 8;
  uint32_t kind;
  std::memcpy(&kind, m0, 4);
  M3 F = {{{p[20], p[21], p[22]}, {p[24], p[25], p[26]}, {p[28], p[29], p[30]}}};
  M3 tau = mzero();
  float vw = p[7];
  if (kind == 0u) {
    float J = clampw(vw * (1.0f + trace3(C)), 0.6f, 1.4f);
    vw = J;
    tau = madd(mscale(m0[2] * fminf(J - 1.0f, 0.02f), I3()), mscale(m1[0] * J, madd(C, transpose(C))));
    F = I3();
  } else {
    F = mulmm(madd(I3(), C), F);
    float mu0 = m0[2], la0 = m0[3];
    if (kind == 1u) {
      float J = fmaxf(determinant(F), 0.05f);
      tau = madd(mscale(mu0, msub(mulmm(F, transpose(F)), I3())), mscale(la0 * si_logf(J), I3()));
    } else if (kind == 2u) {
      M3 U, Vm;
      V3 sv;
      svd3(F, U, sv, Vm);
      V3 s = v3(clampw(sv.x, 1.0f - m1[2], 1.0f + m1[3]), clampw(sv.y, 1.0f - m1[2], 1.0f + m1[3]),
                clampw(sv.z, 1.0f - m1[2], 1.0f + m1[3]));
      float Jold = (sv.x * sv.y) * sv.z;
      float Jnew = (s.x * s.y) * s.z;
      vw = clampw((vw * Jold) / Jnew, 0.6f, 20.0f);
      F = mulmm(mulmm(U, diag3(s)), transpose(Vm));
      float h = si_expf(m1[1] * (1.0f - vw));
      float mu = mu0 * h, la = la0 * h;
      M3 R = mulmm(U, transpose(Vm));
      float J = Jnew;
      tau = madd(mulmm(mscale(2.0f * mu, msub(F, R)), transpose(F)), mscale((la * J) * (J - 1.0f), I3()));
    } else {
      M3 U, Vm;
      V3 sv;
      svd3(F, U, sv, Vm);
      V3 e = v3(si_logf(fmaxf(fabsf(sv.x), 1e-4f)), si_logf(fmaxf(fabsf(sv.y), 1e-4f)), si_logf(fmaxf(fabsf(sv.z), 1e-4f)));
      float tr = (e.x + e.y) + e.z;
      V3 dev = sub(e, v3(tr / 3.0f, tr / 3.0f, tr / 3.0f));
      float devn = length(dev);
      if (tr >= 0.0f) {
        e = v3(0, 0, 0);
      } else {
        float dg = devn + (((3.0f * la0 + 2.0f * mu0) / (2.0f * mu0)) * tr) * m2[0];
        if (dg > 0.0f && devn > 1e-12f) e = sub(e, divs(scale(dg, dev), devn));
      }
      V3 s = v3(si_expf(e.x), si_expf(e.y), si_expf(e.z));
      F = mulmm(mulmm(U, diag3(s)), transpose(Vm));
      float te = (e.x + e.y) + e.z;
      V3 dd = add(scale(2.0f * mu0, e), v3(la0 * te, la0 * te, la0 * te));
      tau = mulmm(mulmm(U, diag3(dd)), transpose(U));
    }
  }
  // mpm.js:557-559
  p[0] = x.x; p[1] = x.y; p[2] = x.z;
  p[4] = v.x; p[5] = v.y; p[6] = v.z; p[7] = vw;
  p[8] = C.c[0].x; p[9] = C.c[0].y; p[10] = C.c[0].z; p[11] = tau.c[0].x;
  p[12] = C.c[1].x; p[13] = C.c[1].y; p[14] = C.c[1].z; p[15] = tau.c[1].y;
  p[16] = C.c[2].x; p[17] = C.c[2].y; p[18] = C.c[2].z; p[19] = tau.c[2].z;
  p[20] = F.c[0].x; p[21] = F.c[0].y; p[22] = F.c[0].z; p[23] = tau.c[1].x;
  p[24] = F.c[1].x; p[25] = F.c[1].y; p[26] = F.c[1].z; p[27] = tau.c[2].x;
  p[28] = F.c[2].x; p[29] = F.c[2].y; p[30] = F.c[2].z; p[31] = tau.c[2].y;
}

void g2p(Scene& S, float vmax) {  // mpm.js:474-560, 598-600
  const int X = S.dims[0], Y = S.dims[1];
  const float* gf = reinterpret_cast<const float*>(S.grid.data());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)S.count; i++) {
    g2p_particle(&S.P[32 * i], S.mats, S.dims, S.bound, vmax, [&](int gx, int gy, int gz) {
      size_t gi = 4 * ((size_t)gx + (size_t)X * ((size_t)gy + (size_t)Y * (size_t)gz));
      return v3(gf[gi], gf[gi + 1], gf[gi + 2]);
    });
  }
}

// Sort key of a particle: Morton code of its stencil base cell floor(x - 1/2), 10 bits per axis.
// The reference keys by 4^3-cell block (mpm.js:567-568) with an atomic, unstable rank
// (mpm.js:570); this restatement refines the key to the base cell and breaks ties by the previous
// position in the array (stable), so the order is a deterministic function of the input.
inline uint32_t particle_key(const float* p) {
  uint32_t b[3];
  for (int a = 0; a < 3; a++) b[a] = (uint32_t)fmaxf(floorf(p[a] - 0.5f), 0.0f);
  return morton3(b[0], b[1], b[2]);
}

// MPM._sort (mpm.js:614-621). Particles end up grouped by 4^3-cell block in Morton order, which is
// the reference's order (its key is the block's Morton code); inside a block they are ordered by
// (rank within their base cell, base cell numbered x fastest: cx + 4 cy + 16 cz), where the rank
// counts earlier particles of the same cell in the input order. Consecutive particles therefore sit
// in different cells of the block. (The reference's sort is not stable, so it fixes no order inside
// a block; this one is the GPU's.)
uint32_t cell_in_block(uint32_t key) {  // Morton bits x0 y0 z0 x1 y1 z1 -> cx + 4 cy + 16 cz
  const uint32_t cx = (key & 1u) | ((key >> 2) & 2u), cy = ((key >> 1) & 1u) | ((key >> 3) & 2u);
  const uint32_t cz = ((key >> 2) & 1u) | ((key >> 4) & 2u);
  return cx + 4u * cy + 16u * cz;
}
void sort_particles(Scene& S) {
  const uint32_t n = S.count;
  std::vector<uint64_t> kv(n);
  for (uint32_t i = 0; i < n; i++) kv[i] = ((uint64_t)particle_key(&S.P[32 * (size_t)i]) << 32) | i;
  std::sort(kv.begin(), kv.end());  // by cell key, ties in input order
  std::vector<uint32_t> order(n);
  for (uint32_t b0 = 0; b0 < n;) {  // one block at a time
    const uint32_t blk = (uint32_t)(kv[b0] >> 38);
    uint32_t b1 = b0;
    while (b1 < n && (uint32_t)(kv[b1] >> 38) == blk) b1++;
    std::vector<std::pair<uint64_t, uint32_t>> loc;  // ((rank << 6) | cell, sorted position)
    uint32_t rank = 0;
    for (uint32_t j = b0; j < b1; j++) {
      rank = (j > b0 && (kv[j] >> 32) == (kv[j - 1] >> 32)) ? rank + 1 : 0;
      loc.push_back({((uint64_t)rank << 6) | cell_in_block((uint32_t)(kv[j] >> 32)), j});
    }
    std::sort(loc.begin(), loc.end());
    for (uint32_t j = b0; j < b1; j++) order[j] = (uint32_t)(kv[loc[j - b0].second] & 0xffffffffu);
    b0 = b1;
  }
  std::vector<float> P2(S.P.size());
  for (uint32_t j = 0; j < n; j++) std::memcpy(&P2[32 * (size_t)j], &S.P[32 * (size_t)order[j]], 128);
  std::copy(S.P.begin() + 32 * (size_t)n, S.P.end(), P2.begin() + 32 * (size_t)n);
  S.P.swap(P2);
}

// ------------------------------------------------------------------------------------------------
// Host-side seeding (mpm.js:170-232, 683-714) in float64, stored as float32.
struct XorShift32 {  // mpm.js:684-685
  uint32_t s;
  explicit XorShift32(uint32_t seed) : s(seed ? seed : 1u) {}
  double next() { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return (double)s / 4294967296.0; }
};

}  // namespace

// ================================================================================================
extern "C" {

float or_expf(float x) { return si_expf(x); }
float or_logf(float x) { return si_logf(x); }
uint32_t or_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
int32_t or_to_fixed(float x) { return to_fixed(x); }
void or_rsqrt_1to2(const float* x, float* y, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) y[i] = rsqrt_1to2(x[i]);
}
void or_bspline(const float fx[3], float w[9]) {
  V3 ww[3];
  bspline(v3(fx[0], fx[1], fx[2]), ww);
  for (int k = 0; k < 3; k++) { w[3 * k] = ww[k].x; w[3 * k + 1] = ww[k].y; w[3 * k + 2] = ww[k].z; }
}

// F, U, V column-major (9 floats: column 0 first). s: 3 floats.
void or_svd3(const float* F9, float* U9, float* s3, float* V9) {
  M3 F;
  for (int c = 0; c < 3; c++) F.c[c] = v3(F9[3 * c], F9[3 * c + 1], F9[3 * c + 2]);
  M3 U, V;
  V3 s;
  svd3(F, U, s, V);
  for (int c = 0; c < 3; c++) {
    U9[3 * c] = U.c[c].x; U9[3 * c + 1] = U.c[c].y; U9[3 * c + 2] = U.c[c].z;
    V9[3 * c] = V.c[c].x; V9[3 * c + 1] = V.c[c].y; V9[3 * c + 2] = V.c[c].z;
  }
  s3[0] = s.x; s3[1] = s.y; s3[2] = s.z;
}

// Exclusive prefix sum of n u32 (src/prims/scan.js:292-306; wrapping u32 arithmetic). May alias.
void or_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; i++) { uint32_t a = in[i]; out[i] = acc; acc += a; }
}

// Counting-sort binning (src/prims/grid.js:70-79: count, exclusive scan, scatter), with a stable
// rank: order[start[k] + r] = i for the r-th (in index order) element with key k.
// start has num_buckets + 1 entries; start[num_buckets] = n.
int or_counting_sort(const uint32_t* keys, uint64_t n, uint32_t num_buckets, uint32_t* order, uint32_t* start) {
  std::vector<uint32_t> cnt((size_t)num_buckets + 1, 0);
  for (uint64_t i = 0; i < n; i++) {
    if (keys[i] >= num_buckets) return -1;
    cnt[keys[i]]++;
  }
  or_scan_u32(cnt.data(), start, (uint64_t)num_buckets + 1);
  std::vector<uint32_t> fill(start, start + num_buckets);
  for (uint64_t i = 0; i < n; i++) order[fill[keys[i]]++] = (uint32_t)i;
  return 0;
}

// Material row (mpm.js:170-196). kind 0 fluid, 1 jelly, 2 snow, 3 sand.
void or_material_row(int kind, double density, double stiffness, double poisson, double viscosity, double hardening,
                     double compression, double stretch, double friction_deg, int ppc_axis, float* out12) {
  const double C_MAX = 0.4;
  double vol = 1.0 / std::pow((double)ppc_axis, 3);
  double mass = density * vol;
  double rho = mass / vol;
  double c = C_MAX * std::min(1.0, std::max(0.01, stiffness));
  double nu = std::min(0.49, std::max(0.0, poisson));
  double M = rho * c * c;
  double mu = M * (1 - 2 * nu) / (2 * (1 - nu));
  double la = M * nu / (1 - nu);
  std::memset(out12, 0, 48);
  uint32_t k = (uint32_t)kind;
  std::memcpy(out12, &k, 4);
  out12[1] = (float)mass;
  out12[2] = (float)(kind == 0 ? rho * c * c : mu);
  out12[3] = (float)la;
  out12[4] = (float)(std::min(4.0, std::max(0.0, viscosity)) * 0.05 * rho);
  out12[5] = (float)hardening;
  out12[6] = (float)compression;
  out12[7] = (float)stretch;
  double phi = (friction_deg * M_PI) / 180;
  double sinp = std::sin(phi);
  out12[8] = (float)(std::sqrt(2.0 / 3.0) * (2 * sinp) / (3 - sinp));
}

// sampleShape (mpm.js:683-714). shape 0 box (p = lo[3], hi[3]), 1 sphere (c[3], r),
// 2 cylinder (c[3], r, h). Writes float32 xyz triples (world units) into out (capacity in points);
// returns the number of points (even if it exceeds the capacity; only capacity are written).
int64_t or_sample_shape(int shape, const double* prm, double spacing, double jitter, uint32_t seed, float* out,
                        int64_t capacity) {
  double lo[3], hi[3];
  double c[3] = {0, 0, 0}, r = 0;
  if (shape == 0) {
    for (int a = 0; a < 3; a++) { lo[a] = prm[a]; hi[a] = prm[3 + a]; }
  } else if (shape == 1) {
    for (int a = 0; a < 3; a++) c[a] = prm[a];
    r = prm[3];
    for (int a = 0; a < 3; a++) { lo[a] = c[a] - r; hi[a] = c[a] + r; }
  } else if (shape == 2) {
    for (int a = 0; a < 3; a++) c[a] = prm[a];
    r = prm[3];
    double h = prm[4];
    lo[0] = c[0] - r; lo[1] = c[1] - h; lo[2] = c[2] - r;
    hi[0] = c[0] + r; hi[1] = c[1] + h; hi[2] = c[2] + r;
  } else {
    return -1;
  }
  int64_t n[3];
  for (int a = 0; a < 3; a++) n[a] = std::max<int64_t>(1, (int64_t)std::floor((hi[a] - lo[a]) / spacing));
  XorShift32 rng(seed);
  int64_t cnt = 0;
  double p[3];
  for (int64_t z = 0; z < n[2]; z++)
    for (int64_t y = 0; y < n[1]; y++)
      for (int64_t x = 0; x < n[0]; x++) {
        p[0] = lo[0] + ((double)x + 0.5 + (rng.next() - 0.5) * jitter) * spacing;
        p[1] = lo[1] + ((double)y + 0.5 + (rng.next() - 0.5) * jitter) * spacing;
        p[2] = lo[2] + ((double)z + 0.5 + (rng.next() - 0.5) * jitter) * spacing;
        bool in = true;
        if (shape == 1) in = std::pow(p[0] - c[0], 2) + std::pow(p[1] - c[1], 2) + std::pow(p[2] - c[2], 2) <= r * r;
        if (shape == 2) in = std::pow(p[0] - c[0], 2) + std::pow(p[2] - c[2], 2) <= r * r;
        if (!in) continue;
        if (cnt < capacity) for (int a = 0; a < 3; a++) out[3 * cnt + a] = (float)p[a];
        cnt++;
      }
  return cnt;
}

// Packing of MPM.add (mpm.js:213-227) for n world-space float32 points. vel: m/s. s = dt/dx.
// noise_sigma (m/s) > 0 adds per-component Gaussian velocity noise from a Box-Muller transform of
// an xorshift32 stream seeded with noise_seed (two uniform pairs per particle, first three normals
// used) — an extension the reference does not have, used by the mixed-material workload.
void or_pack_particles(const float* pts, int64_t n, const double* origin, double dx, const int* dims, double bound,
                       uint32_t mat, const double* vel, double s, double noise_sigma, uint32_t noise_seed,
                       float* rec) {
  XorShift32 rng(noise_seed);
  for (int64_t k = 0; k < n; k++) {
    float* o = rec + 32 * k;
    std::memset(o, 0, 128);
    for (int a = 0; a < 3; a++) {
      double x = ((double)pts[3 * k + a] - origin[a]) / dx;
      double hi = (double)dims[a] - bound;
      o[a] = (float)std::min(hi - 1e-3, std::max(bound + 1e-3, x));
    }
    std::memcpy(o + 3, &mat, 4);
    double vv[3] = {vel[0], vel[1], vel[2]};
    if (noise_sigma > 0) {
      double z[4];
      for (int j = 0; j < 2; j++) {
        double u1 = 1.0 - rng.next();  // (0, 1]
        double u2 = rng.next();
        double rr = std::sqrt(-2.0 * std::log(u1));
        z[2 * j] = rr * std::cos(2 * M_PI * u2);
        z[2 * j + 1] = rr * std::sin(2 * M_PI * u2);
      }
      for (int a = 0; a < 3; a++) vv[a] = vv[a] + noise_sigma * z[a];
    }
    o[4] = (float)(vv[0] * s); o[5] = (float)(vv[1] * s); o[6] = (float)(vv[2] * s);
    o[7] = 1.0f;
    o[20] = 1.0f; o[25] = 1.0f; o[30] = 1.0f;
  }
}

// ---- scene lifecycle ----
void* or_create(int X, int Y, int Z, int ppc_axis, int sort_every, uint32_t cap) {
  Scene* S = new Scene;
  S->dims[0] = X; S->dims[1] = Y; S->dims[2] = Z;
  S->ppc_axis = ppc_axis;
  S->sort_every = sort_every;
  S->cap = cap;
  S->P.assign((size_t)cap * 32, 0.0f);
  S->grid.assign(S->nodes() * 4, 0);
  return S;
}
void or_destroy(void* s) { delete static_cast<Scene*>(s); }
void or_set_materials(void* s, const float* tbl, uint32_t n) {
  Scene* S = static_cast<Scene*>(s);
  std::memset(S->mats, 0, sizeof S->mats);
  std::memcpy(S->mats, tbl, 48 * (size_t)std::min<uint32_t>(n, 8));
}
void or_set_colliders(void* s, const float* tbl, uint32_t n) {
  Scene* S = static_cast<Scene*>(s);
  S->ncol = std::min<uint32_t>(n, 16);
  std::memset(S->cols, 0, sizeof S->cols);
  std::memcpy(S->cols, tbl, 48 * (size_t)S->ncol);
}
void or_write_particles(void* s, const float* rec, uint32_t offset, uint32_t n) {
  Scene* S = static_cast<Scene*>(s);
  std::memcpy(&S->P[32 * (size_t)offset], rec, 128 * (size_t)n);
}
void or_read_particles(void* s, float* rec, uint32_t offset, uint32_t n) {
  Scene* S = static_cast<Scene*>(s);
  std::memcpy(rec, &S->P[32 * (size_t)offset], 128 * (size_t)n);
}
void or_set_count(void* s, uint32_t n) { static_cast<Scene*>(s)->count = n; }
void or_set_step_count(void* s, uint64_t n) { static_cast<Scene*>(s)->step_count = n; }
uint64_t or_get_step_count(void* s) { return static_cast<Scene*>(s)->step_count; }
// MPM.sortEvery = n (mpm.js:101, 591): the sort test of the next substeps uses n
void or_set_sort_every(void* s, int n) { static_cast<Scene*>(s)->sort_every = n > 0 ? n : 1; }
void or_read_grid(void* s, int32_t* words) {
  Scene* S = static_cast<Scene*>(s);
  std::memcpy(words, S->grid.data(), S->grid.size() * 4);
}
void or_write_grid(void* s, const int32_t* words) {
  Scene* S = static_cast<Scene*>(s);
  std::memcpy(S->grid.data(), words, S->grid.size() * 4);
}
void or_p2g(void* s) { p2g(*static_cast<Scene*>(s)); }
void or_grid_update(void* s, const float* g, float floor_friction) { grid_update(*static_cast<Scene*>(s), g, floor_friction); }
void or_g2p(void* s, float vmax) { g2p(*static_cast<Scene*>(s), vmax); }
void or_sort(void* s) { sort_particles(*static_cast<Scene*>(s)); }
void or_rescale(void* s, float k) {  // mpm.js:405-413
  Scene* S = static_cast<Scene*>(s);
  for (uint32_t i = 0; i < S->count; i++)
    for (int q = 1; q <= 4; q++)
      for (int a = 0; a < 3; a++) S->P[32 * (size_t)i + 4 * q + a] *= k;
}
void or_keys(void* s, uint32_t* keys) {
  Scene* S = static_cast<Scene*>(s);
  for (uint32_t i = 0; i < S->count; i++) keys[i] = particle_key(&S->P[32 * (size_t)i]);
}
// MPM.step's substep loop (mpm.js:590-602).
// The grid update's momentum bookkeeping (law_node): on = 1 starts it from zero, 0 stops it.
void or_set_grid_law(void* s, int on) {
  Scene* S = static_cast<Scene*>(s);
  S->law_on = on != 0;
  std::memset(S->law, 0, sizeof S->law);
}
// out[16]: the sums since the last reset ([14] = grid updates); reset = 1 zeroes them.
void or_read_grid_law(void* s, uint64_t* out, int reset) {
  Scene* S = static_cast<Scene*>(s);
  std::memcpy(out, S->law, sizeof S->law);
  if (reset) std::memset(S->law, 0, sizeof S->law);
}
void or_step(void* s, uint32_t substeps, const float* g, float floor_friction, float vmax) {
  Scene* S = static_cast<Scene*>(s);
  if (S->count == 0) return;
  for (uint32_t k = 0; k < substeps; k++) {
    if (S->step_count % (uint64_t)S->sort_every == 0) sort_particles(*S);
    p2g(*S);
    grid_update(*S, g, floor_friction);
    g2p(*S, vmax);
    S->step_count++;
  }
}
int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void or_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
