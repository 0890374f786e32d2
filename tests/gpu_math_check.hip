// GPU check of the kernels' short square root and reciprocal (sqrt_ge1, rcp_ge1,
// shaderintel_amd/csrc/mpm_math.hpp) against the IEEE operations, bit for bit, over every float of a
// range of bit patterns, and a run of the kernels' SVD (svd3) over given matrices for a comparison
// with the oracle's. Built as a small shared library by the package build; tests/test_gpu_math.py
// calls it.
#include <hip/hip_runtime.h>

#include "mpm_math.hpp"

namespace {
__global__ void k_check_sqrt_ge1(uint32_t first, uint32_t count, unsigned long long* bad, uint32_t* first_bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t bits = first + i;
  const float x = __uint_as_float(bits);
  if (__float_as_uint(mpm::sqrt_ge1(x)) != __float_as_uint(sqrtf(x))) {
    atomicAdd(bad, 1ull);
    atomicMin(first_bad, bits);
  }
}
__global__ void k_check_rcp_ge1(uint32_t first, uint32_t count, unsigned long long* bad, uint32_t* first_bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t bits = first + i;
  const float d = __uint_as_float(bits);
  if (__float_as_uint(mpm::rcp_ge1(d)) != __float_as_uint(1.0f / d)) {
    atomicAdd(bad, 1ull);
    atomicMin(first_bad, bits);
  }
}
// svd3 of n column-major 3x3 matrices (9 floats each), one per thread: U, s, V out.
__global__ void k_svd3(const float* __restrict__ F, uint32_t n, float* __restrict__ U, float* __restrict__ s,
                       float* __restrict__ V) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* f = F + 9 * (size_t)i;
  const mpm::m3 M = mpm::mat(mpm::mk(f[0], f[1], f[2]), mpm::mk(f[3], f[4], f[5]), mpm::mk(f[6], f[7], f[8]));
  mpm::m3 u, v;
  mpm::f3 sv;
  mpm::svd3(M, u, sv, v);
  float* uo = U + 9 * (size_t)i;
  float* vo = V + 9 * (size_t)i;
  for (int c = 0; c < 3; c++) {
    uo[3 * c] = u.c[c].x; uo[3 * c + 1] = u.c[c].y; uo[3 * c + 2] = u.c[c].z;
    vo[3 * c] = v.c[c].x; vo[3 * c + 1] = v.c[c].y; vo[3 * c + 2] = v.c[c].z;
  }
  s[3 * (size_t)i] = sv.x; s[3 * (size_t)i + 1] = sv.y; s[3 * (size_t)i + 2] = sv.z;
}
}  // namespace

// svd3 on the GPU of n column-major 3x3 matrices (host arrays: F 9n floats in; U, V 9n and s 3n
// out), in 256-thread blocks so that each wavefront holds 64 consecutive matrices. Returns a
// hipError_t code.
extern "C" int run_svd3(const float* F, uint32_t n, float* U, float* s, float* V) {
  float *dF = nullptr, *dU = nullptr, *ds = nullptr, *dV = nullptr;
  const size_t b9 = 36 * (size_t)n, b3 = 12 * (size_t)n;
  hipError_t e = hipMalloc(&dF, b9 ? b9 : 4);
  if (e == hipSuccess) e = hipMalloc(&dU, b9 ? b9 : 4);
  if (e == hipSuccess) e = hipMalloc(&dV, b9 ? b9 : 4);
  if (e == hipSuccess) e = hipMalloc(&ds, b3 ? b3 : 4);
  if (e == hipSuccess && n) e = hipMemcpy(dF, F, b9, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) {
    hipLaunchKernelGGL(k_svd3, dim3((n + 255u) / 256u), dim3(256), 0, 0, dF, n, dU, ds, dV);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && n) e = hipMemcpy(U, dU, b9, hipMemcpyDeviceToHost);
  if (e == hipSuccess && n) e = hipMemcpy(V, dV, b9, hipMemcpyDeviceToHost);
  if (e == hipSuccess && n) e = hipMemcpy(s, ds, b3, hipMemcpyDeviceToHost);
  hipFree(dF);
  hipFree(dU);
  hipFree(dV);
  hipFree(ds);
  return (int)e;
}

// Mismatches against the IEEE operation over the bit patterns [first, first + count) of
// sqrt_ge1 (which = 0, against sqrtf) or rcp_ge1 (which = 1, against 1.0f / d), and the first
// mismatching pattern (0xffffffff if none). Returns a hipError_t code.
extern "C" int check_short_math(int which, uint32_t first, uint32_t count, unsigned long long* mismatches,
                                uint32_t* first_bad) {
  unsigned long long* d_bad = nullptr;
  uint32_t* d_first = nullptr;
  hipError_t e = hipMalloc(&d_bad, 8);
  if (e == hipSuccess) e = hipMalloc(&d_first, 4);
  if (e == hipSuccess) e = hipMemset(d_bad, 0, 8);
  if (e == hipSuccess) e = hipMemset(d_first, 0xff, 4);
  if (e == hipSuccess && count) {
    const dim3 grid((count + 255u) / 256u), block(256);
    if (which == 0)
      hipLaunchKernelGGL(k_check_sqrt_ge1, grid, block, 0, 0, first, count, d_bad, d_first);
    else
      hipLaunchKernelGGL(k_check_rcp_ge1, grid, block, 0, 0, first, count, d_bad, d_first);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(mismatches, d_bad, 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(first_bad, d_first, 4, hipMemcpyDeviceToHost);
  if (d_bad) hipFree(d_bad);
  if (d_first) hipFree(d_first);
  return (int)e;
}
