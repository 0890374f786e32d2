// GPU check of the kernels' short square root and reciprocal (sqrt_ge1, rcp_ge1,
// shaderintel_amd/csrc/mpm_math.hpp) against the IEEE operations, bit for bit, over every float of a
// range of bit patterns. Built as a small shared library by the package build; tests/test_gpu_math.py
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
}  // namespace

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
