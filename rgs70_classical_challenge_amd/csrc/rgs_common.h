// Shared device helpers and host error plumbing for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "rgs_status.h"

namespace rgs {

#define RGS_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::rgs::fail(_e == hipErrorOutOfMemory ? ::rgs::kNoMemory : ::rgs::kHip, \
                         std::string(#expr) + ": " + hipGetErrorString(_e));       \
  } while (0)

// Checks the launch status of the kernel just enqueued.
#define RGS_LAUNCH_CHECK(what)                                                     \
  do {                                                                             \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return ::rgs::fail(::rgs::kHip, std::string(what) + ": " + hipGetErrorString(_e)); \
  } while (0)

// ---------------------------------------------------------------- device math
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Products round as written (one fma over one product), never as the
// compiler's contraction chooses: every kernel that inlines them gives the
// same bits for the same inputs.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__fmaf_rn(a.x, b.x, -__fmul_rn(a.y, b.y)), __fmaf_rn(a.x, b.y, __fmul_rn(a.y, b.x)));
}
__device__ __forceinline__ float2 cneg_if(float2 a, uint32_t neg) {
  // Flip both signs when neg is 1 (exact; no multiply).
  uint32_t s = neg << 31;
  return make_float2(__uint_as_float(__float_as_uint(a.x) ^ s), __uint_as_float(__float_as_uint(a.y) ^ s));
}
// |a|^2 in float64: both squares are exact for float inputs, one rounding in the add.
__device__ __forceinline__ double prob64(float2 a) {
  double re = (double)a.x, im = (double)a.y;
  return __dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im));
}

}  // namespace rgs
