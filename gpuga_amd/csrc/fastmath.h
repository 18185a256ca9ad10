// Device math used by every kernel.
//
//  * ldu(): load through the constant address space.  With a wave-uniform
//    index the compiler issues a scalar load (s_load) and keeps the value in
//    SGPRs, so branches on it are scalar; used only for buffers that no
//    kernel writes while it runs (the training-set tables).
//  * exp2 / log2 / reciprocal: the CDNA transcendental instructions
//    (v_exp_f32, v_log_f32, v_rcp_f32, about 1 ulp on normal inputs), and
//    tanh from them (the NN2B tabulation's hidden layers).
//  * cos(pi t), sin(pi t) for t in [0, 1] (the cutoff taper): reduced to
//    y = pi (t - 1/2) in [-pi/2, pi/2] and evaluated from the Taylor series of
//    sin y (to y^11) and cos y (to y^12), whose truncation errors on that
//    interval are below 6e-8 and 7e-9.
#pragma once

#include <hip/hip_runtime.h>

namespace gpuga {

template <class T>
__device__ __forceinline__ T ldu(const T* p, long i)
{
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) T CT;
  return ((CT*)p)[i];
#else
  return p[i];
#endif
}

// ldu() when the index is wave-uniform (U), an ordinary per-lane load otherwise.
template <bool U, class T>
__device__ __forceinline__ T ldx(const T* p, long i)
{
  if constexpr (U) return ldu(p, i);
  else return p[i];
}

constexpr float LOG2E = 1.4426950408889634f;
constexpr float PI_F = 3.14159265358979f;

__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// tanh x = 1 - 2 / (e^(2x) + 1) from exp2 and the reciprocal: absolute
// error about 2e-7 (the subtraction from 1 sets it near x = 0); -1 and +1
// for large |x| (e^(2x) flushes to 0 or overflows to +inf)
__device__ __forceinline__ float ftanh(float x)
{
  return 1.0f - 2.0f * frcp(fexp2(x * (2.0f * LOG2E)) + 1.0f);
}

// sin(pi t) and cos(pi t) for 0 <= t <= 1
__device__ __forceinline__ void sincospi01(float t, float& s, float& c)
{
  const float y = PI_F * (t - 0.5f);
  const float y2 = y * y;
  float ps = -1.0f / 39916800.0f;
  ps = ps * y2 + 1.0f / 362880.0f;
  ps = ps * y2 - 1.0f / 5040.0f;
  ps = ps * y2 + 1.0f / 120.0f;
  ps = ps * y2 - 1.0f / 6.0f;
  const float siny = y + y * y2 * ps;
  float pc = 1.0f / 479001600.0f;
  pc = pc * y2 - 1.0f / 3628800.0f;
  pc = pc * y2 + 1.0f / 40320.0f;
  pc = pc * y2 - 1.0f / 720.0f;
  pc = pc * y2 + 1.0f / 24.0f;
  pc = pc * y2 - 0.5f;
  const float cosy = 1.0f + y2 * pc;
  s = cosy;  // sin(pi t) = cos(y)
  c = -siny; // cos(pi t) = -sin(y)
}

} // namespace gpuga
