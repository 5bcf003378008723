// Shared device helpers for the collocation kernels (gfx950, wave64).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace colloc {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Broadcast a double held by lane `src` (uniform) to the whole wave.
__device__ __forceinline__ double readlane_d(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// NaN-propagating max (fmax would drop a NaN and hide a diverged residual).
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// max|.| over a workgroup folded into a u64 slot.  Non-negative doubles order
// like their bit patterns; a NaN (sign cleared by fabs) wins over any number.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* slot, double* sred) {
  v = wave_max(v);
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  if (lane_id() == 0) sred[wave_id()] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sred[0];
    for (int w = 1; w < nw; ++w) m = nmax(sred[w], m);
    atomicMax(slot, (unsigned long long)__double_as_longlong(fabs(m)));
  }
}

}  // namespace colloc
