// Shared device helpers for the collocation kernels (gfx950, wave64).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace colloc {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
// Wave index as a uniform (scalar) value, so that tests on it compile to
// scalar branches instead of exec-mask manipulation.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Broadcast a double held by lane `src` (uniform) to the whole wave.
__device__ __forceinline__ double readlane_d(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// Compile-time loop: f(Int<I>{}) for I in [B, E).  Used where a loop index
// must be a constant (register-array subscripts) even when the body is too
// large for the unroller.
template <int V>
struct Int {
  static constexpr int value = V;
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(Int<B>{});
    static_for<B + 1, E>(f);
  }
}

// 64-bit DPP move within rows of 16 lanes (CTRL: a DPP control code).
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)v, (int)(unsigned)v, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(v >> 32), (int)(v >> 32), CTRL, 0xf, 0xf, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Wave-wide max of a u64 (result uniform): butterflies within rows of 16
// (quad_perm xor 1, xor 2, row_half_mirror, row_mirror) then the four rows.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  unsigned long long o;
  o = dpp_u64<0xB1>(v);  v = o > v ? o : v;
  o = dpp_u64<0x4E>(v);  v = o > v ? o : v;
  o = dpp_u64<0x141>(v); v = o > v ? o : v;
  o = dpp_u64<0x140>(v); v = o > v ? o : v;
  unsigned long long m = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)v, 16 * r);
    const unsigned hi = __builtin_amdgcn_readlane((int)(v >> 32), 16 * r);
    const unsigned long long x = ((unsigned long long)hi << 32) | lo;
    m = x > m ? x : m;
  }
  return m;
}

// Wave-wide max of a u32 (result uniform): DPP butterflies within rows of
// 16 lanes, then the four row results through v_readlane.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  unsigned o;
  o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xf, 0xf, false);  v = o > v ? o : v;
  o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xf, 0xf, false);  v = o > v ? o : v;
  o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xf, 0xf, false); v = o > v ? o : v;
  o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xf, 0xf, false); v = o > v ? o : v;
  const unsigned r0 = __builtin_amdgcn_readlane((int)v, 0), r1 = __builtin_amdgcn_readlane((int)v, 16);
  const unsigned r2 = __builtin_amdgcn_readlane((int)v, 32), r3 = __builtin_amdgcn_readlane((int)v, 48);
  const unsigned m01 = r0 > r1 ? r0 : r1, m23 = r2 > r3 ? r2 : r3;
  return m01 > m23 ? m01 : m23;
}

// Cross-row lane exchanges for a wave seen as 4 rows of 16 lanes (gfx950
// v_permlane16_swap / v_permlane32_swap; VALU, no LDS round trip).
// sum_rows(v): in every lane, the sum of v over the 4 lanes with the same
// lane index modulo 16.
__device__ __forceinline__ double sum_rows(double v) {
  auto l16 = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  auto h16 = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  // rows (0,1,2,3): [0] holds rows (0,0,2,2), [1] rows (1,1,3,3)
  v = __hiloint2double((int)h16[0], (int)l16[0]) + __hiloint2double((int)h16[1], (int)l16[1]);
  auto l32 = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  auto h32 = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  // halves (lo, hi): [0] holds (lo, lo), [1] holds (hi, hi)
  return __hiloint2double((int)h32[0], (int)l32[0]) + __hiloint2double((int)h32[1], (int)l32[1]);
}

// bcast_row<R>(v): in every lane, v of the lane in row R with the same lane
// index modulo 16.
template <int R>
__device__ __forceinline__ double bcast_row(double v) {
  static_assert(R >= 0 && R < 4, "four rows of 16 lanes");
  auto l16 = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  auto h16 = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  const int lo = (int)l16[R & 1], hi = (int)h16[R & 1];     // rows (R', R', R'+2, R'+2), R' = R & 1
  auto l32 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  auto h32 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double((int)h32[R >> 1], (int)l32[R >> 1]);
}

// 1/x in FP64 from the hardware reciprocal estimate and two Newton steps
// (error within a few ulp; used where a pivot's inverse is on a latency
// chain).
__device__ __forceinline__ double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// NaN-propagating max (fmax would drop a NaN and hide a diverged residual).
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || a > b) ? a : b; }
__device__ __forceinline__ float nmax(float a, float b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
  return v;
}

// Broadcast a value held by lane `src` (uniform), either precision.
__device__ __forceinline__ double readlane_r(double v, int src) { return readlane_d(v, src); }
__device__ __forceinline__ float readlane_r(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
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

// Tables of a problem program in the polynomial form (large problems): the
// values of nb base atoms (y | z | p | instance constants | alpha) at a
// point, the nodes (nonlinear functions of polynomials, atoms nb ..
// natoms-1), the monomials (products of atom powers; monomial 0 is 1) and
// the polynomials, in sliced ELL slabs of 64 rows (term t of row q of slab s
// at slab_term_off[s] + 64 t + q).  Node rows are slabs
// [node_slab_off[L], node_slab_off[L + 1]) for the nodes of level L + 1
// (slab row q = node index); output rows are the slabs from out0 on (slab
// row q = slab_rows[64 s + q], the output's destination, -1: padding).
struct TabProg {
  int nb, natoms, nmono, nlev, nslab, out0;
  const int* mono_fac_off;
  const int* mono_fac_atom;
  const signed char* mono_fac_pow;
  const int* mono_lev_off;
  const int* node_slab_off;
  const unsigned char* node_kind;
  const double* node_r;
  const int* slab_term_off;
  const int* slab_rows;
  const double* coef;
  const int* mono;
};

}  // namespace colloc
