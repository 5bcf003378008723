// Register-resident Gauss-Jordan elimination with partial pivoting.
//
// An N x NCOL matrix [S | R] (columns >= N are right-hand sides) lives in the
// registers of a workgroup of W waves: lane l of wave w owns rows l + 64 q
// (q < RQ) and columns w + W cq (cq < CQ).  One step per matrix column k:
//   1. the wave owning column k picks the pivot -- the largest |a_rk| over
//      rows not yet used as pivots -- from the copy of column k in LDS, with a
//      DPP max over a 64-bit key (|a_rk| bits, low 8 mantissa bits replaced by
//      255 - row, so near-exact ties go to the lower row as in LAPACK's idamax);
//   2. every wave reads the multipliers a_rk / pivot from LDS and the pivot
//      row from the lane that holds it (v_readlane, so it is a scalar operand)
//      and removes column k from every row except the pivot row.
// Pivot rows are left unscaled; the solution of unknown k is
// a[pivrow[k]][rhs] / pivval[k].  Matrix columns <= k are skipped (uniformly
// per wave).  Rows never move: pivrow[k] records the physical row that holds
// unknown k.  No register of `a` is addressed with a step-dependent index:
// the wave owning column k+1 copies it to LDS while applying step k.  One
// barrier per step (the LDS column buffer is double-buffered).
#pragma once
#include "colloc_common.h"

namespace colloc {

template <int N, int NCOL, int W>
struct GJ {
  static constexpr int RQ = (N + WAVE - 1) / WAVE;
  static constexpr int CQ = (NCOL + W - 1) / W;
  static constexpr int COLBUF = 2 * RQ * WAVE;   // doubles of LDS for the column buffer
  static_assert(N <= 256, "row indices are packed into 8 bits of the pivot key");

  double a[RQ][CQ];

  // Global column index of column slot cq in this wave.
  __device__ __forceinline__ static int col(int cq) { return wave_id() + W * cq; }
  __device__ __forceinline__ static int row(int q) { return lane_id() + WAVE * q; }

  // Returns nonzero if a zero or non-finite pivot was met.
  // LDS: colbuf (COLBUF doubles), pivrow (N ints), pivval (N doubles).
  __device__ int eliminate(double* colbuf, int* pivrow, double* pivval) {
    const int lane = lane_id();
    const int wave = wave_id();
    unsigned used = 0;
#pragma unroll
    for (int q = 0; q < RQ; ++q)
      if (row(q) >= N) used |= 1u << q;
    if (wave == 0) {
#pragma unroll
      for (int q = 0; q < RQ; ++q) colbuf[lane + WAVE * q] = a[q][0];
    }
    __syncthreads();
    int bad = 0;
    for (int k = 0; k < N; ++k) {
      double* cb = colbuf + (k & 1) * (RQ * WAVE);
      double* cbn = colbuf + ((k + 1) & 1) * (RQ * WAVE);
      if (wave == k % W) {
        unsigned long long key = 0;
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          const unsigned long long kq =
              (((unsigned long long)__double_as_longlong(fabs(cb[lane + WAVE * q]))) & ~0xFFull) |
              (unsigned long long)(255 - row(q));
          if (!((used >> q) & 1u) && kq > key) key = kq;
        }
        key = wave_max_u64(key);
        if (lane == 0) {
          const int pr = key ? 255 - (int)(key & 0xFF) : 0;
          pivrow[k] = pr;
          pivval[k] = cb[pr];
        }
      }
      __syncthreads();
      const int p = pivrow[k];
      const double piv = pivval[k];
      if (!(fabs(piv) > 0.0) || !(fabs(piv) < 1.0e308)) bad = 1;
      const double inv = 1.0 / piv;
      const int pl = p & (WAVE - 1);
      const int pq = p >> 6;
      if (lane == pl) used |= 1u << pq;
      double l[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) l[q] = (q == pq && lane == pl) ? 0.0 : cb[lane + WAVE * q] * inv;
#pragma unroll
      for (int cq = 0; cq < CQ; ++cq) {
        const int c = wave + W * cq;
        if (c >= NCOL) continue;
        if (c < N && c <= k) continue;
        // pivot-row value: every slot of lane pl is read with v_readlane and the
        // choice among them is made on the scalar results
        double pv = readlane_d(a[0][cq], pl);
#pragma unroll
        for (int q = 1; q < RQ; ++q) {
          const double sq = readlane_d(a[q][cq], pl);
          pv = (pq == q) ? sq : pv;
        }
#pragma unroll
        for (int q = 0; q < RQ; ++q) a[q][cq] = fma(-l[q], pv, a[q][cq]);
        if (c == k + 1) {
#pragma unroll
          for (int q = 0; q < RQ; ++q) cbn[lane + WAVE * q] = a[q][cq];
        }
      }
    }
    return bad;
  }
};

}  // namespace colloc
