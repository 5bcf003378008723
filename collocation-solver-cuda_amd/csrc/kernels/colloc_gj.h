// Register-resident Gauss-Jordan elimination with partial pivoting.
//
// An N x NCOL matrix [S | R] (columns >= N are right-hand sides) lives in the
// registers of a workgroup of W waves: lane l of wave w owns rows l + 64 q
// (q < RQ) and columns w + W cq (cq < CQ).  One step per matrix column k:
//   1. the wave owning column k finds the pivot (largest |a_rk| over rows not
//      yet used as pivots; ties go to the lower row, as LAPACK's idamax) and
//      writes column k to LDS;
//   2. every wave reads the multipliers from LDS and the pivot row from the
//      owning lane (v_readlane, so the pivot row is a scalar operand), scales
//      the pivot row by 1/pivot and removes column k from every other row.
// Columns < k of the matrix part hold unit vectors and are skipped (uniformly
// per wave).  Rows are never moved: pivrow[k] records the physical row that
// holds unknown k at the end.  One barrier per step (LDS column buffer is
// double-buffered).
#pragma once
#include "colloc_common.h"

namespace colloc {

template <int N, int NCOL, int W>
struct GJ {
  static constexpr int RQ = (N + WAVE - 1) / WAVE;
  static constexpr int CQ = (NCOL + W - 1) / W;
  static constexpr int CQS = (N + W - 1) / W;
  static constexpr int COLBUF = 2 * RQ * WAVE;   // doubles of LDS for the column buffer

  double a[RQ][CQ];

  // Global column index of column slot cq in this wave.
  __device__ __forceinline__ static int col(int cq) { return wave_id() + W * cq; }
  __device__ __forceinline__ static int row(int q) { return lane_id() + WAVE * q; }

  // Returns nonzero if a zero or non-finite pivot was met.
  // colbuf: COLBUF doubles of LDS; pivrow: N ints of LDS.
  __device__ int eliminate(double* colbuf, int* pivrow) {
    const int lane = lane_id();
    const int wave = wave_id();
    unsigned used = 0;
#pragma unroll
    for (int q = 0; q < RQ; ++q)
      if (row(q) >= N) used |= 1u << q;
    int bad = 0;
#pragma unroll
    for (int kc = 0; kc < CQS; ++kc) {
      for (int kw = 0; kw < W; ++kw) {
        const int k = kw + W * kc;
        if (k >= N) break;
        double* cb = colbuf + (k & 1) * (RQ * WAVE);
        if (wave == kw) {
          double best = -1.0;
          int brow = N;
#pragma unroll
          for (int q = 0; q < RQ; ++q) {
            const double v = a[q][kc];
            cb[lane + WAVE * q] = v;
            const double av = fabs(v);
            if (!((used >> q) & 1u) && av > best) { best = av; brow = row(q); }
          }
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int orow = __shfl_xor(brow, off);
            if (ob > best || (ob == best && orow < brow)) { best = ob; brow = orow; }
          }
          if (lane == 0) pivrow[k] = brow < N ? brow : 0;
        }
        __syncthreads();
        const int p = pivrow[k];
        const double piv = cb[p];
        if (!(fabs(piv) > 0.0) || !(fabs(piv) < 1.0e308)) bad = 1;
        const double inv = 1.0 / piv;
        const int pl = p & (WAVE - 1);
        const int pq = p >> 6;
        if (lane == pl) used |= 1u << pq;
        double l[RQ];
#pragma unroll
        for (int q = 0; q < RQ; ++q) l[q] = cb[lane + WAVE * q] * inv;
#pragma unroll
        for (int cq = 0; cq < CQ; ++cq) {
          const int c = wave + W * cq;
          if (c >= NCOL) continue;
          if (c < N && c <= k) continue;
          double pv = a[0][cq];
#pragma unroll
          for (int q = 1; q < RQ; ++q) pv = (pq == q) ? a[q][cq] : pv;
          const double prow = readlane_d(pv, pl);
#pragma unroll
          for (int q = 0; q < RQ; ++q) {
            const double upd = fma(-l[q], prow, a[q][cq]);
            a[q][cq] = (q == pq && lane == pl) ? prow * inv : upd;
          }
        }
      }
    }
    return bad;
  }
};

}  // namespace colloc
