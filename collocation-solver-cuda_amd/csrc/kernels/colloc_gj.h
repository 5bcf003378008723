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
  __device__ __forceinline__ int eliminate(double* colbuf, int* pivrow, double* pivval) {
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

namespace colloc {

// Blocked Gauss-Jordan elimination with partial pivoting.
//
// Same register layout as GJ (lane l of wave w owns rows l + 64 q and
// columns w + W cq of the N x NCOL matrix [S | R]).  The matrix columns are
// eliminated B at a time:
//   1. the owners copy the B panel columns to LDS (column-major, so that
//      lane-per-row accesses are free of bank conflicts);
//   2. all waves apply B Gauss-Jordan exchange steps to the panel in LDS,
//      wave w updating panel columns w, w + W, ...: at step kk every wave
//      finds the pivot row p among the rows not yet used (largest |a|, ties
//      to the lower row; the search is repeated in every wave so that no
//      broadcast is needed), row p is scaled by 1/a_p,kk, multiples of it are
//      subtracted from the other rows and column kk becomes the exchange
//      column (-a_r,kk / a_p,kk for r != p, 1 / a_p,kk for r = p).  The
//      pivot key keeps 12 mantissa bits of |a|: candidates within a relative
//      2^-12 of the largest are ties and go to the lower row.  The panel
//      is double-buffered by step parity, so a step reads one copy and writes
//      the other and needs one barrier.  After the B steps the panel holds
//      M = E[:, P], the columns of the accumulated row transformation E that
//      belong to the B pivot rows P;
//   3. every wave applies E to its trailing columns: the pivot rows' values
//      Tp (B x trailing) go to LDS, the pivot rows are cleared, and
//      a += M Tp is accumulated in registers from LDS broadcasts, four panel
//      columns at a time.
// The pivot rows end up scaled: row pivrow[k] holds unknown k's solution in
// the right-hand-side columns, so pivval[k] is 1.  Rows never move.
template <int N, int NCOL, int W, int B>
struct GJB {
  static constexpr int RQ = (N + WAVE - 1) / WAVE;
  static constexpr int CQ = (NCOL + W - 1) / W;
  static constexpr int NR = RQ * WAVE;            // rows held (N plus padding)
  static constexpr int PBUF = NR * B;             // doubles of LDS for one copy of the panel (column-major)
  static constexpr int TP = NCOL * B;             // doubles of LDS for the pivot rows
  static constexpr int LDS_DOUBLES = 2 * PBUF + TP;   // plus NR ints for the row flags
  static constexpr int PCW = (B + W - 1) / W;     // panel columns per wave in the sweep
  static_assert(N <= 256, "row indices are packed into 8 bits of the pivot key");
  static_assert(B % 4 == 0, "panel width must be a multiple of 4");

  double a[RQ][CQ];

  __device__ __forceinline__ static int col(int cq) { return wave_id() + W * cq; }
  __device__ __forceinline__ static int row(int q) { return lane_id() + WAVE * q; }

  // Copy matrix columns [k0, k0 + B) (those < N) to the panel buffer; panel
  // columns at or beyond N are written as zeros by wave 0.
  __device__ __forceinline__ void put_panel(int k0, double* pbuf) {
    const int wave = wave_id();
#pragma unroll
    for (int cq = 0; cq < CQ; ++cq) {
      const int c = wave + W * cq;
      if (c < k0 || c >= k0 + B || c >= N) continue;
#pragma unroll
      for (int q = 0; q < RQ; ++q) pbuf[(c - k0) * NR + row(q)] = a[q][cq];
    }
    if (wave == 0) {
      for (int c = (N - k0 > 0 ? N - k0 : 0); c < B; ++c)
#pragma unroll
        for (int q = 0; q < RQ; ++q) pbuf[c * NR + row(q)] = 0.0;
    }
  }

  // All waves: nb Gauss-Jordan exchange steps on the panel, which starts in
  // pb0 and ends in pb0 (nb even) or pb1 (nb odd).  Returns nonzero on a zero
  // or non-finite pivot.
  __device__ __forceinline__ int sweep(int k0, int nb, unsigned& used, double* pb0, double* pb1, int* flag,
                                       int* pivrow, double* pivval) {
    const int lane = lane_id();
    const int wave = wave_id();
    int bad = 0;
    for (int kk = 0; kk < nb; ++kk) {
      const double* src = (kk & 1) ? pb1 : pb0;
      double* dst = (kk & 1) ? pb0 : pb1;
      // every LDS read of the step is issued first: column kk and this wave's
      // own panel columns (the pivot row's values are then taken from
      // registers with v_readlane, so nothing waits on LDS after the search)
      double cv[RQ], own[PCW][RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) cv[q] = src[kk * NR + row(q)];
#pragma unroll
      for (int t = 0; t < PCW; ++t) {
        const int c = wave + W * t;
#pragma unroll
        for (int q = 0; q < RQ; ++q) own[t][q] = c < B ? src[c * NR + row(q)] : 0.0;
      }
      // pivot key: the high word of |a| (exponent and 20 mantissa bits) with
      // its low 8 bits replaced by 255 - row, so candidates within a relative
      // 2^-12 of the largest count as ties and go to the lower row
      unsigned key = 0;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const unsigned kq = ((unsigned)__double2hiint(fabs(cv[q])) & ~0xFFu) | (unsigned)(255 - row(q));
        if (!((used >> q) & 1u) && kq > key) key = kq;
      }
      key = wave_max_u32(key);
      const int p = 255 - (int)(key & 0xFF);
      const int pl = p & (WAVE - 1), pq = p >> 6;
      double sel = cv[0];
#pragma unroll
      for (int q = 1; q < RQ; ++q) sel = (pq == q) ? cv[q] : sel;
      const double piv = readlane_d(sel, pl);
      if (!(fabs(piv) > 0.0) || !(fabs(piv) < 1.0e308)) bad = 1;
      const double inv = fast_rcp(piv);
      double mlt[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) mlt[q] = (q == pq && lane == pl) ? 1.0 - inv : cv[q] * inv;
#pragma unroll
      for (int t = 0; t < PCW; ++t) {
        const int c = wave + W * t;
        if (c >= B) continue;
        double ps = own[t][0];
#pragma unroll
        for (int q = 1; q < RQ; ++q) ps = (pq == q) ? own[t][q] : ps;
        const double pv = readlane_d(ps, pl);
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          double v = fma(-mlt[q], pv, own[t][q]);
          if (c == kk) v = (q == pq && lane == pl) ? inv : -cv[q] * inv;
          dst[c * NR + row(q)] = v;
        }
      }
      if (lane == pl) used |= 1u << pq;
      if (wave == 0 && lane == 0) {
        pivrow[k0 + kk] = p;
        pivval[k0 + kk] = 1.0;
        flag[p] = kk + 1;
      }
      __syncthreads();
    }
    return bad;
  }

  // LDS: lds (LDS_DOUBLES doubles), flag (NR ints, zero on entry), pivrow
  // (N ints), pivval (N doubles).
  __device__ __forceinline__ int eliminate(double* lds, int* flag, int* pivrow, double* pivval) {
    double* pb0 = lds;
    double* pb1 = lds + PBUF;
    double* tp = lds + 2 * PBUF;
    const int wave = wave_id();
    unsigned used = 0;
#pragma unroll
    for (int q = 0; q < RQ; ++q)
      if (row(q) >= N) used |= 1u << q;
    int bad = 0;
    put_panel(0, pb0);
    __syncthreads();
    for (int k0 = 0; k0 < N; k0 += B) {
      const int nb = N - k0 < B ? N - k0 : B;
      bad |= sweep(k0, nb, used, pb0, pb1, flag, pivrow, pivval);
      const double* mb = (nb & 1) ? pb1 : pb0;
      // pivot rows: trailing values to tp, then cleared (their new value is M Tp)
      const int lo = k0 + nb;
      int fl[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) fl[q] = flag[row(q)];
#pragma unroll
      for (int cq = 0; cq < CQ; ++cq) {
        const int c = wave + W * cq;
        if (c < lo || c >= NCOL) continue;
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          if (fl[q]) tp[c * B + fl[q] - 1] = a[q][cq];
          a[q][cq] = fl[q] ? 0.0 : a[q][cq];
        }
        if (nb < B && lane_id() < B - nb) tp[c * B + nb + lane_id()] = 0.0;
      }
      __syncthreads();
      // a += M Tp, four panel columns at a time
      static_for<0, B / 4>([&](auto kc_) {
        constexpr int kc = decltype(kc_)::value;
        if (4 * kc < nb) {
          double L[RQ][4];
#pragma unroll
          for (int q = 0; q < RQ; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) L[q][e] = mb[(4 * kc + e) * NR + row(q)];
#pragma unroll
          for (int cq = 0; cq < CQ; ++cq) {
            const int c = wave + W * cq;
            if (c < lo || c >= NCOL) continue;
            double t4[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) t4[e] = tp[c * B + 4 * kc + e];
#pragma unroll
            for (int q = 0; q < RQ; ++q)
#pragma unroll
              for (int e = 0; e < 4; ++e) a[q][cq] = fma(L[q][e], t4[e], a[q][cq]);
          }
        }
      });
      __syncthreads();
      if (wave == 0 && lane_id() < nb) flag[pivrow[k0 + lane_id()]] = 0;
      if (k0 + B < N) put_panel(k0 + B, pb0);
      __syncthreads();
    }
    return bad;
  }
};

}  // namespace colloc

