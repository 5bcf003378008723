// Large-block kernels (ny > 64: the register-resident condensation and
// cyclic reduction of colloc_kernels.h hold at most 64 rows per lane set).
//
// Condensation (the same elimination as condense_ns, solver:840-1001
// restated): per interval,
//   1. colloc_l_stage_elim: one workgroup per (interval, stage j): the
//      algebraic block G_j = g_z^-1 [g_y | g_pa | f_a,g] by Gauss-Jordan with
//      partial pivoting in LDS, then T_j = [h_y | h_pa | f_a,h] - h_z G_j;
//   2. colloc_l_series_init: R (the right-hand sides of S U = R) from T;
//   3. colloc_l_series_term, once per term: U <- R + h (A^ x I) diag(M_k) U
//      with the products Y_k = M_k U_k on matrix cores (64 x 64 output tiles,
//      one workgroup per (interval, tile), every stage block's product in the
//      workgroup so the mixing needs no exchange); U alternates between the
//      record's U block and a second buffer; a term's largest change and
//      largest entry per interval go to slots that the next term reads, so
//      an interval that has converged is skipped;
//   4. colloc_l_series_out: the final U into the record, the BABD row.
// Cyclic reduction (solve_babd_system.py:93-258 restated as in cr_pair /
// cr_subst / cr_final / cr_back): the substitution product G_b [G_a | H_a |
// b_a] on matrix cores in 64 x 64 tiles; the Householder pairing on a work
// matrix in global memory (column-parallel reflector applications); the
// boundary system by Gauss-Jordan in global memory; back substitution per
// record with the triangular solve in LDS.
#pragma once

namespace colloc {

constexpr int NSK = M - 1;                       // stage blocks of the series
constexpr int L_TILE = 64;                       // output tile of the matrix-core products
constexpr int L_KC = 32;                         // k-chunk staged in LDS
constexpr int L_THREADS = 256;
constexpr int L_NRT = (NY + L_TILE - 1) / L_TILE;
constexpr int L_NCT = (NVA + L_TILE - 1) / L_TILE;
constexpr int L_TILES = L_NRT * L_NCT;
constexpr int L_MAX_TERMS = 48;
constexpr int L_SLOT = 10;                       // u64 per interval: 4 x (change, size), the converged term, -
constexpr int CS_XC = NY + NPH + 1;              // columns of [G_a | H_a | b_a]
constexpr int L_SCT = (CS_XC + L_TILE - 1) / L_TILE;
constexpr int L_STILES = L_NRT * L_SCT;
constexpr int L_WC = 3 * NY + NPH + 1;           // columns of a Householder pair's work matrix
constexpr int L_QR_THREADS = 512;
constexpr int L_FIN_THREADS = 512;

template <class R> __device__ __forceinline__ R l_tol();
template <> __device__ __forceinline__ double l_tol<double>() { return 0x1p-50; }
template <> __device__ __forceinline__ float l_tol<float>() { return 0x1p-21f; }

typedef float l_f4 __attribute__((ext_vector_type(4)));
typedef double l_d4 __attribute__((ext_vector_type(4)));
template <class T> struct MfmaOp;
// The 16 x 16 x 4 products take A[i][k] from lane i + 16 k and B[k][j] from
// lane j + 16 k; the result's element v in lane l is row row(l, v), column
// l % 16, where the float32 form holds four consecutive rows per lane and the
// float64 form rows four apart (measured on gfx950).
template <> struct MfmaOp<float> {
  typedef l_f4 V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct MfmaOp<double> {
  typedef l_d4 V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};

// LDS of tile_gemm: A chunk 64 x L_KC and B chunk L_KC x 64, padded rows.
constexpr int L_SA = L_TILE * (L_KC + 1);
constexpr int L_SB = L_KC * (L_TILE + 1);

// acc += A B for a 64 x 64 output tile, A (64 x K) and B (K x 64) read
// through A(r, k) and B(k, c) (which return 0 outside the matrices), staged
// in LDS in chunks of L_KC; the next chunk's global loads are issued before
// the current chunk's products, so their latency overlaps the matrix cores.
// Wave w owns rows 32 (w / 2) .. +32 and columns 32 (w % 2) .. +32 of the
// tile as 2 x 2 matrix-core tiles of 16 x 16: element v of acc[ti][tj] in
// lane l is row 32 (w / 2) + 16 ti + MfmaOp<T>::row(l, v), column
// 32 (w % 2) + 16 tj + l % 16.  Every thread of the workgroup of L_THREADS
// calls it.
constexpr int L_FA = L_TILE * L_KC / L_THREADS;   // A elements a thread stages per chunk
constexpr int L_FB = L_KC * L_TILE / L_THREADS;   // B elements
static_assert(L_FA * L_THREADS == L_TILE * L_KC && L_FB * L_THREADS == L_KC * L_TILE, "tile_gemm staging");
template <class T, class FA, class FB>
__device__ __forceinline__ void tile_gemm(int K, FA&& A, FB&& B, T* sA, T* sB,
                                          typename MfmaOp<T>::V (&acc)[2][2]) {
  const int tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
  T ra[L_FA], rb[L_FB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int it = 0; it < L_FA; ++it) {
      const int e = tid + it * L_THREADS;
      const int r = e / L_KC, k = e - (e / L_KC) * L_KC;
      ra[it] = (k0 + k < K) ? A(r, k0 + k) : (T)0;
    }
#pragma unroll
    for (int it = 0; it < L_FB; ++it) {
      const int e = tid + it * L_THREADS;
      const int k = e / L_TILE, c = e - (e / L_TILE) * L_TILE;
      rb[it] = (k0 + k < K) ? B(k0 + k, c) : (T)0;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += L_KC) {
    __syncthreads();   // the previous chunk's reads are done
#pragma unroll
    for (int it = 0; it < L_FA; ++it) {
      const int e = tid + it * L_THREADS;
      sA[(e / L_KC) * (L_KC + 1) + (e - (e / L_KC) * L_KC)] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < L_FB; ++it) {
      const int e = tid + it * L_THREADS;
      sB[(e / L_TILE) * (L_TILE + 1) + (e - (e / L_TILE) * L_TILE)] = rb[it];
    }
    __syncthreads();
    if (k0 + L_KC < K) fetch(k0 + L_KC);
#pragma unroll
    for (int s = 0; s < L_KC / 4; ++s) {
      T a[2], b[2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) a[ti] = sA[(r0 + 16 * ti + (lane & 15)) * (L_KC + 1) + 4 * s + (lane >> 4)];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) b[tj] = sB[(4 * s + (lane >> 4)) * (L_TILE + 1) + c0 + 16 * tj + (lane & 15)];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = MfmaOp<T>::op(a[ti], b[tj], acc[ti][tj]);
    }
  }
}

// Workgroup max of two non-negative values folded into u64 slots (as the
// bits of doubles; a NaN, sign cleared, wins over any number).
__device__ __forceinline__ void l_block_max2(double d, double x, unsigned long long* sd, unsigned long long* sx) {
  __shared__ double red[2][L_THREADS / WAVE];
  d = wave_max(d);
  x = wave_max(x);
  if (lane_id() == 0) {
    red[0][wave_id()] = d;
    red[1][wave_id()] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / WAVE;
    for (int w = 1; w < nw; ++w) {
      d = nmax(red[0][w], d);
      x = nmax(red[1][w], x);
    }
    atomicMax(sd, (unsigned long long)__double_as_longlong(fabs(d)));
    atomicMax(sx, (unsigned long long)__double_as_longlong(fabs(x)));
  }
}

__device__ __forceinline__ double slot_d(unsigned long long v) { return __longlong_as_double((long long)v); }

// XCD-aware workgroup order: the hardware deals consecutive workgroups to
// the 8 XCDs in turn, each with its own L2.  Workgroup g becomes logical
// work unit L(g) such that the units an XCD runs are consecutive, so the
// tiles of one interval (or pair), which read the same operand blocks, share
// that XCD's L2.  A bijection of [0, G).
constexpr int L_XCDS = 8;
__device__ __forceinline__ int xcd_unit(int g, int G) {
  const int x = g % L_XCDS, slot = g / L_XCDS;
  const int q = G / L_XCDS, rem = G - q * L_XCDS;
  const int base = x * q + (x < rem ? x : rem);
  return base + slot;
}

// ------------------------------------------------------ stage eliminations
// One workgroup per (listed interval, stage j), computing in R.  The
// algebraic block: every wave holds [g_z | its quarter of g_y g_pa f_a,g]
// in registers (lane = column, the NZ rows per lane) and runs the same
// Gauss-Jordan with partial pivoting (the largest |entry| of the column,
// the first on ties; the pivot column broadcast by v_readlane), so the four
// waves take identical pivot decisions with no exchange; G_j goes to LDS and
// the record.  Then T_j = [h_y | h_pa | f_a,h] - h_z G_j on matrix cores.
// Status bit 1 on a zero pivot.
constexpr int L_GZW = NVA + 1;                            // padded row of G_j in LDS
constexpr int L_SE_CPW = (NVA + 3) / 4;                   // right-hand-side columns per wave
constexpr int L_SE_SL = (NZ + L_SE_CPW + WAVE - 1) / WAVE; // register slices of a wave's columns
template <class R>
__device__ __forceinline__ void l_stage_elim(int nint, Batch bt, const R* __restrict__ jac,
                                             const double* __restrict__ fa, R* __restrict__ rec,
                                             int* __restrict__ status) {
  __shared__ R gz[NZ1 * L_GZW];   // G_j
  __shared__ __attribute__((aligned(16))) R sA[L_SA];
  __shared__ __attribute__((aligned(16))) R sB[L_SB];
  const int item = blockIdx.x / M, j = blockIdx.x - (blockIdx.x / M) * M;
  if (item >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, item);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  const int tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const R* Dfj = jac + ((size_t)i * M + j) * JSTAGE;
  const R* Dgj = Dfj + NY * NJC;
  const double* faj = fa + (size_t)i * M * NYZ + j * NYZ;
  R* recT = rec + (size_t)i * REC + NS * NVA;
  R* recG = recT + M * NY * NVA;
  int bad = 0;
  if (NZ > 0) {
    // column of slice sl of this lane: g_z column (< NZ) or right-hand side q
    R gc[L_SE_SL][NZ1];
    int qcol[L_SE_SL];
#pragma unroll
    for (int sl = 0; sl < L_SE_SL; ++sl) {
      const int cc = lane + WAVE * sl;
      const int q = cc < NZ ? -1 : w * L_SE_CPW + (cc - NZ);
      const bool isq = cc >= NZ && cc - NZ < L_SE_CPW && q < NVA;
      qcol[sl] = isq ? q : -2;
#pragma unroll
      for (int r = 0; r < NZ; ++r) {
        R v = (R)0;
        if (cc < NZ) v = Dgj[r * NJC + NY + cc];
        else if (isq) v = q < NY ? Dgj[r * NJC + q] : (q < NY + NPA ? Dgj[r * NJC + NY + NZ + pact(q - NY)] : (R)faj[NY + r]);
        gc[sl][r] = v;
      }
    }
    // every step at compile time (static_for), so that the register arrays
    // are indexed by constants
    static_for<0, NZ>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      constexpr int ks = k / WAVE, kl = k % WAVE;
      // pivot: the lane holding column k finds the largest |entry| at or below
      // row k (the first on ties); every lane learns the row
      int prl = k;
      {
        R best = fabs(gc[ks][k]);
#pragma unroll
        for (int r = k + 1; r < NZ; ++r)
          if (fabs(gc[ks][r]) > best) { best = fabs(gc[ks][r]); prl = r; }
      }
      const int pr = __builtin_amdgcn_readlane(prl, kl);
#pragma unroll
      for (int r = k + 1; r < NZ; ++r) {
        const bool sw = r == pr;
#pragma unroll
        for (int sl = 0; sl < L_SE_SL; ++sl) {
          const R t1 = gc[sl][r];
          gc[sl][r] = sw ? gc[sl][k] : gc[sl][r];
          gc[sl][k] = sw ? t1 : gc[sl][k];
        }
      }
      const R cpr = readlane_r(gc[ks][k], kl);
      if (!(fabs(cpr) > (R)0)) bad = 1;
      const R inv = (R)1 / cpr;
      R pk[L_SE_SL];
#pragma unroll
      for (int sl = 0; sl < L_SE_SL; ++sl) pk[sl] = gc[sl][k] * inv;
      // eliminate column k from the other rows, eight multipliers at a time
#pragma unroll
      for (int r0 = 0; r0 < NZ; r0 += 8) {
        R cr[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) cr[q] = (r0 + q < NZ) ? readlane_r(gc[ks][r0 + q < NZ ? r0 + q : 0], kl) : (R)0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int r = r0 + q;
          if (r >= NZ || r == k) continue;
#pragma unroll
          for (int sl = 0; sl < L_SE_SL; ++sl) gc[sl][r] = fma(-cr[q], pk[sl], gc[sl][r]);
        }
      }
#pragma unroll
      for (int sl = 0; sl < L_SE_SL; ++sl) gc[sl][k] = pk[sl];
    });
    // [I | G_j]: this wave's columns of G_j
#pragma unroll
    for (int sl = 0; sl < L_SE_SL; ++sl) {
      const int q = qcol[sl];
      if (q < 0) continue;
#pragma unroll
      for (int r = 0; r < NZ; ++r) {
        gz[r * L_GZW + q] = gc[sl][r];
        recG[(j * NZ + r) * NVA + q] = gc[sl][r];
      }
    }
  }
  __syncthreads();
  // T_j = [h_y | h_pa | f_a,h] - h_z G_j: the product on matrix cores in
  // 64 x 64 tiles (K = NZ)
  for (int tile = 0; tile < L_TILES; ++tile) {
    const int rbase = (tile / L_NCT) * L_TILE, cbase = (tile - (tile / L_NCT) * L_NCT) * L_TILE;
    typename MfmaOp<R>::V acc[2][2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[ti][tj][v] = (R)0;
    if (NZ > 0)
      tile_gemm<R>(NZ,
          [&](int r, int z) { return rbase + r < NY ? Dfj[(rbase + r) * NJC + NY + z] : (R)0; },
          [&](int z, int c) { return cbase + c < NVA ? gz[z * L_GZW + cbase + c] : (R)0; },
          sA, sB, acc);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = rbase + 32 * (w >> 1) + 16 * ti + MfmaOp<R>::row(lane, v);
          const int q = cbase + 32 * (w & 1) + 16 * tj + (lane & 15);
          if (r >= NY || q >= NVA) continue;
          const R hb = q < NY ? Dfj[r * NJC + q] : (q < NY + NPA ? Dfj[r * NJC + NY + NZ + pact(q - NY)] : (R)faj[r]);
          recT[(j * NY + r) * NVA + q] = hb - acc[ti][tj][v];
        }
  }
  if (bad && tid == 0) atomicOr(status, 1);
}

// ---------------------------------------------------------------- series
// R_k (k = 0 .. NSK-1, stage k+1) of the listed intervals:
//   R_k = [I | 0] + h A[k+1][0] T_0 + sum_{j>=1} h A[k+1][j] [0 | T_j[:, NY:]],
// written to the record's U block (the series' first iterate).
// One workgroup per (listed interval, stage block); slots cleared.
template <class R>
__device__ __forceinline__ void l_series_init(int nint, Batch bt, const double* __restrict__ t, R* __restrict__ rec,
                                              unsigned long long* __restrict__ slots) {
  const int item = blockIdx.x / NSK, k = blockIdx.x - (blockIdx.x / NSK) * NSK;
  if (item >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, item);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  const R h = (R)(t[i + ib + 1] - t[i + ib]);
  R* recU = rec + (size_t)i * REC;
  const R* recT = recU + NS * NVA;
  const int js = k + 1;
  for (int e = threadIdx.x; e < NY * NVA; e += blockDim.x) {
    const int r = e / NVA, c = e - (e / NVA) * NVA;
    R v = h * (R)c_rkA[js * M] * recT[r * NVA + c];
    if (c < NY) {
      if (c == r) v += (R)1;
    } else {
      for (int jj = 1; jj < M; ++jj) v += h * (R)c_rkA[js * M + jj] * recT[(jj * NY + r) * NVA + c];
    }
    recU[(k * NY + r) * NVA + c] = v;
  }
  if (k == 0 && threadIdx.x < L_SLOT) slots[(size_t)i * L_SLOT + threadIdx.x] = 0;
}

// The stopping rule of ns_stop (colloc_kernels.h) on the slots after term
// `term`: its change within the tolerance, or, in float32 from the second
// term on, a contraction against the previous term's change that bounds the
// remaining terms by a quarter of the tolerance.
template <class R>
__device__ __forceinline__ bool l_stop(const unsigned long long* s, int term, double tol) {
  const int set = term % 4;
  const double d = slot_d(s[2 * set]), x = slot_d(s[2 * set + 1]);
  if (d <= tol * x) return true;
  if constexpr (sizeof(R) == 4) {
    if (term >= 2) {
      const double dp = slot_d(s[2 * ((term - 1) % 4)]);
      return dp > 0.0 && d < 0.5 * dp && 8.0 * d * d <= tol * x * dp;
    }
  }
  return false;
}

// Term `term` (>= 1) of the series for tile (rt, ct) of the listed interval:
// Unew_k = R_k + sum_q h A[k+1][q+1] M_{q+1} Uold_q on rows rt, columns ct.
// Uold / Unew: the U blocks (stride ustride_old / ustride_new per interval);
// slots[i]: [4 x (largest change, largest entry) by term mod 4 | converged
// term | -]; a term reads the sets of the two terms before it, writes its
// own and clears the next term's; active[term]: intervals that computed
// this term.
template <class R>
__device__ __forceinline__ void l_series_term(int nint, Batch bt, const double* __restrict__ t,
                                              const R* __restrict__ rec,
                                              const R* __restrict__ uold, long long ustride_old,
                                              R* __restrict__ unew, long long ustride_new,
                                              unsigned long long* __restrict__ slots,
                                              unsigned long long* __restrict__ active, int term) {
  __shared__ __attribute__((aligned(16))) R sA[L_SA];
  __shared__ __attribute__((aligned(16))) R sB[L_SB];
  const int unit = xcd_unit(blockIdx.x, gridDim.x);
  const int item = unit / L_TILES, tile = unit - (unit / L_TILES) * L_TILES;
  if (item >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, item);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  unsigned long long* s = slots + (size_t)i * L_SLOT;
  const double tol = (double)l_tol<R>();
  if (term >= 2 && l_stop<R>(s, term - 1, tol)) {
    if (tile == 0 && threadIdx.x == 0 && s[8] == 0) s[8] = (unsigned long long)(term - 1);
    return;
  }
  if (tile == 0 && threadIdx.x == 0) {
    const int nx = (term + 1) % 4;
    s[2 * nx] = 0;
    s[2 * nx + 1] = 0;
    atomicAdd(active + term, 1ull);
  }
  const int rt = tile / L_NCT, ct = tile - (tile / L_NCT) * L_NCT;
  const int rbase = rt * L_TILE, cbase = ct * L_TILE;
  const R* recT = rec + (size_t)i * REC + NS * NVA;
  const R* uo = uold + (size_t)i * ustride_old;
  typename MfmaOp<R>::V acc[NSK][2][2];
#pragma unroll
  for (int q = 0; q < NSK; ++q)
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[q][ti][tj][v] = (R)0;
#pragma unroll
  for (int q = 0; q < NSK; ++q) {
    const R* Mq = recT + (size_t)(q + 1) * NY * NVA;   // M_{q+1} = T_{q+1}[:, :NY]
    const R* Uq = uo + (size_t)q * NY * NVA;
    tile_gemm<R>(NY,
        [&](int r, int k) { return rbase + r < NY ? Mq[(rbase + r) * NVA + k] : (R)0; },
        [&](int k, int c) { return cbase + c < NVA ? Uq[k * NVA + cbase + c] : (R)0; },
        sA, sB, acc[q]);
  }
  const R h = (R)(t[i + ib + 1] - t[i + ib]);
  const int lane = lane_id(), w = wave_id();
  R* un = unew + (size_t)i * ustride_new;
  double dmax = 0.0, xmax = 0.0;
  // R_k = [I | 0] + h A[k+1][0] T_0 + sum_{j>=1} h A[k+1][j] [0 | T_j[:, NY:]],
  // formed here from T_0 (one read for the NSK blocks) instead of stored
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rbase + 32 * (w >> 1) + 16 * ti + MfmaOp<R>::row(lane, v);
        const int col = cbase + 32 * (w & 1) + 16 * tj + (lane & 15);
        if (row >= NY || col >= NVA) continue;
        const R t0 = recT[row * NVA + col];
#pragma unroll
        for (int k = 0; k < NSK; ++k) {
          R x = h * (R)c_rkA[(k + 1) * M] * t0;
          if (col < NY) {
            if (col == row) x += (R)1;
          } else {
            for (int jj = 1; jj < M; ++jj) x += h * (R)c_rkA[(k + 1) * M + jj] * recT[(jj * NY + row) * NVA + col];
          }
#pragma unroll
          for (int q = 0; q < NSK; ++q) x = fma(h * (R)c_rkA[(k + 1) * M + q + 1], acc[q][ti][tj][v], x);
          const size_t e = (size_t)(k * NY + row) * NVA + col;
          dmax = nmax(fabs((double)x - (double)uo[e]), dmax);
          xmax = nmax(fabs((double)x), xmax);
          un[e] = x;
        }
      }
  l_block_max2(dmax, xmax, s + 2 * (term % 4), s + 2 * (term % 4) + 1);
}

// After `nterms` terms: the term at which each listed interval converged
// (its U is the record's U block for even terms, ubuf's for odd ones), the
// final U into the record, the BABD row [G | C = I | H | b] from the last
// stage block; counters[1] += terms, counters[2] += intervals not converged
// (left without a row: the runtime reports them).
template <class R>
__device__ __forceinline__ void l_series_out(int nint, Batch bt, const double* __restrict__ fb,
                                             double* __restrict__ rows, R* __restrict__ rec,
                                             const R* __restrict__ ubuf, const unsigned long long* __restrict__ slots,
                                             int nterms, unsigned long long* __restrict__ counters) {
  __shared__ int sterm;
  if ((int)blockIdx.x >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, blockIdx.x);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  const unsigned long long* s = slots + (size_t)i * L_SLOT;
  if (threadIdx.x == 0) {
    int tc = (int)s[8];
    if (tc == 0 && nterms >= 1 && l_stop<R>(s, nterms, (double)l_tol<R>())) tc = nterms;
    sterm = tc;
    if (tc == 0) atomicAdd(counters + 2, 1ull);
    else atomicAdd(counters + 1, (unsigned long long)tc);
  }
  __syncthreads();
  const int tc = sterm;
  if (tc == 0) return;
  R* recU = rec + (size_t)i * REC;
  if (tc & 1) {
    const R* src = ubuf + (size_t)i * NS * NVA;
    for (int e = threadIdx.x; e < NS * NVA; e += blockDim.x) recU[e] = src[e];
  }
  const R* ul = ((tc & 1) ? ubuf + (size_t)i * NS * NVA : recU) + (size_t)(NSK - 1) * NY * NVA;
  double* rowi = rows + (size_t)i * NY * CRW;
  for (int e = threadIdx.x; e < NY * CRW; e += blockDim.x) {
    const int r = e / CRW, c = e - (e / CRW) * CRW;
    double v;
    if (c < NY) v = -(double)ul[r * NVA + c];
    else if (c < 2 * NY) v = (c - NY == r) ? 1.0 : 0.0;
    else if (c < CRW - 1) {
      if (HZ) continue;
      int a = -1;
      for (int q = 0; q < NPA; ++q)
        if (pact(q) == c - 2 * NY) a = q;
      v = a >= 0 ? -(double)ul[r * NVA + NY + a] : 0.0;
    } else {
      v = (double)ul[r * NVA + NVA - 1] - fb[(size_t)i * NY + r];
    }
    rowi[e] = v;
  }
}

// ------------------------------------------------------ cyclic reduction
// Substitution pairing (see cr_subst): out = [-G_b G_a | C_b | H_b - G_b H_a
// | b_b - G_b b_a] for tile (rt, ct) of the pair's NY x (NY + NPH + 1)
// product; workgroup = (output row of the level, tile).  TI / TO: stored
// type of the input / output rows; the product is computed in TO (float32
// matrix cores for the FP32 cyclic reduction).
template <class TI, class TO>
__device__ __forceinline__ void l_cr_subst(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                           const int* __restrict__ act, const TI* __restrict__ rows_in,
                                           TO* __restrict__ rows_out) {
  __shared__ __attribute__((aligned(16))) TO sA[L_SA];
  __shared__ __attribute__((aligned(16))) TO sB[L_SB];
  const int unit = xcd_unit(blockIdx.x, gridDim.x);
  const int wg = unit / L_STILES, tile = unit - (unit / L_STILES) * L_STILES;
  const int ib = oinst[wg];
  if (act && !act[ib]) return;
  const CRSeg sg = seg[ib];
  if (!sg.subst) return;
  const int k = wg - sg.out_off;
  const int nrows = sg.nrows;
  rows_in += (size_t)sg.in_off * NY * CRW;
  rows_out += (size_t)sg.out_off * NY * CRW;
  const int npairs = nrows >> 1;
  const bool wc = sg.subst == 1;
  const int rt = tile / L_SCT, ct = tile - (tile / L_SCT) * L_SCT;
  if (k == npairs) {   // carry the odd last row (rows of tile rt, columns of tile ct's share)
    if (ct != 0) return;
    const TI* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    TO* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = threadIdx.x; e < L_TILE * CRW; e += blockDim.x) {
      const int r = rt * L_TILE + e / CRW, c = e - (e / CRW) * CRW;
      if (r >= NY) break;
      const size_t x = (size_t)r * CRW + c;
      if (c < NY || c == CRW - 1 || (!HZ && c >= 2 * NY)) dst[x] = (TO)src[x];
      else if (wc && c < 2 * NY) dst[x] = (c - NY == r) ? (TO)1 : (TO)0;
    }
    return;
  }
  const TI* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const TI* rb = ra + NY * CRW;
  TO* out = rows_out + (size_t)k * NY * CRW;
  const int rbase = rt * L_TILE, cbase = ct * L_TILE;
  typename MfmaOp<TO>::V acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[ti][tj][v] = (TO)0;
  tile_gemm<TO>(NY,
      [&](int r, int q) { return rbase + r < NY ? (TO)rb[(size_t)(rbase + r) * CRW + q] : (TO)0; },
      [&](int q, int c) {
        const int col = cbase + c;
        if (col < NY) return (TO)ra[(size_t)q * CRW + col];                 // G_a
        if (col < NY + NPH) return (TO)ra[(size_t)q * CRW + NY + col];      // H_a (after C_a)
        if (col == NY + NPH) return (TO)ra[(size_t)q * CRW + CRW - 1];      // b_a
        return (TO)0;
      },
      sA, sB, acc);
  const int lane = lane_id(), w = wave_id();
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rbase + 32 * (w >> 1) + 16 * ti + MfmaOp<TO>::row(lane, v);
        const int col = cbase + 32 * (w & 1) + 16 * tj + (lane & 15);
        if (r >= NY) continue;
        const TO a = acc[ti][tj][v];
        if (col < NY) out[(size_t)r * CRW + col] = -a;
        else if (col < NY + NPH) out[(size_t)r * CRW + NY + col] = (TO)rb[(size_t)r * CRW + NY + col] - a;
        else if (col == NY + NPH) out[(size_t)r * CRW + CRW - 1] = (TO)rb[(size_t)r * CRW + CRW - 1] - a;
      }
  if (wc && ct == 0) {
    for (int e = threadIdx.x; e < L_TILE * NY; e += blockDim.x) {
      const int r = rbase + e / NY, c = e - (e / NY) * NY;
      if (r < NY) out[(size_t)r * CRW + NY + c] = (r == c) ? (TO)1 : (TO)0;
    }
  }
}

// Householder pairing (see cr_pair; the reference's qr,
// matrix_factorization_cuda.py:5-82: g = -sign(v_k) |v|, stop at |v|^2 <
// eps): the pair's 2NY x L_WC block [x_M | x_L | x_R | p | rhs] in work
// (global memory), one workgroup of L_QR_THREADS; thread (column slot,
// row quarter) applies each reflector to a column's quarter, the quarters'
// dot products summed through LDS.  TI / TO: stored type of the input rows
// / of the output rows and records (the arithmetic is float64 either way).
template <class TI, class TO>
__device__ __forceinline__ void l_cr_pair_global(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                          const int* __restrict__ act, const TI* __restrict__ rows_in,
                                          TO* __restrict__ rows_out, TO* __restrict__ recs,
                                          double* __restrict__ work) {
  constexpr int R2 = 2 * NY;
  constexpr int NQ = 4;                          // row quarters
  constexpr int CSL = L_QR_THREADS / NQ;         // column slots
  constexpr int RQ = (R2 + NQ - 1) / NQ;         // rows per quarter
  __shared__ double sv[R2];
  __shared__ double wpart[NQ][CSL];
  __shared__ double red[L_QR_THREADS / WAVE];
  __shared__ double sbeta, sg_, stf;
  const int ib = oinst[blockIdx.x];
  if (act && !act[ib]) return;
  const CRSeg sg = seg[ib];
  if (sg.subst) return;
  const int k = blockIdx.x - sg.out_off;
  const int nrows = sg.nrows;
  rows_in += (size_t)sg.in_off * NY * CRW;
  rows_out += (size_t)sg.out_off * NY * CRW;
  recs += (size_t)sg.rec_off * NY * RECW;
  const int npairs = nrows >> 1;
  const int tid = threadIdx.x;
  if (k == npairs) {
    const TI* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    TO* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = tid; e < NY * CRW; e += blockDim.x) dst[e] = (TO)src[e];
    return;
  }
  double* W = work + (size_t)blockIdx.x * R2 * L_WC;
  const TI* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const TI* rb = ra + NY * CRW;
  for (int e = tid; e < R2 * L_WC; e += blockDim.x) {
    const int r = e / L_WC, c = e - (e / L_WC) * L_WC;
    const bool top = r < NY;
    const TI* src = top ? ra + (size_t)r * CRW : rb + (size_t)(r - NY) * CRW;
    const int pcol = c < 3 * NY + NPH ? c : RECW - 1;
    double v;
    if (c < NY) v = top ? (double)src[NY + c] : (double)src[c];
    else if (c < 2 * NY) v = top ? (double)src[c - NY] : 0.0;
    else if (c < 3 * NY) v = top ? 0.0 : (double)src[c - NY];
    else v = (double)src[pcol - NY];
    W[e] = v;
  }
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  const int cs = tid % CSL, qtr = tid / CSL;
  const int r0 = qtr * RQ, r1 = min(R2, r0 + RQ);
  for (int kk = 0; kk < NY; ++kk) {
    // |column kk below the pivot|^2 and the pivot entry
    double part = 0.0;
    if (tid >= kk && tid < R2) {
      const double a = W[(size_t)tid * L_WC + kk];
      part = a * a;
    }
    part = wave_sum(part);
    if (lane_id() == 0) red[wave_id()] = part;
    __syncthreads();
    if (tid == 0) {
      double tf = 0.0;
      for (int w = 0; w < L_QR_THREADS / WAVE; ++w) tf += red[w];
      const double akk = W[(size_t)kk * L_WC + kk];
      const double nrm = sqrt(tf);
      const double g = akk >= 0.0 ? -nrm : nrm;
      const double v0 = akk - g;
      const double vs = tf - akk * akk + v0 * v0;
      sbeta = vs < eps ? 0.0 : -2.0 / vs;
      sg_ = g;
      stf = v0;
    }
    __syncthreads();
    const double beta = sbeta;
    if (beta == 0.0) break;   // the reference stops reflecting here (uniform)
    if (tid < R2) sv[tid] = tid == kk ? stf : (tid > kk ? W[(size_t)tid * L_WC + kk] : 0.0);
    __syncthreads();
    if (tid < R2 && tid >= kk) W[(size_t)tid * L_WC + kk] = tid == kk ? sg_ : 0.0;
    for (int c0 = kk + 1; c0 < L_WC; c0 += CSL) {
      const int c = c0 + cs;
      double wsum = 0.0;
      if (c < L_WC)
        for (int r = max(r0, kk); r < r1; ++r) wsum = fma(sv[r], W[(size_t)r * L_WC + c], wsum);
      wpart[qtr][cs] = wsum;
      __syncthreads();
      if (c < L_WC) {
        double wt = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) wt += wpart[q][cs];
        wt *= beta;
        for (int r = max(r0, kk); r < r1; ++r) W[(size_t)r * L_WC + c] = fma(wt, sv[r], W[(size_t)r * L_WC + c]);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  TO* rec = recs + (size_t)k * NY * RECW;
  TO* out = rows_out + (size_t)k * NY * CRW;
  for (int e = tid; e < R2 * L_WC; e += blockDim.x) {
    const int r = e / L_WC, c = e - (e / L_WC) * L_WC;
    const int pcol = c < 3 * NY + NPH ? c : RECW - 1;
    if (r < NY) rec[(size_t)r * RECW + pcol] = (TO)W[e];
    else if (c >= NY) out[(size_t)(r - NY) * CRW + pcol - NY] = (TO)W[e];
  }
}

// Householder pairing by panels (ny <= 192): the pair's 2NY x L_WC block
// [x_M | x_L | x_R | p | rhs] (work, device memory) is reduced L_PW
// columns at a time.  Each panel is loaded to LDS and factored there (the
// reference's qr step per column, matrix_factorization_cuda.py:5-82: g =
// -sign(v_k) |v|, beta = -2 / |v|^2, stop at |v|^2 < eps), its reflectors
// kept in LDS; then the panel's reflectors are applied in order to every
// column after it, each column held in registers by four lanes (one per
// 16-lane row of a wave, NY / 2 rows each), the dot products summed across
// the four by lane exchanges.  The trailing columns are read and written
// once per panel instead of twice per reflector.
constexpr int L_PW = 16;                          // panel width
constexpr int L_PARTS = 8;                        // lanes per trailing column (one per 8-lane group of a wave)
constexpr int L_RPL = (2 * NY + L_PARTS - 1) / L_PARTS;   // rows per lane of a trailing column
template <int RPL, class TI, class TO>
__device__ __forceinline__ void l_cr_pair_panels(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                                 const int* __restrict__ act, const TI* __restrict__ rows_in,
                                                 TO* __restrict__ rows_out, TO* __restrict__ recs,
                                                 double* __restrict__ work) {
  constexpr int R2 = 2 * NY;
  constexpr int NW = L_QR_THREADS / WAVE;
  __shared__ double Pn[R2 * (L_PW + 1)];      // the panel, padded rows
  __shared__ double V[L_PW * R2];             // its reflectors (zero above their pivot row)
  __shared__ double Bt[L_PW];                 // and their betas
  __shared__ double red[NW];
  __shared__ double sdot[L_PW];
  __shared__ int snref, sstop;
  const int ib = oinst[blockIdx.x];
  if (act && !act[ib]) return;
  const CRSeg sg = seg[ib];
  if (sg.subst) return;
  const int k = blockIdx.x - sg.out_off;
  const int nrows = sg.nrows;
  rows_in += (size_t)sg.in_off * NY * CRW;
  rows_out += (size_t)sg.out_off * NY * CRW;
  recs += (size_t)sg.rec_off * NY * RECW;
  const int npairs = nrows >> 1;
  const int tid = threadIdx.x, lane = lane_id(), w = wave_id();
  if (k == npairs) {
    const TI* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    TO* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = tid; e < NY * CRW; e += blockDim.x) dst[e] = (TO)src[e];
    return;
  }
  double* W = work + (size_t)blockIdx.x * R2 * L_WC;
  const TI* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const TI* rb = ra + NY * CRW;
  for (int e = tid; e < R2 * L_WC; e += blockDim.x) {
    const int r = e / L_WC, c = e - (e / L_WC) * L_WC;
    const bool top = r < NY;
    const TI* src = top ? ra + (size_t)r * CRW : rb + (size_t)(r - NY) * CRW;
    const int pcol = c < 3 * NY + NPH ? c : RECW - 1;
    double v;
    if (c < NY) v = top ? (double)src[NY + c] : (double)src[c];
    else if (c < 2 * NY) v = top ? (double)src[c - NY] : 0.0;
    else if (c < 3 * NY) v = top ? 0.0 : (double)src[c - NY];
    else v = (double)src[pcol - NY];
    W[e] = v;
  }
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  const int part = lane >> 3, c8 = lane & 7;
  for (int kk0 = 0; kk0 < NY; kk0 += L_PW) {
    const int npw = min(L_PW, NY - kk0);
    for (int e = tid; e < R2 * L_PW; e += blockDim.x) {
      const int r = e / L_PW, j = e - (e / L_PW) * L_PW;
      Pn[r * (L_PW + 1) + j] = j < npw ? W[(size_t)r * L_WC + kk0 + j] : 0.0;
    }
    if (tid == 0) { snref = 0; sstop = 0; }
    __syncthreads();
    // factor the panel: column j's reflector, then its application to the
    // panel's later columns (wave w takes columns j + 1 + w, + NW, ...)
    for (int j = 0; j < npw; ++j) {
      const int kk = kk0 + j;
      double part2 = 0.0;
      for (int r = kk + tid; r < R2; r += blockDim.x) {
        const double a = Pn[r * (L_PW + 1) + j];
        part2 = fma(a, a, part2);
      }
      part2 = wave_sum(part2);
      if (lane == 0) red[w] = part2;
      __syncthreads();
      if (tid == 0) {
        double tf = 0.0;
        for (int q = 0; q < NW; ++q) tf += red[q];
        const double akk = Pn[kk * (L_PW + 1) + j];
        const double nrm = sqrt(tf);
        const double g = akk >= 0.0 ? -nrm : nrm;
        const double v0 = akk - g;
        const double vs = tf - akk * akk + v0 * v0;
        if (vs < eps) {
          sstop = 1;   // the reference stops reflecting here
        } else {
          Bt[j] = -2.0 / vs;
          sdot[0] = g;
          sdot[1] = v0;
          snref = j + 1;
        }
      }
      __syncthreads();
      if (sstop) break;
      const double g = sdot[0], v0 = sdot[1], beta = Bt[j];
      for (int r = tid; r < R2; r += blockDim.x)
        V[j * R2 + r] = r == kk ? v0 : (r > kk ? Pn[r * (L_PW + 1) + j] : 0.0);
      __syncthreads();
      for (int jj = j + 1 + w; jj < npw; jj += NW) {
        double d = 0.0;
        for (int r = kk + lane; r < R2; r += WAVE) d = fma(V[j * R2 + r], Pn[r * (L_PW + 1) + jj], d);
        d = wave_sum(d) * beta;
        for (int r = kk + lane; r < R2; r += WAVE) Pn[r * (L_PW + 1) + jj] = fma(d, V[j * R2 + r], Pn[r * (L_PW + 1) + jj]);
      }
      for (int r = kk + tid; r < R2; r += blockDim.x) Pn[r * (L_PW + 1) + j] = r == kk ? g : 0.0;
      __syncthreads();
    }
    const int nref = snref;
    const bool stop = sstop != 0;
    // the factored panel columns back (those reflected; a stopped column and
    // the ones after it keep their values)
    for (int e = tid; e < R2 * L_PW; e += blockDim.x) {
      const int r = e / L_PW, j = e - (e / L_PW) * L_PW;
      if (j < nref) W[(size_t)r * L_WC + kk0 + j] = Pn[r * (L_PW + 1) + j];
      else if (stop && j < npw) {
        // not reflected: left as it is in W (columns >= the stopped one get no
        // further reflectors either)
      }
    }
    // the panel's reflectors on the trailing columns: the panel's unreflected
    // columns (after a stop) and every column after the panel
    const int cfirst = stop ? kk0 + nref : kk0 + npw;
    if (nref > 0) {
      for (int c0 = cfirst; c0 < L_WC; c0 += NW * 8) {
        const int c = c0 + w * 8 + c8;
        const bool live = c < L_WC;
        double a[RPL];
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
          const int r = part * RPL + i;
          a[i] = (live && r < R2) ? W[(size_t)r * L_WC + c] : 0.0;
        }
        for (int j = 0; j < nref; ++j) {
          const double* vj = V + j * R2 + part * RPL;
          double dd[4] = {0.0, 0.0, 0.0, 0.0};
          static_for<0, RPL>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (part * RPL + i < R2) dd[i & 3] = fma(vj[i], a[i], dd[i & 3]);
          });
          double d = (dd[0] + dd[1]) + (dd[2] + dd[3]);
          d += __shfl_xor(d, 8);
          d += __shfl_xor(d, 16);
          d += __shfl_xor(d, 32);
          d *= Bt[j];
          static_for<0, RPL>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (part * RPL + i < R2) a[i] = fma(d, vj[i], a[i]);
          });
        }
        if (live) {
#pragma unroll
          for (int i = 0; i < RPL; ++i) {
            const int r = part * RPL + i;
            if (r < R2) W[(size_t)r * L_WC + c] = a[i];
          }
        }
      }
    }
    __syncthreads();
    if (stop) break;
  }
  TO* rec = recs + (size_t)k * NY * RECW;
  TO* out = rows_out + (size_t)k * NY * CRW;
  for (int e = tid; e < R2 * L_WC; e += blockDim.x) {
    const int r = e / L_WC, c = e - (e / L_WC) * L_WC;
    const int pcol = c < 3 * NY + NPH ? c : RECW - 1;
    if (r < NY) rec[(size_t)r * RECW + pcol] = (TO)W[e];
    else if (c >= NY) out[(size_t)(r - NY) * CRW + pcol - NY] = (TO)W[e];
  }
}

template <class TI, class TO>
__device__ __forceinline__ void l_cr_pair(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                          const int* __restrict__ act, const TI* __restrict__ rows_in,
                                          TO* __restrict__ rows_out, TO* __restrict__ recs,
                                          double* __restrict__ work) {
  if constexpr (L_RPL <= 48) l_cr_pair_panels<L_RPL, TI, TO>(oinst, seg, act, rows_in, rows_out, recs, work);
  else l_cr_pair_global<TI, TO>(oinst, seg, act, rows_in, rows_out, recs, work);
}

// Boundary system (see cr_final): [C G H | b] of the last row and
// [B_n B_0 V_n | -r_bc] of the boundary conditions, NFIN x (NFIN + 1) in
// work (device memory), solved by right-looking blocked LU with partial
// pivoting (the largest |entry| of the column at or below the diagonal, the
// first on ties; the row exchanges of a panel applied to the columns after
// it, so the right-hand side column receives L^-1 with the trailing
// updates), then blocked back substitution.  One workgroup of
// L_FIN_THREADS per instance; status bit 4 on a zero pivot.  TI: stored type
// of the last level's rows.
constexpr int L_PB = 16;                              // panel width
constexpr int L_PP = L_PB + 1;                        // panel row stride in LDS (odd: a column's rows in distinct banks)
constexpr int L_FC = NFIN + 1;                        // columns of the augmented system
template <class TI>
__device__ __forceinline__ void l_cr_final(Batch bt, const TI* __restrict__ rows, const double* __restrict__ Dr,
                                           const double* __restrict__ rbc, double* __restrict__ dy,
                                           double* __restrict__ dp, int* __restrict__ status,
                                           double* __restrict__ work) {
  constexpr int FC = L_FC;
  constexpr int DRW = 2 * NY + NP;
  __shared__ double P[NFIN * L_PP];        // the panel (rows c0 .. NFIN-1 used), then L21
  __shared__ double S[L_PB * FC];          // the panel rows of the columns after the panel (U12)
  __shared__ double xs[NFIN];
  __shared__ double rv[L_FIN_THREADS / WAVE];
  __shared__ int ri[L_FIN_THREADS / WAVE];
  __shared__ int piv[L_PB];
  __shared__ int trow[2 * L_PB], tsrc[2 * L_PB];
  __shared__ int ntouch;
  const int ib = blockIdx.x;
  if (skipped(bt, ib)) return;
  const TI* row = rows + (size_t)ib * NY * CRW;
  Dr += (size_t)ib * DRSZ;
  rbc += (size_t)ib * NRB;
  double* F = work + (size_t)ib * NFIN * FC;
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int e = tid; e < NFIN * FC; e += nthr) {
    const int r = e / FC, cc = e - (e / FC) * FC;
    double v = 0.0;
    if (r < NY) {
      if (cc < NY) v = (double)row[r * CRW + NY + cc];
      else if (cc < 2 * NY) v = (double)row[r * CRW + cc - NY];
      else if (cc < NFIN) v = HZ ? 0.0 : (double)row[r * CRW + cc];
      else v = (double)row[r * CRW + CRW - 1];
    } else {
      const int rr = r - NY;
      if (cc < NY) v = Dr[rr * DRW + NY + cc];
      else if (cc < 2 * NY) v = Dr[rr * DRW + cc - NY];
      else if (cc < NFIN) v = Dr[rr * DRW + cc];
      else v = -rbc[rr];
    }
    F[e] = v;
  }
  __syncthreads();
  int bad = 0;
  for (int c0 = 0; c0 < NFIN; c0 += L_PB) {
    const int nb = min(L_PB, NFIN - c0);
    const int ce = c0 + nb;   // first column after the panel
    for (int e = tid; e < (NFIN - c0) * L_PB; e += nthr) {
      const int r = c0 + e / L_PB, j = e - (e / L_PB) * L_PB;
      P[r * L_PP + j] = j < nb ? F[(size_t)r * FC + c0 + j] : 0.0;
    }
    __syncthreads();
    // unblocked LU of the panel
    for (int j = 0; j < nb; ++j) {
      const int k = c0 + j;
      double best = -1.0;
      int pr = NFIN;
      for (int r = k + tid; r < NFIN; r += nthr) {
        const double a = fabs(P[r * L_PP + j]);
        if (a > best || (a == best && r < pr)) { best = a; pr = r; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int op = __shfl_xor(pr, off);
        if (ob > best || (ob == best && op < pr)) { best = ob; pr = op; }
      }
      if (lane_id() == 0) { rv[wave_id()] = best; ri[wave_id()] = pr; }
      __syncthreads();
      if (tid == 0) {
        for (int w = 1; w < nthr / WAVE; ++w)
          if (rv[w] > rv[0] || (rv[w] == rv[0] && ri[w] < ri[0])) { rv[0] = rv[w]; ri[0] = ri[w]; }
        piv[j] = ri[0];
      }
      __syncthreads();
      pr = piv[j];
      if (pr != k && tid < L_PB) {
        const double t0 = P[k * L_PP + tid];
        P[k * L_PP + tid] = P[pr * L_PP + tid];
        P[pr * L_PP + tid] = t0;
      }
      __syncthreads();
      const double pv = P[k * L_PP + j];
      if (!(fabs(pv) > 0.0)) bad = 1;
      const double inv = 1.0 / pv;
      for (int e = tid; e < (NFIN - k - 1) * L_PB; e += nthr) {
        const int r = k + 1 + e / L_PB, jj = e - (e / L_PB) * L_PB;
        if (jj <= j || jj >= nb) continue;
        // multiplier of row r (recomputed per entry: the column is read before it is scaled)
        P[r * L_PP + jj] = fma(-(P[r * L_PP + j] * inv), P[k * L_PP + jj], P[r * L_PP + jj]);
      }
      __syncthreads();
      for (int r = k + 1 + tid; r < NFIN; r += nthr) P[r * L_PP + j] *= inv;
      __syncthreads();
    }
    // the panel's row exchanges on the columns after it, composed into one
    // permutation of the rows they touch (the panel's rows, then each pivot
    // row below the panel once): row trow[t] receives row tsrc[t]'s value
    if (tid == 0) {
      int n = nb;
      for (int j = 0; j < nb; ++j) {
        trow[j] = c0 + j;
        tsrc[j] = c0 + j;
      }
      for (int j = 0; j < nb; ++j) {
        const int pr = piv[j];   // >= c0 + j
        int a = pr - c0;
        if (pr >= ce) {
          a = -1;
          for (int t = nb; t < n; ++t)
            if (trow[t] == pr) a = t;
          if (a < 0) {
            a = n++;
            trow[a] = pr;
            tsrc[a] = pr;
          }
        }
        const int t0 = tsrc[j];
        tsrc[j] = tsrc[a];
        tsrc[a] = t0;
      }
      ntouch = n;
    }
    // U11 back to F (the back substitution reads it)
    for (int e = tid; e < nb * L_PB; e += nthr) {
      const int j = e / L_PB, jj = e - (e / L_PB) * L_PB;
      if (jj >= j && jj < nb) F[(size_t)(c0 + j) * FC + c0 + jj] = P[(c0 + j) * L_PP + jj];
    }
    __syncthreads();
    // per column c after the panel: the touched rows read once, U12 =
    // L11^-1 F[c0:ce, c] by forward substitution in registers, the rows
    // below the panel that the exchanges touched written back
    const int nt = ntouch;
    for (int c = ce + tid; c < FC; c += nthr) {
      double v[2 * L_PB];
#pragma unroll
      for (int t = 0; t < 2 * L_PB; ++t) v[t] = t < nt ? F[(size_t)tsrc[t] * FC + c] : 0.0;
#pragma unroll
      for (int j = 0; j < L_PB; ++j) {
        double u = 0.0;
        if (j < nb) {
          u = v[j];
#pragma unroll
          for (int q = 0; q < j; ++q) u = fma(-P[(c0 + j) * L_PP + q], v[q], u);
          v[j] = u;
          F[(size_t)(c0 + j) * FC + c] = u;
        }
        S[j * FC + c] = u;
      }
#pragma unroll
      for (int t = 0; t < 2 * L_PB; ++t)
        if (t >= nb && t < nt) F[(size_t)trow[t] * FC + c] = v[t];
    }
    __syncthreads();
    // trailing update F[ce:, ce:] -= L21 U12, 4 x 4 register tiles
    constexpr int TR = 4, TCL = 4;
    const int nr = NFIN - ce, nc = FC - ce;
    const int tr = (nr + TR - 1) / TR, tc = (nc + TCL - 1) / TCL;
    for (int t = tid; t < tr * tc; t += nthr) {
      const int rt = t / tc, ct = t - (t / tc) * tc;
      const int r0 = ce + rt * TR, cb = ce + ct * TCL;
      double acc[TR][TCL], fv[TR][TCL];
#pragma unroll
      for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int q = 0; q < TCL; ++q) {   // the tile's loads in flight during the products
          fv[a][q] = r0 + a < NFIN && cb + q < FC ? F[(size_t)(r0 + a) * FC + cb + q] : 0.0;
          acc[a][q] = 0.0;
        }
#pragma nounroll
      for (int j = 0; j < L_PB; ++j) {
        double l[TR], u[TCL];
#pragma unroll
        for (int a = 0; a < TR; ++a) l[a] = r0 + a < NFIN ? P[(r0 + a) * L_PP + j] : 0.0;
#pragma unroll
        for (int q = 0; q < TCL; ++q) u[q] = cb + q < FC ? S[j * FC + cb + q] : 0.0;
#pragma unroll
        for (int a = 0; a < TR; ++a)
#pragma unroll
          for (int q = 0; q < TCL; ++q) acc[a][q] = fma(l[a], u[q], acc[a][q]);
      }
#pragma unroll
      for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int q = 0; q < TCL; ++q)
          if (r0 + a < NFIN && cb + q < FC) F[(size_t)(r0 + a) * FC + cb + q] = fv[a][q] - acc[a][q];
    }
    __syncthreads();
  }
  // back substitution U x = F[:, NFIN], a block of L_PB rows at a time
  for (int r = tid; r < NFIN; r += nthr) xs[r] = F[(size_t)r * FC + NFIN];
  __syncthreads();
  for (int b0 = ((NFIN - 1) / L_PB) * L_PB; b0 >= 0; b0 -= L_PB) {
    const int be = min(b0 + L_PB, NFIN);
    if (tid < WAVE) {   // the diagonal block, one wave: lane j holds row b0 + j (read at once)
      const int r = b0 + tid;
      double u[L_PB];
#pragma unroll
      for (int j = 0; j < L_PB; ++j) u[j] = r < be && b0 + j < be ? F[(size_t)r * FC + b0 + j] : 0.0;
      double v = r < be ? xs[r] : 0.0;
#pragma unroll
      for (int j = L_PB - 1; j >= 0; --j) {
        const int c = b0 + j;
        if (c < be) {
          const double xc = __shfl(v, j) / __shfl(u[j], j);   // lane j's diagonal entry
          if (r == c) v = xc;
          else if (r < c) v = fma(-u[j], xc, v);
        }
      }
      if (r < be) xs[r] = v;
    }
    __syncthreads();
    for (int r = tid; r < b0; r += nthr) {
      double u[L_PB];
#pragma unroll
      for (int j = 0; j < L_PB; ++j) u[j] = b0 + j < be ? F[(size_t)r * FC + b0 + j] : 0.0;
      double v = xs[r];
#pragma unroll
      for (int j = 0; j < L_PB; ++j)
        if (b0 + j < be) v = fma(-u[j], xs[b0 + j], v);
      xs[r] = v;
    }
    __syncthreads();
  }
  double* x0 = dy + (size_t)bt.noff[ib] * NY;
  double* xn = dy + (size_t)(bt.noff[ib + 1] - 1) * NY;
  dp += (size_t)ib * NP;
  for (int r = tid; r < NFIN; r += nthr) {
    const double v = xs[r];
    if (r < NY) xn[r] = v;
    else if (r < 2 * NY) x0[r - NY] = v;
    else dp[r - 2 * NY] = v;
  }
  if (bad && tid == 0) atomicOr(status, 4);
}

// Back substitution of one level (see cr_back), one workgroup per output
// row: the right-hand side per row, then R's triangular solve column by
// column in LDS.  TR / TI: stored type of the records / of the level's input
// rows (float64 arithmetic).
template <class TR, class TI>
__device__ __forceinline__ void l_cr_back(int level, Batch bt, const int* __restrict__ oinst,
                                          const CRSeg* __restrict__ seg, const TR* __restrict__ recs,
                                          const TI* __restrict__ rows_in, double* __restrict__ dy,
                                          const double* __restrict__ dp) {
  __shared__ double rhs[NY];
  __shared__ double xs;
  const int ib = oinst[blockIdx.x];
  if (skipped(bt, ib)) return;
  const CRSeg sg = seg[ib];
  const int k = blockIdx.x - sg.out_off;
  if (k >= (sg.nrows >> 1)) return;
  const int tid = threadIdx.x;
  const int base = bt.noff[ib];
  const int nint = bt.noff[ib + 1] - base - 1;
  const int mid = base + ((2 * k + 1) << level);
  const int left = base + ((2 * k) << level);
  const int right = base + min((2 * k + 2) << level, nint);
  const double* xl = dy + (size_t)left * NY;
  const double* xr = dy + (size_t)right * NY;
  dp += (size_t)ib * NP;
  if (sg.subst) {
    const TI* ra = rows_in + (size_t)(sg.in_off + 2 * k) * NY * CRW;
    for (int r = tid; r < NY; r += blockDim.x) {
      double v = (double)ra[(size_t)r * CRW + CRW - 1];
      for (int c = 0; c < NY; ++c) v = fma(-(double)ra[(size_t)r * CRW + c], xl[c], v);
      for (int c = 0; c < NPH; ++c) v = fma(-(double)ra[(size_t)r * CRW + 2 * NY + c], dp[c], v);
      dy[(size_t)mid * NY + r] = v;
    }
    return;
  }
  const TR* rec = recs + (size_t)(sg.rec_off + k) * NY * RECW;
  for (int r = tid; r < NY; r += blockDim.x) {
    const TR* rw = rec + (size_t)r * RECW;
    double v = (double)rw[RECW - 1];
    for (int c = 0; c < NY; ++c) v = fma(-(double)rw[NY + c], xl[c], v);
    for (int c = 0; c < NY; ++c) v = fma(-(double)rw[2 * NY + c], xr[c], v);
    for (int c = 0; c < NPH; ++c) v = fma(-(double)rw[3 * NY + c], dp[c], v);
    rhs[r] = v;
  }
  __syncthreads();
  for (int c = NY - 1; c >= 0; --c) {
    if (tid == 0) xs = rhs[c] / (double)rec[(size_t)c * RECW + c];
    __syncthreads();
    const double x = xs;
    for (int r = tid; r < c; r += blockDim.x) rhs[r] = fma(-(double)rec[(size_t)r * RECW + c], x, rhs[r]);
    if (tid == 0) rhs[c] = x;
    __syncthreads();
  }
  for (int r = tid; r < NY; r += blockDim.x) dy[(size_t)mid * NY + r] = rhs[r];
}

}  // namespace colloc

// ------------------------------------------------------------ entry points
extern "C" {
// the register-resident Gauss-Jordan wants the whole register file of its SIMD
__global__ __launch_bounds__(colloc::L_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void colloc_l_stage_elim(int nint, colloc::Batch bt,
    const double* jac, const double* fa, double* rec, int* status) {
  colloc::l_stage_elim<double>(nint, bt, jac, fa, rec, status);
}
__global__ __launch_bounds__(colloc::L_THREADS) __attribute__((amdgpu_waves_per_eu(1, 2)))
void colloc_l_stage_elim32(int nint, colloc::Batch bt,
    const float* jac, const double* fa, float* rec, int* status) {
  colloc::l_stage_elim<float>(nint, bt, jac, fa, rec, status);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_init(int nint, colloc::Batch bt,
    const double* t, double* rec, unsigned long long* slots) {
  colloc::l_series_init<double>(nint, bt, t, rec, slots);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_init32(int nint, colloc::Batch bt,
    const double* t, float* rec, unsigned long long* slots) {
  colloc::l_series_init<float>(nint, bt, t, rec, slots);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_term(int nint, colloc::Batch bt,
    const double* t, const double* rec, const double* uold, long long so, double* unew, long long sn,
    unsigned long long* slots, unsigned long long* active, int term) {
  colloc::l_series_term<double>(nint, bt, t, rec, uold, so, unew, sn, slots, active, term);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_term32(int nint, colloc::Batch bt,
    const double* t, const float* rec, const float* uold, long long so, float* unew, long long sn,
    unsigned long long* slots, unsigned long long* active, int term) {
  colloc::l_series_term<float>(nint, bt, t, rec, uold, so, unew, sn, slots, active, term);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_out(int nint, colloc::Batch bt,
    const double* fb, double* rows, double* rec, const double* ubuf, const unsigned long long* slots, int nterms,
    unsigned long long* counters) {
  colloc::l_series_out<double>(nint, bt, fb, rows, rec, ubuf, slots, nterms, counters);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_out32(int nint, colloc::Batch bt,
    const double* fb, double* rows, float* rec, const float* ubuf, const unsigned long long* slots, int nterms,
    unsigned long long* counters) {
  colloc::l_series_out<float>(nint, bt, fb, rows, rec, ubuf, slots, nterms, counters);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_subst(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, double* rows_out) {
  colloc::l_cr_subst<double, double>(oinst, seg, act, rows_in, rows_out);
}
__global__ __launch_bounds__(colloc::L_QR_THREADS) void colloc_l_cr_pair(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, double* rows_out, double* recs, double* work) {
  colloc::l_cr_pair<double, double>(oinst, seg, act, rows_in, rows_out, recs, work);
}
__global__ __launch_bounds__(colloc::L_FIN_THREADS) void colloc_l_cr_final(colloc::Batch bt, const double* rows,
    const double* Dr, const double* rbc, double* dy, double* dp, int* status, double* work) {
  colloc::l_cr_final<double>(bt, rows, Dr, rbc, dy, dp, status, work);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_back(int level, colloc::Batch bt, const int* oinst,
    const colloc::CRSeg* seg, const double* recs, const double* rows_in, double* dy, const double* dp) {
  colloc::l_cr_back<double, double>(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
// the FP32 cyclic reduction: level 0 reads the condensation's float64 rows
// (_df), the levels after it float32 rows (_ff); records are float32
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_subst_df(const int* oinst,
    const colloc::CRSeg* seg, const int* act, const double* rows_in, float* rows_out) {
  colloc::l_cr_subst<double, float>(oinst, seg, act, rows_in, rows_out);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_subst_ff(const int* oinst,
    const colloc::CRSeg* seg, const int* act, const float* rows_in, float* rows_out) {
  colloc::l_cr_subst<float, float>(oinst, seg, act, rows_in, rows_out);
}
__global__ __launch_bounds__(colloc::L_QR_THREADS) void colloc_l_cr_pair_df(const int* oinst,
    const colloc::CRSeg* seg, const int* act, const double* rows_in, float* rows_out, float* recs, double* work) {
  colloc::l_cr_pair<double, float>(oinst, seg, act, rows_in, rows_out, recs, work);
}
__global__ __launch_bounds__(colloc::L_QR_THREADS) void colloc_l_cr_pair_ff(const int* oinst,
    const colloc::CRSeg* seg, const int* act, const float* rows_in, float* rows_out, float* recs, double* work) {
  colloc::l_cr_pair<float, float>(oinst, seg, act, rows_in, rows_out, recs, work);
}
__global__ __launch_bounds__(colloc::L_FIN_THREADS) void colloc_l_cr_final_f(colloc::Batch bt, const float* rows,
    const double* Dr, const double* rbc, double* dy, double* dp, int* status, double* work) {
  colloc::l_cr_final<float>(bt, rows, Dr, rbc, dy, dp, status, work);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_back_fd(int level, colloc::Batch bt,
    const int* oinst, const colloc::CRSeg* seg, const float* recs, const double* rows_in, double* dy,
    const double* dp) {
  colloc::l_cr_back<float, double>(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_back_ff(int level, colloc::Batch bt,
    const int* oinst, const colloc::CRSeg* seg, const float* recs, const float* rows_in, double* dy,
    const double* dp) {
  colloc::l_cr_back<float, float>(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
// [0..3] as the register-resident kernels' (none here), [4] table-form
// points per workgroup, [5] large-block kernels, [6] tiles per interval of
// the series term, [7] tiles per pair of the substitution product
__device__ int colloc_caps[8] = {0, 0, colloc::HZ ? 1 : 0, 1, COLLOC_TABLE_FORM ? colloc::TAB_G : 0, 1,
                                 colloc::L_TILES, colloc::L_STILES};
}  // extern "C"
