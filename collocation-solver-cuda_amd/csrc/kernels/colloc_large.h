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
constexpr int L_SLOT = 8;                        // u64 per interval: 3 x (change, size), the converged term
constexpr int CS_XC = NY + NPH + 1;              // columns of [G_a | H_a | b_a]
constexpr int L_SCT = (CS_XC + L_TILE - 1) / L_TILE;
constexpr int L_STILES = L_NRT * L_SCT;
constexpr int L_WC = 3 * NY + NPH + 1;           // columns of a Householder pair's work matrix
constexpr int L_QR_THREADS = 1024;
constexpr int L_FIN_THREADS = 1024;

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
// in LDS in chunks of L_KC.  Wave w owns rows 32 (w / 2) .. +32 and columns
// 32 (w % 2) .. +32 of the tile as 2 x 2 matrix-core tiles of 16 x 16:
// element v of acc[ti][tj] in lane l is row 32 (w / 2) + 16 ti +
// MfmaOp<T>::row(l, v), column 32 (w % 2) + 16 tj + l % 16.  Every thread of
// the (4-wave) workgroup calls it.
template <class T, class FA, class FB>
__device__ __forceinline__ void tile_gemm(int K, FA&& A, FB&& B, T* sA, T* sB,
                                          typename MfmaOp<T>::V (&acc)[2][2]) {
  const int tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
  for (int k0 = 0; k0 < K; k0 += L_KC) {
    __syncthreads();   // the previous chunk's reads are done
    for (int e = tid; e < L_TILE * L_KC; e += blockDim.x) {
      const int r = e / L_KC, k = e - (e / L_KC) * L_KC;
      sA[r * (L_KC + 1) + k] = (k0 + k < K) ? A(r, k0 + k) : (T)0;
    }
    for (int e = tid; e < L_KC * L_TILE; e += blockDim.x) {
      const int k = e / L_TILE, c = e - (e / L_TILE) * L_TILE;
      sB[k * (L_TILE + 1) + c] = (k0 + k < K) ? B(k0 + k, c) : (T)0;
    }
    __syncthreads();
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

// ------------------------------------------------------ stage eliminations
// One workgroup per (listed interval, stage j), computing in R: [g_z | g_y
// g_pa f_a,g] (NZ rows) in LDS, Gauss-Jordan with partial pivoting (the
// largest |entry| of the column, first on ties), then T_j row by row.
// Writes G_j and T_j to the record (type R); status bit 1 on a zero pivot.
constexpr int L_GZW = NZ + NVA + 1;   // padded row of the algebraic block
template <class R>
__device__ __forceinline__ void l_stage_elim(int nint, Batch bt, const R* __restrict__ jac,
                                             const double* __restrict__ fa, R* __restrict__ rec,
                                             int* __restrict__ status) {
  __shared__ R gz[NZ1 * L_GZW];
  __shared__ R hz[NY * NZ1];
  __shared__ R fcol[NZ1];
  __shared__ int spiv;
  const int item = blockIdx.x / M, j = blockIdx.x - (blockIdx.x / M) * M;
  if (item >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, item);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  const int tid = threadIdx.x;
  const R* Dfj = jac + ((size_t)i * M + j) * JSTAGE;
  const R* Dgj = Dfj + NY * NJC;
  const double* faj = fa + (size_t)i * M * NYZ + j * NYZ;
  R* recT = rec + (size_t)i * REC + NS * NVA;
  R* recG = recT + M * NY * NVA;
  for (int e = tid; e < NZ * (NZ + NVA); e += blockDim.x) {
    const int r = e / (NZ + NVA), cc = e - (e / (NZ + NVA)) * (NZ + NVA);
    R v;
    if (cc < NZ) v = Dgj[r * NJC + NY + cc];
    else {
      const int q = cc - NZ;
      v = q < NY ? Dgj[r * NJC + q] : (q < NY + NPA ? Dgj[r * NJC + NY + NZ + pact(q - NY)] : (R)faj[NY + r]);
    }
    gz[r * L_GZW + cc] = v;
  }
  for (int e = tid; e < NY * NZ; e += blockDim.x) {
    const int r = e / NZ1, z = e - (e / NZ1) * NZ1;
    hz[e] = Dfj[r * NJC + NY + z];
  }
  __syncthreads();
  int bad = 0;
  for (int k = 0; k < NZ; ++k) {
    if (tid < WAVE) {
      // pivot: the largest |g_z[r][k]|, r >= k, the first on ties
      R best = (R)-1;
      int pr = k;
      for (int r = k + tid; r < NZ; r += WAVE) {
        const R a = fabs(gz[r * L_GZW + k]);
        if (a > best) { best = a; pr = r; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const R ob = __shfl_xor(best, off);
        const int op = __shfl_xor(pr, off);
        if (ob > best || (ob == best && op < pr)) { best = ob; pr = op; }
      }
      if (tid == 0) spiv = pr;
    }
    __syncthreads();
    const int pr = spiv;
    if (pr != k)
      for (int c = tid; c < NZ + NVA; c += blockDim.x) {
        const R t0 = gz[k * L_GZW + c];
        gz[k * L_GZW + c] = gz[pr * L_GZW + c];
        gz[pr * L_GZW + c] = t0;
      }
    __syncthreads();
    const R piv = gz[k * L_GZW + k];
    if (!(fabs(piv) > (R)0)) bad = 1;
    const R inv = (R)1 / piv;
    if (tid < NZ) fcol[tid] = gz[tid * L_GZW + k];
    __syncthreads();
    for (int c = tid; c < NZ + NVA; c += blockDim.x) gz[k * L_GZW + c] *= inv;
    __syncthreads();
    for (int e = tid; e < NZ * (NZ + NVA); e += blockDim.x) {
      const int r = e / (NZ + NVA), c = e - (e / (NZ + NVA)) * (NZ + NVA);
      if (r != k) gz[r * L_GZW + c] = fma(-fcol[r], gz[k * L_GZW + c], gz[r * L_GZW + c]);
    }
    __syncthreads();
  }
  for (int e = tid; e < NZ * NVA; e += blockDim.x) {
    const int r = e / NVA, q = e - (e / NVA) * NVA;
    recG[(j * NZ + r) * NVA + q] = gz[r * L_GZW + NZ + q];
  }
  // T_j = [h_y | h_pa | f_a,h] - h_z G_j
  for (int e = tid; e < NY * NVA; e += blockDim.x) {
    const int r = e / NVA, q = e - (e / NVA) * NVA;
    R v = q < NY ? Dfj[r * NJC + q] : (q < NY + NPA ? Dfj[r * NJC + NY + NZ + pact(q - NY)] : (R)faj[r]);
    for (int z = 0; z < NZ; ++z) v = fma(-hz[r * NZ1 + z], gz[z * L_GZW + NZ + q], v);
    recT[(j * NY + r) * NVA + q] = v;
  }
  if (bad && tid == 0) atomicOr(status, 1);
}

// ---------------------------------------------------------------- series
// R_k (k = 0 .. NSK-1, stage k+1) of the listed intervals:
//   R_k = [I | 0] + h A[k+1][0] T_0 + sum_{j>=1} h A[k+1][j] [0 | T_j[:, NY:]],
// written to rbuf and to the record's U block (the series' first iterate).
// One workgroup per (listed interval, stage block); slots cleared.
template <class R>
__device__ __forceinline__ void l_series_init(int nint, Batch bt, const double* __restrict__ t, R* __restrict__ rec,
                                              R* __restrict__ rbuf, unsigned long long* __restrict__ slots) {
  const int item = blockIdx.x / NSK, k = blockIdx.x - (blockIdx.x / NSK) * NSK;
  if (item >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, item);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  const R h = (R)(t[i + ib + 1] - t[i + ib]);
  R* recU = rec + (size_t)i * REC;
  const R* recT = recU + NS * NVA;
  R* rb = rbuf + (size_t)i * NS * NVA;
  const int js = k + 1;
  for (int e = threadIdx.x; e < NY * NVA; e += blockDim.x) {
    const int r = e / NVA, c = e - (e / NVA) * NVA;
    R v = h * (R)c_rkA[js * M] * recT[r * NVA + c];
    if (c < NY) {
      if (c == r) v += (R)1;
    } else {
      for (int jj = 1; jj < M; ++jj) v += h * (R)c_rkA[js * M + jj] * recT[(jj * NY + r) * NVA + c];
    }
    rb[(k * NY + r) * NVA + c] = v;
    recU[(k * NY + r) * NVA + c] = v;
  }
  if (k == 0 && threadIdx.x < L_SLOT) slots[(size_t)i * L_SLOT + threadIdx.x] = 0;
}

__device__ __forceinline__ bool l_converged(const unsigned long long* s, int set, double tol) {
  const double d = slot_d(s[2 * set]), x = slot_d(s[2 * set + 1]);
  return d <= tol * x;
}

// Term `term` (>= 1) of the series for tile (rt, ct) of the listed interval:
// Unew_k = R_k + sum_q h A[k+1][q+1] M_{q+1} Uold_q on rows rt, columns ct.
// Uold / Unew: the U blocks (stride ustride_old / ustride_new per interval);
// slots[i]: [3 x (largest change, largest entry) by term mod 3 | converged
// term | -]; active[term]: intervals that computed this term.
template <class R>
__device__ __forceinline__ void l_series_term(int nint, Batch bt, const double* __restrict__ t,
                                              const R* __restrict__ rec, const R* __restrict__ rbuf,
                                              const R* __restrict__ uold, long long ustride_old,
                                              R* __restrict__ unew, long long ustride_new,
                                              unsigned long long* __restrict__ slots,
                                              unsigned long long* __restrict__ active, int term) {
  __shared__ __attribute__((aligned(16))) R sA[L_SA];
  __shared__ __attribute__((aligned(16))) R sB[L_SB];
  const int item = blockIdx.x / L_TILES, tile = blockIdx.x - (blockIdx.x / L_TILES) * L_TILES;
  if (item >= launch_intervals(bt, nint)) return;
  const int i = listed_interval(bt, item);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  unsigned long long* s = slots + (size_t)i * L_SLOT;
  const double tol = (double)l_tol<R>();
  if (term >= 2 && l_converged(s, (term - 1) % 3, tol)) {
    if (tile == 0 && threadIdx.x == 0 && s[6] == 0) s[6] = (unsigned long long)(term - 1);
    return;
  }
  if (tile == 0 && threadIdx.x == 0) {
    const int nx = (term + 1) % 3;
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
  const R* rb = rbuf + (size_t)i * NS * NVA;
  R* un = unew + (size_t)i * ustride_new;
  double dmax = 0.0, xmax = 0.0;
#pragma unroll
  for (int k = 0; k < NSK; ++k) {
    R wk[NSK];
#pragma unroll
    for (int q = 0; q < NSK; ++q) wk[q] = h * (R)c_rkA[(k + 1) * M + q + 1];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = rbase + 32 * (w >> 1) + 16 * ti + MfmaOp<R>::row(lane, v);
          const int col = cbase + 32 * (w & 1) + 16 * tj + (lane & 15);
          if (row >= NY || col >= NVA) continue;
          const size_t e = (size_t)(k * NY + row) * NVA + col;
          R x = rb[e];
#pragma unroll
          for (int q = 0; q < NSK; ++q) x = fma(wk[q], acc[q][ti][tj][v], x);
          dmax = nmax(fabs((double)x - (double)uo[e]), dmax);
          xmax = nmax(fabs((double)x), xmax);
          un[e] = x;
        }
  }
  l_block_max2(dmax, xmax, s + 2 * (term % 3), s + 2 * (term % 3) + 1);
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
    int tc = (int)s[6];
    if (tc == 0 && nterms >= 1 && l_converged(s, nterms % 3, (double)l_tol<R>())) tc = nterms;
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
// product; workgroup = (output row of the level, tile).
__device__ __forceinline__ void l_cr_subst(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                           const int* __restrict__ act, const double* __restrict__ rows_in,
                                           double* __restrict__ rows_out) {
  __shared__ __attribute__((aligned(16))) double sA[L_SA];
  __shared__ __attribute__((aligned(16))) double sB[L_SB];
  const int wg = blockIdx.x / L_STILES, tile = blockIdx.x - (blockIdx.x / L_STILES) * L_STILES;
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
    const double* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    double* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = threadIdx.x; e < L_TILE * CRW; e += blockDim.x) {
      const int r = rt * L_TILE + e / CRW, c = e - (e / CRW) * CRW;
      if (r >= NY) break;
      const size_t x = (size_t)r * CRW + c;
      if (c < NY || c == CRW - 1 || (!HZ && c >= 2 * NY)) dst[x] = src[x];
      else if (wc && c < 2 * NY) dst[x] = (c - NY == r) ? 1.0 : 0.0;
    }
    return;
  }
  const double* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const double* rb = ra + NY * CRW;
  double* out = rows_out + (size_t)k * NY * CRW;
  const int rbase = rt * L_TILE, cbase = ct * L_TILE;
  typename MfmaOp<double>::V acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[ti][tj][v] = 0.0;
  tile_gemm<double>(NY,
      [&](int r, int q) { return rbase + r < NY ? rb[(size_t)(rbase + r) * CRW + q] : 0.0; },
      [&](int q, int c) {
        const int col = cbase + c;
        if (col < NY) return ra[(size_t)q * CRW + col];                 // G_a
        if (col < NY + NPH) return ra[(size_t)q * CRW + NY + col];      // H_a (after C_a)
        if (col == NY + NPH) return ra[(size_t)q * CRW + CRW - 1];      // b_a
        return 0.0;
      },
      sA, sB, acc);
  const int lane = lane_id(), w = wave_id();
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rbase + 32 * (w >> 1) + 16 * ti + MfmaOp<double>::row(lane, v);
        const int col = cbase + 32 * (w & 1) + 16 * tj + (lane & 15);
        if (r >= NY) continue;
        const double a = acc[ti][tj][v];
        if (col < NY) out[(size_t)r * CRW + col] = -a;
        else if (col < NY + NPH) out[(size_t)r * CRW + NY + col] = rb[(size_t)r * CRW + NY + col] - a;
        else if (col == NY + NPH) out[(size_t)r * CRW + CRW - 1] = rb[(size_t)r * CRW + CRW - 1] - a;
      }
  if (wc && ct == 0) {
    for (int e = threadIdx.x; e < L_TILE * NY; e += blockDim.x) {
      const int r = rbase + e / NY, c = e - (e / NY) * NY;
      if (r < NY) out[(size_t)r * CRW + NY + c] = (r == c) ? 1.0 : 0.0;
    }
  }
}

// Householder pairing (see cr_pair; the reference's qr,
// matrix_factorization_cuda.py:5-82: g = -sign(v_k) |v|, stop at |v|^2 <
// eps): the pair's 2NY x L_WC block [x_M | x_L | x_R | p | rhs] in work
// (global memory), one workgroup of L_QR_THREADS; thread (column slot,
// row quarter) applies each reflector to a column's quarter, the quarters'
// dot products summed through LDS.
__device__ __forceinline__ void l_cr_pair(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                          const int* __restrict__ act, const double* __restrict__ rows_in,
                                          double* __restrict__ rows_out, double* __restrict__ recs,
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
    const double* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    double* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = tid; e < NY * CRW; e += blockDim.x) dst[e] = src[e];
    return;
  }
  double* W = work + (size_t)blockIdx.x * R2 * L_WC;
  const double* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const double* rb = ra + NY * CRW;
  for (int e = tid; e < R2 * L_WC; e += blockDim.x) {
    const int r = e / L_WC, c = e - (e / L_WC) * L_WC;
    const bool top = r < NY;
    const double* src = top ? ra + (size_t)r * CRW : rb + (size_t)(r - NY) * CRW;
    const int pcol = c < 3 * NY + NPH ? c : RECW - 1;
    double v;
    if (c < NY) v = top ? src[NY + c] : src[c];
    else if (c < 2 * NY) v = top ? src[c - NY] : 0.0;
    else if (c < 3 * NY) v = top ? 0.0 : src[c - NY];
    else v = src[pcol - NY];
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
  double* rec = recs + (size_t)k * NY * RECW;
  double* out = rows_out + (size_t)k * NY * CRW;
  for (int e = tid; e < R2 * L_WC; e += blockDim.x) {
    const int r = e / L_WC, c = e - (e / L_WC) * L_WC;
    const int pcol = c < 3 * NY + NPH ? c : RECW - 1;
    if (r < NY) rec[(size_t)r * RECW + pcol] = W[e];
    else if (c >= NY) out[(size_t)(r - NY) * CRW + pcol - NY] = W[e];
  }
}

// Boundary system (see cr_final) by Gauss-Jordan with partial pivoting on
// the (NFIN) x (NFIN + 1) matrix in work (global memory), one workgroup of
// L_FIN_THREADS per instance; status bit 4 on a zero pivot.
__device__ __forceinline__ void l_cr_final(Batch bt, const double* __restrict__ rows, const double* __restrict__ Dr,
                                           const double* __restrict__ rbc, double* __restrict__ dy,
                                           double* __restrict__ dp, int* __restrict__ status,
                                           double* __restrict__ work) {
  constexpr int FC = NFIN + 1;
  constexpr int DRW = 2 * NY + NP;
  __shared__ double fcol[NFIN];
  __shared__ double prow[FC];
  __shared__ double rv[L_FIN_THREADS / WAVE];
  __shared__ int ri[L_FIN_THREADS / WAVE];
  const int ib = blockIdx.x;
  if (skipped(bt, ib)) return;
  const double* row = rows + (size_t)ib * NY * CRW;
  Dr += (size_t)ib * DRSZ;
  rbc += (size_t)ib * NRB;
  double* F = work + (size_t)ib * NFIN * FC;
  const int tid = threadIdx.x;
  for (int e = tid; e < NFIN * FC; e += blockDim.x) {
    const int r = e / FC, cc = e - (e / FC) * FC;
    double v = 0.0;
    if (r < NY) {
      if (cc < NY) v = row[r * CRW + NY + cc];
      else if (cc < 2 * NY) v = row[r * CRW + cc - NY];
      else if (cc < NFIN) v = HZ ? 0.0 : row[r * CRW + cc];
      else v = row[r * CRW + CRW - 1];
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
  for (int k = 0; k < NFIN; ++k) {
    // pivot: the largest |F[r][k]|, r >= k (the first on ties)
    double best = -1.0;
    int pr = NFIN;
    for (int r = k + tid; r < NFIN; r += blockDim.x) {
      const double a = fabs(F[(size_t)r * FC + k]);
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
      for (int w = 1; w < (int)blockDim.x / WAVE; ++w)
        if (rv[w] > rv[0] || (rv[w] == rv[0] && ri[w] < ri[0])) { rv[0] = rv[w]; ri[0] = ri[w]; }
    }
    __syncthreads();
    pr = ri[0];
    if (pr != k)
      for (int c = tid; c < FC; c += blockDim.x) {
        const double t0 = F[(size_t)k * FC + c];
        F[(size_t)k * FC + c] = F[(size_t)pr * FC + c];
        F[(size_t)pr * FC + c] = t0;
      }
    __syncthreads();
    const double piv = F[(size_t)k * FC + k];
    if (!(fabs(piv) > 0.0)) bad = 1;
    const double inv = 1.0 / piv;
    for (int c = tid; c < FC; c += blockDim.x) prow[c] = F[(size_t)k * FC + c] * inv;
    for (int r = tid; r < NFIN; r += blockDim.x) fcol[r] = F[(size_t)r * FC + k];
    __syncthreads();
    for (int e = tid; e < NFIN * FC; e += blockDim.x) {
      const int r = e / FC, c = e - (e / FC) * FC;
      if (c < k) continue;
      F[e] = r == k ? prow[c] : fma(-fcol[r], prow[c], F[e]);
    }
    __syncthreads();
  }
  double* x0 = dy + (size_t)bt.noff[ib] * NY;
  double* xn = dy + (size_t)(bt.noff[ib + 1] - 1) * NY;
  dp += (size_t)ib * NP;
  for (int r = tid; r < NFIN; r += blockDim.x) {
    const double v = F[(size_t)r * FC + NFIN];
    if (r < NY) xn[r] = v;
    else if (r < 2 * NY) x0[r - NY] = v;
    else dp[r - 2 * NY] = v;
  }
  if (bad && tid == 0) atomicOr(status, 4);
}

// Back substitution of one level (see cr_back), one workgroup per output
// row: the right-hand side per row, then R's triangular solve column by
// column in LDS.
__device__ __forceinline__ void l_cr_back(int level, Batch bt, const int* __restrict__ oinst,
                                          const CRSeg* __restrict__ seg, const double* __restrict__ recs,
                                          const double* __restrict__ rows_in, double* __restrict__ dy,
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
    const double* ra = rows_in + (size_t)(sg.in_off + 2 * k) * NY * CRW;
    for (int r = tid; r < NY; r += blockDim.x) {
      double v = ra[(size_t)r * CRW + CRW - 1];
      for (int c = 0; c < NY; ++c) v = fma(-ra[(size_t)r * CRW + c], xl[c], v);
      for (int c = 0; c < NPH; ++c) v = fma(-ra[(size_t)r * CRW + 2 * NY + c], dp[c], v);
      dy[(size_t)mid * NY + r] = v;
    }
    return;
  }
  const double* rec = recs + (size_t)(sg.rec_off + k) * NY * RECW;
  for (int r = tid; r < NY; r += blockDim.x) {
    const double* rw = rec + (size_t)r * RECW;
    double v = rw[RECW - 1];
    for (int c = 0; c < NY; ++c) v = fma(-rw[NY + c], xl[c], v);
    for (int c = 0; c < NY; ++c) v = fma(-rw[2 * NY + c], xr[c], v);
    for (int c = 0; c < NPH; ++c) v = fma(-rw[3 * NY + c], dp[c], v);
    rhs[r] = v;
  }
  __syncthreads();
  for (int c = NY - 1; c >= 0; --c) {
    if (tid == 0) xs = rhs[c] / rec[(size_t)c * RECW + c];
    __syncthreads();
    const double x = xs;
    for (int r = tid; r < c; r += blockDim.x) rhs[r] = fma(-rec[(size_t)r * RECW + c], x, rhs[r]);
    if (tid == 0) rhs[c] = x;
    __syncthreads();
  }
  for (int r = tid; r < NY; r += blockDim.x) dy[(size_t)mid * NY + r] = rhs[r];
}

}  // namespace colloc

// ------------------------------------------------------------ entry points
extern "C" {
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_stage_elim(int nint, colloc::Batch bt,
    const double* jac, const double* fa, double* rec, int* status) {
  colloc::l_stage_elim<double>(nint, bt, jac, fa, rec, status);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_stage_elim32(int nint, colloc::Batch bt,
    const float* jac, const double* fa, float* rec, int* status) {
  colloc::l_stage_elim<float>(nint, bt, jac, fa, rec, status);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_init(int nint, colloc::Batch bt,
    const double* t, double* rec, double* rbuf, unsigned long long* slots) {
  colloc::l_series_init<double>(nint, bt, t, rec, rbuf, slots);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_init32(int nint, colloc::Batch bt,
    const double* t, float* rec, float* rbuf, unsigned long long* slots) {
  colloc::l_series_init<float>(nint, bt, t, rec, rbuf, slots);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_term(int nint, colloc::Batch bt,
    const double* t, const double* rec, const double* rbuf, const double* uold, long long so, double* unew,
    long long sn, unsigned long long* slots, unsigned long long* active, int term) {
  colloc::l_series_term<double>(nint, bt, t, rec, rbuf, uold, so, unew, sn, slots, active, term);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_series_term32(int nint, colloc::Batch bt,
    const double* t, const float* rec, const float* rbuf, const float* uold, long long so, float* unew,
    long long sn, unsigned long long* slots, unsigned long long* active, int term) {
  colloc::l_series_term<float>(nint, bt, t, rec, rbuf, uold, so, unew, sn, slots, active, term);
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
  colloc::l_cr_subst(oinst, seg, act, rows_in, rows_out);
}
__global__ __launch_bounds__(colloc::L_QR_THREADS) void colloc_l_cr_pair(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, double* rows_out, double* recs, double* work) {
  colloc::l_cr_pair(oinst, seg, act, rows_in, rows_out, recs, work);
}
__global__ __launch_bounds__(colloc::L_FIN_THREADS) void colloc_l_cr_final(colloc::Batch bt, const double* rows,
    const double* Dr, const double* rbc, double* dy, double* dp, int* status, double* work) {
  colloc::l_cr_final(bt, rows, Dr, rbc, dy, dp, status, work);
}
__global__ __launch_bounds__(colloc::L_THREADS) void colloc_l_cr_back(int level, colloc::Batch bt, const int* oinst,
    const colloc::CRSeg* seg, const double* recs, const double* rows_in, double* dy, const double* dp) {
  colloc::l_cr_back(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
// [0..3] as the register-resident kernels' (none here), [4] table-form
// points per workgroup, [5] large-block kernels, [6] tiles per interval of
// the series term, [7] tiles per pair of the substitution product
__device__ int colloc_caps[8] = {0, 0, colloc::HZ ? 1 : 0, 1, COLLOC_TABLE_FORM ? colloc::TAB_G : 0, 1,
                                 colloc::L_TILES, colloc::L_STILES};
}  // extern "C"
