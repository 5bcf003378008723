// Collocation BVP-DAE kernels for gfx950, specialised per problem at JIT time.
//
// The including translation unit defines, before including this file:
//   namespace prob { f, g, Df, Dg, r, Dr }   -- the generated callbacks
//   struct Cfg { NY, NZ, NP, NC, M, CW, FW; }  -- sizes and launch shapes
//   __constant__ tables  c_rkA[M*M], c_rkb[M], c_rkc[M],
//                        c_gL[(M+1)*M], c_gI[(M+1)*M], c_gw[M+1], c_L0[M], c_L1[M]
//
// Device memory layouts (row-major, float64) for a batch of B instances laid
// end to end (N nodes and Nint = N - B intervals in total, see Batch):
//   t[N]; y[N][NY]; ydot[Nint*M][NY]; ztil[Nint*M][NZ]; p[B][NP]
//   ytil[Nint*M][NY]; fa[Nint][M*(NY+NZ)]; fb[Nint][NY]; rbc[B][NY+NP]
//   jac[Nint*M][JSTAGE]      stage block = Df (NY x NJC) then Dg (NZ x NJC), float64
//                            (float32 for the FP32 direction)
//   rows[Nint][NY][CRW]      BABD row  [G | C | H | b]  (G x_left + C x_right + H p = b)
//   rec[Nint][REC]           condensation record [U (NS x NVA) | T (M NY x NVA) | G (M NZ x NVA)],
//                            float64 (float32 for the FP32 direction)
//   crrec[Nint-1][NY][RECW]  cyclic-reduction record [R | Gp | Ep | Jp | d]
#pragma once
#include "colloc_common.h"
#include "colloc_gj.h"

namespace colloc {

constexpr int NY = Cfg::NY, NZ = Cfg::NZ, NP = Cfg::NP, NC = Cfg::NC, M = Cfg::M;
constexpr int NZ1 = NZ > 0 ? NZ : 1;
constexpr int NP1 = NP > 0 ? NP : 1;
constexpr int NC1 = NC > 0 ? NC : 1;
constexpr int NYZ = NY + NZ;
constexpr int NJC = NY + NZ + NP;          // columns of Df / Dg
constexpr int JSTAGE = (NY + NZ) * NJC;    // doubles per stage Jacobian block
constexpr int NPA = Cfg::NPA;              // parameters the stage functions depend on
constexpr int NVA = NY + NPA + 1;          // condensed columns [dy_i | dp (active) | 1]
constexpr int NS = (M - 1) * NY;           // size of the condensed stage system
constexpr int REC = (NS + M * NY + M * NZ) * NVA;
constexpr int CRW = 2 * NY + NP + 1;       // BABD row width
constexpr int RECW = 3 * NY + NP + 1;      // cyclic-reduction record width
constexpr int NFIN = 2 * NY + NP;          // boundary (final) system size
constexpr int CR_PARTS = 4;                          // lanes per record column in cr_pair (one per 16-lane row)
constexpr int CR_COLS = WAVE / CR_PARTS;              // columns per wave
constexpr int CR_ROWS = (2 * NY + CR_PARTS - 1) / CR_PARTS;   // rows per lane
// No stage function depends on a parameter (NPA = 0): every H block of the
// BABD rows is identically zero at every level (the condensation gives
// H = 0 and both pairings keep it), so the kernels neither form, read nor
// write it; NPH is the count of H columns they carry.
constexpr bool HZ = NPA == 0;
constexpr int NPH = HZ ? 0 : NP;
constexpr int RECW_L = 3 * NY + NPH + 1;   // record columns cr_pair carries
constexpr int CR_THREADS = (RECW_L + CR_COLS - 1) / CR_COLS * WAVE;

static_assert(M >= 2, "need at least two collocation stages");

// Parameter j of the condensed column q = NY + j (j < NPA) is Cfg::PACT[j].
__device__ __forceinline__ int pact(int j) { return Cfg::PACT[j]; }
__host__ __device__ constexpr bool p_active(int j) {
  for (int a = 0; a < NPA; ++a)
    if (Cfg::PACT[a] == j) return true;
  return false;
}
static_assert(CR_PARTS == 4, "cr_pair sums over the four 16-lane rows of a wave");
#if !COLLOC_LARGE
static_assert(NY <= 64, "cyclic-reduction kernels hold one row per lane");
#endif

constexpr int DRSZ = (NY + NP) * (2 * NY + NP);   // one instance's Dr
constexpr int NRB = NY + NP;                       // one instance's r_bc

// A batch of B >= 1 instances of the problem laid end to end.  Instance b
// owns nodes [noff[b], noff[b+1]) of the node arrays (t, y, the residual of
// segres, z) and the intervals [noff[b] - b, noff[b+1] - b - 1) of the
// interval / stage arrays; iinst[i] is the instance of interval i (its left
// node is i + iinst[i]) and ninst[n] the instance of node n.  p, c, r_bc and
// Dr hold one block per instance.  act (may be null) marks the instances a
// kernel works on; the others are skipped and their outputs left as they are.
// ilist (may be null) lists the intervals a per-interval or per-stage kernel
// is launched over (nlist of them, those of the instances still iterating),
// so that converged instances cost no workgroups; null: all intervals.  When
// the list is built on the device (the graph-replayed Newton iteration), its
// length is read from *nlistp and the launch covers every interval, the
// work units past the list returning at once.
struct Batch {
  const int* iinst;
  const int* ninst;
  const int* noff;
  const int* act;
  int B;
  const int* ilist;
  int nlist;
  const int* nlistp;
};
__device__ __forceinline__ bool skipped(const Batch& bt, int b) { return bt.act && !bt.act[b]; }
// Intervals a launch covers and the interval of its k-th work unit.
__device__ __forceinline__ int launch_intervals(const Batch& bt, int nint) {
  return bt.ilist ? (bt.nlistp ? *bt.nlistp : bt.nlist) : nint;
}
__device__ __forceinline__ int listed_interval(const Batch& bt, int k) { return bt.ilist ? bt.ilist[k] : k; }
// A workgroup-per-interval kernel's block past the launch list.
__device__ __forceinline__ bool beyond_list(const Batch& bt, int nint) {
  return (int)blockIdx.x >= launch_intervals(bt, nint);
}

// One CR level of one instance: its rows start at in_off of the input
// buffer, its ceil(nrows / 2) output rows at out_off, its nrows / 2 records
// at rec_off; subst != 0: the level pairs by substitution (cr_subst), else
// by Householder QR (cr_pair); subst == 2: the instance's next level pairs
// by substitution too, so the output rows' C block (I) is left unwritten.
struct CRSeg {
  int in_off, out_off, rec_off, nrows, subst;
};

// -------------------------------------------------------------------- fill
__device__ __forceinline__ void fill_u64(unsigned long long* p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// max|v| per instance folded into slot[b] (u64 slots, as block_max_to).  b < 0:
// the lane has nothing to add.  The lanes that share the instance of the
// wave's first contributing lane reduce together; any other lane (a wave
// spanning an instance boundary) adds its own value.  Every lane of the wave
// must call this.
__device__ __forceinline__ void seg_max_to(double v, int b, unsigned long long* __restrict__ slot) {
  const unsigned long long m = __ballot(b >= 0);
  if (m == 0) return;
  const int first = __ffsll((long long)m) - 1;
  const int b0 = __shfl(b, first);
  double w = (b == b0) ? v : 0.0;
  w = wave_max(w);
  if (lane_id() == first) atomicMax(slot + b0, (unsigned long long)__double_as_longlong(fabs(w)));
  if (b >= 0 && b != b0) atomicMax(slot + b, (unsigned long long)__double_as_longlong(fabs(v)));
}

#if !COLLOC_TABLE_FORM
// ---------------------------------------------------------------- residual
// One thread per (interval, stage); thread Nint*M + b evaluates r of
// instance b.  Restates collocation_update_kernel + compute_f_q_kernel1/2
// (solver:478-559) and the host max-norm of f_q (solver:116), one norm per
// instance folded into norm[b].
template <bool FULL>
__device__ __forceinline__ void residual(int nint, Batch bt, const double* __restrict__ t,
                         const double* __restrict__ y, const double* __restrict__ ydot,
                         const double* __restrict__ ztil, const double* __restrict__ p,
                         const double* __restrict__ c, double alpha,
                         double* __restrict__ ytil, double* __restrict__ fa, double* __restrict__ fb,
                         double* __restrict__ rbc, unsigned long long* __restrict__ norm) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nst = launch_intervals(bt, nint) * M;
  double amax = 0.0;
  int b = -1;
  int i = 0, j = 0, st = 0;   // interval, stage, stage row
  if (gid < nst) {
    i = listed_interval(bt, gid / M);
    j = gid - (gid / M) * M;
    st = i * M + j;
    b = bt.iinst[i];
    if (skipped(bt, b)) b = -1;
  }
  if (b >= 0) {
    const int nd = i + b;
    const double h = t[nd + 1] - t[nd];
    const double* yi = y + (size_t)nd * NY;
    const double* yd = ydot + (size_t)i * M * NY;
    const double* pb = p + (size_t)b * NP;
    const double* cb = c + (size_t)b * NC;
    double yt[NY], zt[NZ1], fo[NY], go[NZ1];
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) s += c_rkA[j * M + k] * yd[k * NY + l];
      yt[l] = yi[l] + h * s;
    }
#pragma unroll
    for (int l = 0; l < NZ; ++l) zt[l] = ztil[(size_t)st * NZ + l];
    prob::f(yt, zt, pb, cb, fo);
    prob::g(yt, zt, pb, cb, alpha, go);
    double* fai = fa + (size_t)st * NYZ;
#pragma unroll
    for (int l = 0; l < NY; ++l) {
      const double v = fo[l] - yd[j * NY + l];
      amax = nmax(fabs(v), amax);
      if (FULL) fai[l] = v;
    }
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
      amax = nmax(fabs(go[l]), amax);
      if (FULL) fai[NY + l] = go[l];
    }
    if (FULL) {
#pragma unroll
      for (int l = 0; l < NY; ++l) ytil[(size_t)st * NY + l] = yt[l];
    }
    if (j == M - 1) {
#pragma unroll
      for (int l = 0; l < NY; ++l) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) s += c_rkb[k] * yd[k * NY + l];
        const double v = y[(size_t)(nd + 1) * NY + l] - yi[l] - h * s;
        amax = nmax(fabs(v), amax);
        if (FULL) fb[(size_t)i * NY + l] = v;
      }
    }
  } else if (gid >= nst && gid < nst + bt.B && !skipped(bt, gid - nst)) {
    b = gid - nst;
    double ro[NY + NP];
    prob::r(y + (size_t)bt.noff[b] * NY, y + (size_t)(bt.noff[b + 1] - 1) * NY, p + (size_t)b * NP,
            c + (size_t)b * NC, ro);
#pragma unroll
    for (int l = 0; l < NY + NP; ++l) {
      amax = nmax(fabs(ro[l]), amax);
      if (FULL) rbc[(size_t)b * NRB + l] = ro[l];
    }
  }
  seg_max_to(amax, b, norm);
}

// ---------------------------------------------------------------- jacobian
// The values of a wave's 64 stages go to the stage blocks in groups of
// JAC_CHUNK = 16 (prob::jac_part<k> hands them on in that order): each lane
// puts its 16 values in LDS, then lane l stores value l % 16 of the stages
// (l / 16) + 4 t, t < 16, so 16 consecutive lanes write 16 entries of one
// stage (consecutive addresses where the entries are) instead of 64 lanes
// writing 64 stage blocks apart.
template <class TJ>
struct JacSink {
  TJ* jac;
  TJ* buf;                       // this wave's LDS: [WAVE][JAC_CHUNK + 1]
  size_t st_t[16];               // stage of the lanes (l / 16) + 4 t of this wave
  unsigned valid;                // bit t: that stage is evaluated
  template <int N>
  __device__ __forceinline__ void put(int first, const TJ* v) {
    static_assert(prob::JAC_CHUNK == 16, "four 16-lane rows per store");
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) buf[lane * 17 + j] = v[j];
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");   // one wave: the LDS writes above precede the reads below
    const int j = lane & 15;
    if (j < N) {
      const size_t k = (size_t)prob::jac_dyn_idx[first + j];
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (valid & (1u << t)) jac[st_t[t] * JSTAGE + k] = buf[((lane >> 4) + 4 * t) * 17 + j];
    }
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");   // and the reads precede the next put's writes
  }
};

// prob::JAC_PARTS threads share a stage: workgroup g evaluates part
// g % JAC_PARTS of the dynamic entries (the same part in the whole
// workgroup, so no divergence) for the stages of workgroup g / JAC_PARTS.
// One thread per stage; thread Nint*M + b of part 0 writes Dr of instance b.
// Restates construct_jacobian_kernel (solver:680-797) without materialising
// J, V, D, W (the condensation kernel assembles what it needs).  TJ: the
// stored type of the stage Jacobians (double, or float for the FP32
// direction), which is also the type their state-dependent entries are
// evaluated in.
template <class TJ>
__device__ __forceinline__ void jacobian(int nint, Batch bt, const double* __restrict__ y,
                         const double* __restrict__ ytil, const double* __restrict__ ztil,
                         const double* __restrict__ p, const double* __restrict__ c, double alpha,
                         TJ* __restrict__ jac, double* __restrict__ Dr) {
  __shared__ TJ jbuf[256 / WAVE][WAVE * 17];
  const int part = blockIdx.x % prob::JAC_PARTS;
  const int gid = (blockIdx.x / prob::JAC_PARTS) * blockDim.x + threadIdx.x;
  const int nst = launch_intervals(bt, nint) * M;
  bool valid = gid < nst;
  int st = 0, b = 0;
  if (valid) {
    const int i = listed_interval(bt, gid / M);
    st = i * M + (gid - (gid / M) * M);
    b = bt.iinst[i];
    valid = !skipped(bt, b);
    if (!valid) { st = 0; b = 0; }
  }
  if (__ballot(valid) != 0) {   // the whole wave runs the part (its lanes exchange values)
    JacSink<TJ> sink;
    sink.jac = jac;
    sink.buf = jbuf[wave_id()];
    sink.valid = 0;
    const int lane = lane_id();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int src = (lane >> 4) + 4 * t;
      sink.st_t[t] = (size_t)__shfl(st, src);
      if (__shfl((int)valid, src)) sink.valid |= 1u << t;
    }
    double yt[NY], zt[NZ1];
#pragma unroll
    for (int l = 0; l < NY; ++l) yt[l] = ytil[(size_t)st * NY + l];
#pragma unroll
    for (int l = 0; l < NZ; ++l) zt[l] = ztil[(size_t)st * NZ + l];
    // the dynamic entries only: the static ones (zeros, compile-time
    // constants) are in place since jacobian_static filled the buffer
    prob::jac_dyn_part<TJ>(part, yt, zt, p + (size_t)b * NP, c + (size_t)b * NC, alpha, sink);
  }
  if (gid >= nst && gid < nst + bt.B && part == 0) {
    const int ib = gid - nst;
    if (skipped(bt, ib)) return;
    prob::Dr(y + (size_t)bt.noff[ib] * NY, y + (size_t)(bt.noff[ib + 1] - 1) * NY, p + (size_t)ib * NP,
             c + (size_t)ib * NC, Dr + (size_t)ib * DRSZ);
  }
}

#endif  // !COLLOC_TABLE_FORM

// The static entries of every stage Jacobian block (prob::jac_static_mask /
// jac_static_val: entries that depend on nothing but the problem's
// compile-time constants) for stages [0, nst) of the buffer, written once per
// buffer and stored type; workgroup g fills stages g, g + G, ... with
// consecutive lanes on consecutive entries.
template <class TJ>
__device__ __forceinline__ void jacobian_static(long long nst, TJ* __restrict__ jac) {
  for (long long st = blockIdx.x; st < nst; st += gridDim.x) {
    TJ* out = jac + (size_t)st * JSTAGE;
    for (int k = threadIdx.x; k < JSTAGE; k += blockDim.x)
      if (prob::jac_static_mask[k]) out[k] = (TJ)prob::jac_static_val[k];
  }
}

#if !COLLOC_LARGE
// Stage eliminations of one interval (the first phase of both condensation
// kernels): wave j eliminates the algebraic unknowns of stage j, computing
// in R (double, or float for the FP32 direction).  Writes T_j to sT (LDS,
// rows of NVA) and recT, G_j to recG (records of type TR); returns nonzero on
// a zero pivot in some g_z.
template <class R, int NW, class TJ, class TR>
__device__ __forceinline__ int stage_eliminate(int i, const TJ* __restrict__ jac, const double* __restrict__ fa,
                                               R* sT, TR* __restrict__ recT, TR* __restrict__ recG) {
  const int lane = lane_id(), wave = wave_id();
  int bad = 0;
  // Lane l holds columns l + 64 s of [g_z | g_y g_p f_a,g] (NZ rows each);
  // Gauss-Jordan with partial pivoting and physical row swaps, the pivot
  // column broadcast by v_readlane, so the wave needs no barrier.  The same
  // wave then forms T_j = [h_y | h_p | f_a,h] - h_z G_j for its columns.
  constexpr int NCZ = NZ + NVA;
  constexpr int SL = (NCZ + WAVE - 1) / WAVE;
  for (int j = wave; j < M; j += NW) {
    const TJ* Dfj = jac + ((size_t)i * M + j) * JSTAGE;
    const TJ* Dgj = Dfj + NY * NJC;
    const double* faj = fa + (size_t)i * M * NYZ + j * NYZ;
    R gc[SL][NZ1];
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const int cc = lane + WAVE * sl;
#pragma unroll
      for (int r = 0; r < NZ; ++r) {
        double v = 0.0;   // loaded as stored, kept in R
        if (cc < NZ) v = (double)Dgj[r * NJC + NY + cc];
        else if (cc < NCZ) {
          const int q = cc - NZ;
          v = q < NY ? (double)Dgj[r * NJC + q]
                     : (q < NY + NPA ? (double)Dgj[r * NJC + NY + NZ + pact(q - NY)] : faj[NY + r]);
        }
        gc[sl][r] = (R)v;
      }
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      R ck[NZ1];
#pragma unroll
      for (int r = 0; r < NZ; ++r) ck[r] = readlane_r(gc[0][r], k);
      int pr = k;
      R best = fabs(ck[k]);
#pragma unroll
      for (int r = k + 1; r < NZ; ++r)
        if (fabs(ck[r]) > best) { best = fabs(ck[r]); pr = r; }
      // exchange rows k and pr by per-row conditional swaps (at most one fires)
#pragma unroll
      for (int r = k + 1; r < NZ; ++r) {
        const bool sw = r == pr;
        const R t0 = ck[r];
        ck[r] = sw ? ck[k] : ck[r];
        ck[k] = sw ? t0 : ck[k];
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const R t1 = gc[sl][r];
          gc[sl][r] = sw ? gc[sl][k] : gc[sl][r];
          gc[sl][k] = sw ? t1 : gc[sl][k];
        }
      }
      const R cpr = ck[k];
      if (!(fabs(cpr) > (R)0)) bad |= 1;
      const R inv = (R)1 / cpr;
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        const R pk = gc[sl][k] * inv;
#pragma unroll
        for (int r = 0; r < NZ; ++r) gc[sl][r] = (r == k) ? pk : fma(-ck[r], pk, gc[sl][r]);
      }
    }
    // gc now holds [I | G_j]; write G_j and form T_j column by column
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const int q = lane + WAVE * sl - NZ;
      if (q < 0 || q >= NVA) continue;
#pragma unroll
      for (int r = 0; r < NZ; ++r) recG[(j * NZ + r) * NVA + q] = (TR)gc[sl][r];
    }
    // T_j = [h_y | h_p | f_a,h] - h_z G_j, rows in batches of TB whose
    // Jacobian loads are all issued before any of them is used
    constexpr int TB = 5;
    for (int r0 = 0; r0 < NY; r0 += TB) {
      R hz[TB][NZ1], hb[TB][SL];
#pragma unroll
      for (int rr = 0; rr < TB; ++rr) {
        const int r = r0 + rr < NY ? r0 + rr : NY - 1;
        const TJ* hrow = Dfj + r * NJC;
#pragma unroll
        for (int z = 0; z < NZ; ++z) hz[rr][z] = (R)hrow[NY + z];
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const int q = lane + WAVE * sl - NZ;
          // h_y, then the active h_p columns after the NZ z-columns
          const int src = q < NY ? q : NY + NZ + pact(q - NY < NPA ? q - NY : 0);
          hb[rr][sl] = (R)((q >= 0 && q < NY + NPA) ? (double)hrow[src] : (q == NY + NPA ? faj[r] : 0.0));
        }
      }
#pragma unroll
      for (int rr = 0; rr < TB; ++rr) {
        const int r = r0 + rr;
        if (r >= NY) continue;
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const int q = lane + WAVE * sl - NZ;
          if (q < 0 || q >= NVA) continue;
          R v = hb[rr][sl];
#pragma unroll
          for (int z = 0; z < NZ; ++z) v = fma(-hz[rr][z], gc[sl][z], v);
          sT[(j * NY + r) * NVA + q] = v;
          recT[(j * NY + r) * NVA + q] = (TR)v;
        }
      }
    }
  }
  return bad;
}

// Stage eliminations with every global load issued up front: rows [r0, r1)
// of stage j of interval i (the whole stage when r0 = 0, r1 = NY), computing
// in R.  Each lane slice holds one column of [g_z | g_y g_p f_a,g] / of
// [h_z | h_y h_p f_a,h] (same column map for the g and h rows), loaded from
// a clamped, always valid address so that no load sits behind a branch;
// h_z of the rows and f_a of the stage go to the wave's LDS scratch `scr`
// (SE_SCR values) and are read back as broadcasts.  RL: a compile-time bound
// on r1 - r0.  G_j is written when write_g.  Returns nonzero on a zero pivot.
constexpr int SE_SCR = NY * NZ1 + NYZ;
template <class R, int RL, class TJ, class TR>
__device__ __forceinline__ int stage_rows(int i, int j, int r0, int r1, bool write_g, const TJ* __restrict__ jac,
                                          const double* __restrict__ fa, R* sT, TR* __restrict__ recT,
                                          TR* __restrict__ recG, R* scr) {
  constexpr int NCZ = NZ + NVA;
  constexpr int SL = (NCZ + WAVE - 1) / WAVE;
  const int lane = lane_id();
  const TJ* Dfj = jac + ((size_t)i * M + j) * JSTAGE;
  const TJ* Dgj = Dfj + NY * NJC;
  const double* faj = fa + (size_t)i * M * NYZ + j * NYZ;
  int src[SL];
  bool rhs[SL];
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) {
    const int cc = lane + WAVE * sl, q = cc - NZ;
    src[sl] = cc < NZ ? NY + cc : (q < NY ? q : (q < NY + NPA ? NY + NZ + pact(q - NY < NPA ? q - NY : 0) : 0));
    rhs[sl] = q == NY + NPA;
  }
  // every load first: g rows, this range's h rows, h_z, f_a
  R gc[SL][NZ1];
  R hc[RL][SL];
#pragma unroll
  for (int r = 0; r < NZ; ++r)
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) gc[sl][r] = (R)Dgj[r * NJC + src[sl]];
#pragma unroll
  for (int rr = 0; rr < RL; ++rr) {
    const int r = r0 + rr < NY ? r0 + rr : NY - 1;
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) hc[rr][sl] = (R)Dfj[r * NJC + src[sl]];
  }
  constexpr int NHZ = RL * NZ;
  R hz[(NHZ + WAVE - 1) / WAVE];
  R fv[(NYZ + WAVE - 1) / WAVE];
#pragma unroll
  for (int k = 0; k < (NHZ + WAVE - 1) / WAVE; ++k) {
    const int e = lane + WAVE * k;
    const int r = r0 + e / NZ1 < NY ? r0 + e / NZ1 : NY - 1;
    hz[k] = (R)Dfj[r * NJC + NY + (e % NZ1 < NZ ? e % NZ1 : 0)];
  }
#pragma unroll
  for (int k = 0; k < (NYZ + WAVE - 1) / WAVE; ++k) {
    const int e = lane + WAVE * k;
    fv[k] = (R)faj[e < NYZ ? e : 0];
  }
#pragma unroll
  for (int k = 0; k < (NHZ + WAVE - 1) / WAVE; ++k)
    if (lane + WAVE * k < NHZ) scr[lane + WAVE * k] = hz[k];
#pragma unroll
  for (int k = 0; k < (NYZ + WAVE - 1) / WAVE; ++k)
    if (lane + WAVE * k < NYZ) scr[NY * NZ1 + lane + WAVE * k] = fv[k];
  // the rhs column of the g rows from f_a,g
#pragma unroll
  for (int sl = 0; sl < SL; ++sl)
#pragma unroll
    for (int r = 0; r < NZ; ++r)
      if (rhs[sl]) gc[sl][r] = readlane_r(fv[(NY + r) / WAVE], (NY + r) % WAVE);
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");   // the scratch writes above happen before the reads below (one wave: LDS in order)
  // Gauss-Jordan on [g_z | ...] with partial pivoting (lanes = columns)
  int bad = 0;
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    R ck[NZ1];
#pragma unroll
    for (int r = 0; r < NZ; ++r) ck[r] = readlane_r(gc[0][r], k);
    int pr = k;
    R best = fabs(ck[k]);
#pragma unroll
    for (int r = k + 1; r < NZ; ++r)
      if (fabs(ck[r]) > best) { best = fabs(ck[r]); pr = r; }
#pragma unroll
    for (int r = k + 1; r < NZ; ++r) {
      const bool sw = r == pr;
      const R t0 = ck[r];
      ck[r] = sw ? ck[k] : ck[r];
      ck[k] = sw ? t0 : ck[k];
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        const R t1 = gc[sl][r];
        gc[sl][r] = sw ? gc[sl][k] : gc[sl][r];
        gc[sl][k] = sw ? t1 : gc[sl][k];
      }
    }
    const R cpr = ck[k];
    if (!(fabs(cpr) > (R)0)) bad |= 1;
    const R inv = (R)1 / cpr;
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const R pk = gc[sl][k] * inv;
#pragma unroll
      for (int r = 0; r < NZ; ++r) gc[sl][r] = (r == k) ? pk : fma(-ck[r], pk, gc[sl][r]);
    }
  }
  if (write_g) {
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const int q = lane + WAVE * sl - NZ;
      if (q < 0 || q >= NVA) continue;
#pragma unroll
      for (int r = 0; r < NZ; ++r) recG[(j * NZ + r) * NVA + q] = (TR)gc[sl][r];
    }
  }
  // T_j rows: [h_y | h_p | f_a,h] - h_z G_j
#pragma unroll
  for (int rr = 0; rr < RL; ++rr) {
    const int r = r0 + rr;
    if (r >= r1) break;
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const int q = lane + WAVE * sl - NZ;
      R v = rhs[sl] ? scr[NY * NZ1 + r] : hc[rr][sl];
#pragma unroll
      for (int z = 0; z < NZ; ++z) v = fma(-scr[rr * NZ + z], gc[sl][z], v);
      if (q >= 0 && q < NVA) {
        sT[(j * NY + r) * NVA + q] = v;
        recT[(j * NY + r) * NVA + q] = (TR)v;
      }
    }
  }
  return bad;
}

// The stage eliminations of one interval over NW waves: with NW >= M, wave j
// does stage j; otherwise every wave does whole stages w, w + NW, ... below
// (M / NW) NW, and the remaining stages are split by rows over all waves
// (each wave repeats that stage's small g_z elimination).  scr: NW * SE_SCR
// values of LDS.
template <class R, int NW, class TJ, class TR>
__device__ __forceinline__ int stage_eliminate_all(int i, const TJ* __restrict__ jac, const double* __restrict__ fa,
                                                   R* sT, TR* __restrict__ recT, TR* __restrict__ recG, R* scr) {
  const int wave = wave_id();
  R* my = scr + wave * SE_SCR;
  int bad = 0;
  if constexpr (NW >= M) {
    if (wave < M) bad |= stage_rows<R, NY, TJ, TR>(i, wave, 0, NY, true, jac, fa, sT, recT, recG, my);
  } else {
    constexpr int FULL = (M / NW) * NW;
    constexpr int RS = (NY + NW - 1) / NW;
    for (int j = wave; j < FULL; j += NW)
      bad |= stage_rows<R, NY, TJ, TR>(i, j, 0, NY, true, jac, fa, sT, recT, recG, my);
    for (int j = FULL; j < M; ++j) {
      const int r0 = wave * RS < NY ? wave * RS : NY;
      const int r1 = r0 + RS < NY ? r0 + RS : NY;
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");   // the previous stage's scratch reads are done
      bad |= stage_rows<R, RS, TJ, TR>(i, j, r0, r1, wave == 0, jac, fa, sT, recT, recG, my);
    }
  }
  return bad;
}

// stage_rows for float32 with T_j = [h_y | h_p | f_a,h] - h_z G_j on matrix
// cores: h_z rows as A operands (row 16 ri + l % 16, column 4 ks + l / 16),
// G_j as B operands from the wave's LDS copy, and [h_y | h_p | f_a,h] loaded
// straight into the accumulator layout (lane l, value v of tile (ri, ci):
// row r0 + 16 ri + 4 (l / 16) + v, column 16 ci + l % 16).  The g_z
// elimination is stage_rows'.  scr: SE_SCR_MF floats of LDS.
constexpr int SE_KZ = (NZ + 3) / 4;                 // k-steps over the algebraic unknowns
constexpr int SE_GW = ((NVA + 15) / 16) * 16;       // padded columns of the LDS copy of G_j
constexpr int SE_SCR_MF = NYZ + SE_KZ * 4 * SE_GW;
// The loads of stage_rows_mf: the g rows (lanes = columns), h_z as A
// operands, [h_y | h_p] in the accumulator layout, f_a.  Split from the
// arithmetic so that a wave can have two stages' loads in flight.
typedef float se_f4 __attribute__((ext_vector_type(4)));
template <int RL>
struct StageLoadMF {
  static constexpr int SL = (NZ + NVA + WAVE - 1) / WAVE;
  static constexpr int RTL = (RL + 15) / 16;
  static constexpr int CT = SE_GW / 16;
  static constexpr int FV = (NYZ + WAVE - 1) / WAVE;
  float gc[SL][NZ1];
  float az[RTL][SE_KZ];
  se_f4 hv[RTL][CT];
  float fv[FV];
};

__device__ __forceinline__ int se_src(int cc) {   // Jacobian column of [g_z | g_y | g_pa | -] column cc
  const int q = cc - NZ;
  return cc < NZ ? NY + cc : (q < NY ? q : (q < NY + NPA ? NY + NZ + pact(q - NY < NPA ? q - NY : 0) : 0));
}

template <int RL>
__device__ __forceinline__ void stage_load_mf(int i, int j, int r0, int r1, const float* __restrict__ jac,
                                              const double* __restrict__ fa, StageLoadMF<RL>& L) {
  typedef StageLoadMF<RL> SL_;
  const int lane = lane_id(), g = lane >> 4, c16 = lane & 15;
  const float* Dfj = jac + ((size_t)i * M + j) * JSTAGE;
  const float* Dgj = Dfj + NY * NJC;
  const double* faj = fa + (size_t)i * M * NYZ + j * NYZ;
#pragma unroll
  for (int r = 0; r < NZ; ++r)
#pragma unroll
    for (int sl = 0; sl < SL_::SL; ++sl) L.gc[sl][r] = Dgj[r * NJC + se_src(lane + WAVE * sl)];
#pragma unroll
  for (int ri = 0; ri < SL_::RTL; ++ri)
#pragma unroll
    for (int ks = 0; ks < SE_KZ; ++ks) {
      const int row = r0 + 16 * ri + c16, z = 4 * ks + g;
      const float v = Dfj[(row < NY ? row : NY - 1) * NJC + NY + (z < NZ ? z : 0)];
      L.az[ri][ks] = (row < r1 && z < NZ) ? v : 0.0f;
    }
#pragma unroll
  for (int ri = 0; ri < SL_::RTL; ++ri)
#pragma unroll
    for (int ci = 0; ci < SL_::CT; ++ci) {
      const int q = 16 * ci + c16;
      const int sq = q < NY ? q : (q < NY + NPA ? NY + NZ + pact(q - NY < NPA ? q - NY : 0) : 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + 16 * ri + 4 * g + v;
        L.hv[ri][ci][v] = Dfj[(row < NY ? row : NY - 1) * NJC + sq];
      }
    }
#pragma unroll
  for (int k = 0; k < SL_::FV; ++k) {
    const int e = lane + WAVE * k;
    L.fv[k] = (float)faj[e < NYZ ? e : 0];
  }
}

// stage_rows for float32 with T_j = [h_y | h_p | f_a,h] - h_z G_j on matrix
// cores, from the loads of stage_load_mf: h_z rows as A operands (row
// 16 ri + l % 16, column 4 ks + l / 16), G_j as B operands from the wave's
// LDS copy, and [h_y | h_p | f_a,h] in the accumulator layout (lane l,
// value v of tile (ri, ci): row r0 + 16 ri + 4 (l / 16) + v, column
// 16 ci + l % 16).  The g_z elimination is stage_rows'.  scr: SE_SCR_MF
// floats of LDS.
template <int RL, class TR>
__device__ __forceinline__ int stage_compute_mf(int i, int j, int r0, int r1, bool write_g, StageLoadMF<RL>& L,
                                                float* sT, TR* __restrict__ recT, TR* __restrict__ recG, float* scr) {
  typedef StageLoadMF<RL> SL_;
  constexpr int SL = SL_::SL;
  constexpr int RTL = SL_::RTL;
  constexpr int CT = SL_::CT;
  const int lane = lane_id(), g = lane >> 4, c16 = lane & 15;
  float* sfa = scr;             // f_a of the stage
  float* sG = scr + NYZ;        // G_j, SE_KZ * 4 rows of SE_GW
  auto& gc = L.gc;
  bool rhs[SL];
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) rhs[sl] = lane + WAVE * sl - NZ == NY + NPA;
#pragma unroll
  for (int k = 0; k < SL_::FV; ++k)
    if (lane + WAVE * k < NYZ) sfa[lane + WAVE * k] = L.fv[k];
#pragma unroll
  for (int sl = 0; sl < SL; ++sl)
#pragma unroll
    for (int r = 0; r < NZ; ++r)
      if (rhs[sl]) gc[sl][r] = readlane_r(L.fv[(NY + r) / WAVE], (NY + r) % WAVE);
  // Gauss-Jordan on [g_z | ...] with partial pivoting (lanes = columns)
  int bad = 0;
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    float ck[NZ1];
#pragma unroll
    for (int r = 0; r < NZ; ++r) ck[r] = readlane_r(gc[0][r], k);
    int pr = k;
    float best = fabsf(ck[k]);
#pragma unroll
    for (int r = k + 1; r < NZ; ++r)
      if (fabsf(ck[r]) > best) { best = fabsf(ck[r]); pr = r; }
#pragma unroll
    for (int r = k + 1; r < NZ; ++r) {
      const bool sw = r == pr;
      const float t0 = ck[r];
      ck[r] = sw ? ck[k] : ck[r];
      ck[k] = sw ? t0 : ck[k];
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        const float t1 = gc[sl][r];
        gc[sl][r] = sw ? gc[sl][k] : gc[sl][r];
        gc[sl][k] = sw ? t1 : gc[sl][k];
      }
    }
    const float cpr = ck[k];
    if (!(fabsf(cpr) > 0.0f)) bad |= 1;
    const float inv = 1.0f / cpr;
#pragma unroll
    for (int sl = 0; sl < SL; ++sl) {
      const float pk = gc[sl][k] * inv;
#pragma unroll
      for (int r = 0; r < NZ; ++r) gc[sl][r] = (r == k) ? pk : fmaf(-ck[r], pk, gc[sl][r]);
    }
  }
  // G_j to the records and to the LDS copy (zero padding rows and columns)
  for (int e = lane; e < SE_KZ * 4 * SE_GW; e += WAVE) sG[e] = 0.0f;
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) {
    const int q = lane + WAVE * sl - NZ;
    if (q < 0 || q >= NVA) continue;
#pragma unroll
    for (int r = 0; r < NZ; ++r) {
      sG[r * SE_GW + q] = gc[sl][r];
      if (write_g) recG[(j * NZ + r) * NVA + q] = (TR)gc[sl][r];
    }
  }
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");   // one wave: the LDS writes above precede the reads below
  float bg[SE_KZ][CT];
#pragma unroll
  for (int ks = 0; ks < SE_KZ; ++ks)
#pragma unroll
    for (int ci = 0; ci < CT; ++ci) bg[ks][ci] = sG[(4 * ks + g) * SE_GW + 16 * ci + c16];
  // T = H - h_z G, tile by tile
#pragma unroll
  for (int ri = 0; ri < RTL; ++ri)
#pragma unroll
    for (int ci = 0; ci < CT; ++ci) {
      se_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < SE_KZ; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(L.az[ri][ks], bg[ks][ci], acc, 0, 0, 0);
      const int q = 16 * ci + c16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + 16 * ri + 4 * g + v;
        if (row >= r1 || q >= NVA) continue;
        const float h = (q == NY + NPA) ? sfa[row] : L.hv[ri][ci][v];
        const float tv = h - acc[v];
        sT[(j * NY + row) * NVA + q] = tv;
        recT[(j * NY + row) * NVA + q] = (TR)tv;
      }
    }
  return bad;
}

// stage_eliminate_all for the matrix-core kernel (float32): the same
// assignment of stages and row ranges to waves.  With one stage per wave
// and one stage split by rows (m = NW + 1), both stages' loads are issued
// before either is computed.  scr: NW * SE_SCR_MF floats.
template <int NW, class TR>
__device__ __forceinline__ int stage_eliminate_mf(int i, const float* __restrict__ jac, const double* __restrict__ fa,
                                                  float* sT, TR* __restrict__ recT, TR* __restrict__ recG, float* scr) {
  const int wave = wave_id();
  float* my = scr + wave * SE_SCR_MF;
  int bad = 0;
  if constexpr (NW >= M) {
    if (wave < M) {
      StageLoadMF<NY> L;
      stage_load_mf<NY>(i, wave, 0, NY, jac, fa, L);
      bad |= stage_compute_mf<NY, TR>(i, wave, 0, NY, true, L, sT, recT, recG, my);
    }
  } else {
    constexpr int FULL = (M / NW) * NW;
    constexpr int RS = (NY + NW - 1) / NW;
    const int r0 = wave * RS < NY ? wave * RS : NY;
    const int r1 = r0 + RS < NY ? r0 + RS : NY;
    if constexpr (FULL == NW && M - FULL == 1) {
      StageLoadMF<NY> L;
      StageLoadMF<RS> E;
      stage_load_mf<NY>(i, wave, 0, NY, jac, fa, L);
      stage_load_mf<RS>(i, FULL, r0, r1, jac, fa, E);
      bad |= stage_compute_mf<NY, TR>(i, wave, 0, NY, true, L, sT, recT, recG, my);
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");   // the previous stage's scratch reads are done
      bad |= stage_compute_mf<RS, TR>(i, FULL, r0, r1, wave == 0, E, sT, recT, recG, my);
    } else {
      for (int j = wave; j < FULL; j += NW) {
        StageLoadMF<NY> L;
        stage_load_mf<NY>(i, j, 0, NY, jac, fa, L);
        bad |= stage_compute_mf<NY, TR>(i, j, 0, NY, true, L, sT, recT, recG, my);
      }
      for (int j = FULL; j < M; ++j) {
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");   // the previous stage's scratch reads are done
        StageLoadMF<RS> E;
        stage_load_mf<RS>(i, j, r0, r1, jac, fa, E);
        bad |= stage_compute_mf<RS, TR>(i, j, r0, r1, wave == 0, E, sT, recT, recG, my);
      }
    }
  }
  return bad;
}

// --------------------------------------------------------------- condense
// One workgroup per interval.  Condenses the stage unknowns out of the
// interval's linearised collocation equations (the role of
// reduce_jacobian_parallel_kernel, solver:887-1001) using the stage
// structure of W instead of a dense K x K LU:
//   per stage j:  G_j = g_z^-1 [g_y | g_p | f_a,g]        (NZ x NVA)
//                 T_j = [h_y | h_p | f_a,h] - h_z G_j      (NY x NVA)
//   so that  dz_j = -G_j [u_j; dp; 1],  dydot_j = T_j [u_j; dp; 1]
//   with u_j the stage-state increment.  Lobatto IIIA has a_0k = 0, so
//   u_0 = dy_i and the unknowns u_1..u_{M-1} satisfy
//     u_j - h sum_{k>=1} a_jk M_k u_k = (E_y + h a_j0 T_0 + h sum_{k>=1} a_jk [0|N_k|s_k]) v
//   with M_k = T_k[:, :NY], [N_k|s_k] = T_k[:, NY:], v = [dy_i; dp; 1].  That
//   NS x NS system with NVA right-hand sides is solved by register-resident
//   Gauss-Jordan.  Because b = A[M-1, :] (c_{M-1} = 1), the continuity row is
//   dy_{i+1} - u_{M-1} = -f_b, giving the BABD row
//   G = -U_{M-1}[:, :NY], C = I, H = -U_{M-1}[:, NY:NY+NPA] (zero in the other
//   parameter columns), b = -f_b + U_{M-1}[:, NVA-1].  Only the NPA parameters
//   the stage functions depend on get columns (g_p = f_p = 0 for the rest).
template <int CW, class TJ>
__device__ __forceinline__ void condense_one(int i, const int* __restrict__ iinst, const double* __restrict__ t,
                         const TJ* __restrict__ jac, const double* __restrict__ fa, const double* __restrict__ fb,
                         double* __restrict__ rows, TJ* __restrict__ rec, int* __restrict__ status) {
  using Solver = GJB<NS, NS + NVA, CW, 16>;
  constexpr int NT = CW * WAVE;
  // sT (phases 1-2) and the elimination's panel / pivot-row buffers (phase 3)
  // share one LDS area
  constexpr int SMEM = M * NY * NVA > Solver::LDS_DOUBLES ? M * NY * NVA : Solver::LDS_DOUBLES;
  __shared__ double smem[SMEM];
  __shared__ int sflag[Solver::NR];
  double* sT = smem;
  __shared__ int pivrow[NS];
  __shared__ double pivval[NS];
  __shared__ int invperm[Solver::RQ * WAVE];
  __shared__ int sbad;

  const int tid = threadIdx.x;
  const int lane = lane_id(), wave = wave_id();
  if (tid == 0) sbad = 0;
  const int nd = i + iinst[i];
  const double h = t[nd + 1] - t[nd];
  TJ* recU = rec + (size_t)i * REC;   // records stored like the Jacobians
  TJ* recT = recU + NS * NVA;
  TJ* recG = recT + M * NY * NVA;
  int bad = 0;

  // ---- phase 1: wave j eliminates the algebraic unknowns of stage j.
  bad |= stage_eliminate<double, CW, TJ, TJ>(i, jac, fa, sT, recT, recG);
  __syncthreads();

  // ---- phase 2: assemble [S | RHS] into registers
  Solver gj;
#pragma unroll
  for (int q = 0; q < Solver::RQ; ++q) {
    const int r = Solver::row(q);
    const int jr = r / NY + 1;
    const int ar = r - (jr - 1) * NY;
#pragma unroll
    for (int cq = 0; cq < Solver::CQ; ++cq) {
      const int cc = Solver::col(cq);
      double v = 0.0;
      if (r < NS && cc < NS + NVA) {
        if (cc < NS) {
          const int kc = cc / NY + 1;
          const int bc = cc - (kc - 1) * NY;
          v = -h * c_rkA[jr * M + kc] * sT[(kc * NY + ar) * NVA + bc];
          if (jr == kc && ar == bc) v += 1.0;
        } else {
          const int qc = cc - NS;
          v = h * c_rkA[jr * M] * sT[ar * NVA + qc];
          if (qc < NY) {
            if (qc == ar) v += 1.0;
          } else {
            for (int k = 1; k < M; ++k) v += h * c_rkA[jr * M + k] * sT[(k * NY + ar) * NVA + qc];
          }
        }
      }
      gj.a[q][cq] = v;
    }
  }
  __syncthreads();

  // ---- phase 3: Gauss-Jordan
  for (int k = tid; k < Solver::RQ * WAVE; k += NT) {
    invperm[k] = -1;
    sflag[k] = 0;
  }
  if (gj.eliminate(smem, sflag, pivrow, pivval)) bad |= 2;
  __syncthreads();
  for (int k = tid; k < NS; k += NT) invperm[pivrow[k]] = k;
  if (bad) atomicOr(&sbad, bad);
  __syncthreads();

  // ---- phase 4: outputs
  double* rowi = rows + (size_t)i * NY * CRW;
#pragma unroll
  for (int q = 0; q < Solver::RQ; ++q) {
    const int r = Solver::row(q);
    if (r >= NS) continue;
    const int kk = invperm[r];
    if (kk < 0 || kk >= NS) continue;   // only after a failed pivot search (non-finite column)
    const bool last = kk >= (M - 2) * NY;
    const int a = kk - (M - 2) * NY;
#pragma unroll
    for (int cq = 0; cq < Solver::CQ; ++cq) {
      const int cc = Solver::col(cq);
      if (cc < NS || cc >= NS + NVA) continue;
      const int qc = cc - NS;
      const double v = gj.a[q][cq] / pivval[kk];
      recU[kk * NVA + qc] = (TJ)v;
      if (last) {
        if (qc < NY) rowi[a * CRW + qc] = -v;
        else if (qc < NY + NPA) rowi[a * CRW + 2 * NY + pact(qc - NY)] = -v;
        else rowi[a * CRW + CRW - 1] = v - fb[(size_t)i * NY + a];
      }
    }
  }
  // C = I and the inactive H columns: colloc_rows_constant, once per mesh
  if (tid == 0 && sbad) atomicOr(status, sbad);
}

// The intervals of the work queue (queue[0 .. *nq)), the workgroups taking
// every gridDim.x-th: a launch of any size covers them.
template <int CW, class TJ>
__device__ __forceinline__ void condense(const int* __restrict__ iinst, const int* __restrict__ queue,
                         const unsigned long long* __restrict__ nq, const double* __restrict__ t,
                         const TJ* __restrict__ jac, const double* __restrict__ fa,
                         const double* __restrict__ fb, double* __restrict__ rows, TJ* __restrict__ rec,
                         int* __restrict__ status) {
  const int n = (int)*nq;
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    if (k != (int)blockIdx.x) __syncthreads();   // the previous interval's LDS reads are done
    condense_one<CW, TJ>(queue[k], iinst, t, jac, fa, fb, rows, rec, status);
  }
}

// The work queue of the Gauss-Jordan condensation: the series kernels append
// the intervals they leave (queue[counters[0]++] = i; counters[2] counts
// them too, summed over a whole solve where counters[0] is reset per
// launch); colloc_condense walks the queue.
__device__ __forceinline__ void queue_gauss_jordan(int i, int* queue, unsigned long long* counters) {
  const unsigned long long q = atomicAdd(counters, 1ull);
  queue[q] = i;
  atomicAdd(counters + 2, 1ull);
}

// --------------------------------------------------- series condensation
// The stage system S U = R of condense, S = I - h (A^ x I) diag(M_k) with
// A^ = A[1:, 1:] and M_k = T_k[:, :NY], solved by the fixed-point iteration
//   U <- R + h (A^ x I) diag(M_k) U,
// the Neumann series of S^-1 applied to R.  ||h (A^ x I) diag(M_k)|| is O(h):
// on fine meshes a few terms reach FP64 accuracy.  The iteration stops when a
// term changes U by at most 2^-50 of its largest entry (a few units of FP64
// rounding: each term recomputes U from R, so the change does not go to 0);
// an interval that has not converged after NS_MAX_TERMS terms (or produced a
// non-finite value) is queued for condense (queue_gauss_jordan).
// Work item (k, r, h) of the workgroup owns row r of stage block k and the
// columns [h NS_CPT, (h + 1) NS_CPT) of U.  U and the products Y_k = M_k U_k
// pass through one LDS buffer whose rows are padded to NVAP columns, so every
// work item reads whole, aligned column groups.
constexpr int NS_MAX_TERMS = 40;
constexpr int NS_THREADS = 512;
constexpr int NSK = M - 1;                                 // stage blocks
constexpr int NS_CPT = 24;                                 // columns per work item
constexpr int NS_NH = (NVA + NS_CPT - 1) / NS_CPT;
constexpr int NVAP = NS_NH * NS_CPT;                       // padded row of the U buffer
constexpr int NS_ITEMS = NSK * NY * NS_NH;
// the U buffer, which also holds the stage eliminations' scratch before the series
constexpr int NS_BUF = NSK * NY * NVAP > (NS_THREADS / WAVE) * SE_SCR ? NSK * NY * NVAP : (NS_THREADS / WAVE) * SE_SCR;
template <class R>
constexpr bool ns_ok() {
  return NY <= 48 && NS_ITEMS <= NS_THREADS && (M * NY * NVA + NS_BUF) * (int)sizeof(R) <= 150 * 1024;
}
constexpr bool NS_OK = ns_ok<double>();
// Series convergence: a term changes U by at most this fraction of its
// largest entry (a few units of the arithmetic's rounding).
template <class R> __device__ __forceinline__ R ns_tol();
template <> __device__ __forceinline__ double ns_tol<double>() { return 0x1p-50; }
template <> __device__ __forceinline__ float ns_tol<float>() { return 0x1p-21f; }
// The series' stopping rule after a term: dm is the term's largest change,
// xm the largest entry of the new iterate, dprev the previous term's change
// (0 before the second term).  Stop when the change is within the
// tolerance; in float32 also when the observed contraction q = dm / dprev
// is below 1/2 and bounds what the remaining terms would add (at most
// dm q / (1 - q) <= 2 dm q) by a quarter of the tolerance: the iterate is
// then as accurate as the rule asks, one term before its change is.  The
// float64 series keeps the change test alone.
template <class R>
__device__ __forceinline__ bool ns_stop(R dm, R dprev, R xm) {
  if (dm <= ns_tol<R>() * xm) return true;
  if constexpr (sizeof(R) == 4)
    return dprev > (R)0 && dm < (R)0.5 * dprev && (R)8 * dm * dm <= ns_tol<R>() * xm * dprev;
  return false;
}

// R = double: the float64 condensation.  R = float: the FP32 direction (the
// stage eliminations and the series in float, the records and the BABD row
// stored in float64); an interval whose float series does not converge is
// left to the float64 Gauss-Jordan kernel, like a float64 one.
template <class R>
__device__ __forceinline__ void condense_ns(int item, Batch bt, const double* __restrict__ t,
                                            const R* __restrict__ jac,
                                            const double* __restrict__ fa, const double* __restrict__ fb,
                                            double* __restrict__ rows, R* __restrict__ rec,
                                            int* __restrict__ status, int* __restrict__ fallback,
                                            unsigned long long* __restrict__ counters) {
  constexpr int NWV = NS_THREADS / WAVE;
  // one LDS array with buf at its start, so that the U-buffer addresses fit
  // the 16-bit offsets of the LDS instructions
  __shared__ R smem_ns[NS_BUF + M * NY * NVA];
  R* buf = smem_ns;
  R* sT = smem_ns + NS_BUF;
  __shared__ R sred[2 * NWV];
  __shared__ int sbad;
  const int i = listed_interval(bt, item);
  const int tid = threadIdx.x;
  const int lane = lane_id(), wave = wave_id();
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;   // inactive instance: nothing to condense, nothing for Gauss-Jordan
  if (tid == 0) sbad = 0;
  const R h = (R)(t[i + ib + 1] - t[i + ib]);
  R* recU = rec + (size_t)i * REC;
  R* recT = recU + NS * NVA;
  R* recG = recT + M * NY * NVA;
  const int bad = stage_eliminate_all<R, NWV, R, R>(i, jac, fa, sT, recT, recG, buf);
  __syncthreads();

  const bool act = tid < NS_ITEMS;
  const int it = act ? tid : 0;
  const int k = it / (NY * NS_NH);
  const int rem = it - k * NY * NS_NH;
  const int r = rem / NS_NH;
  const int c0 = (rem - r * NS_NH) * NS_CPT;
  const int js = k + 1;
  R wk[NSK];                           // h A[k+1][k'+1]
#pragma unroll
  for (int q = 0; q < NSK; ++q) wk[q] = h * (R)c_rkA[js * M + q + 1];
  // R = h A[k+1][0] T_0 + (c < NY ? I : h sum_j A[k+1][j] T_j), kept in
  // registers; it is also the first iterate
  R rv[NS_CPT], xo[NS_CPT];
  R* mine = buf + (k * NY + r) * NVAP;
#pragma unroll
  for (int cc = 0; cc < NS_CPT; ++cc) {
    const int c = c0 + cc < NVA ? c0 + cc : NVA - 1;
    R v = h * (R)c_rkA[js * M] * sT[r * NVA + c];
    if (c < NY) {
      if (c == r) v += (R)1;
    } else {
#pragma unroll
      for (int j = 1; j < M; ++j) v += h * (R)c_rkA[js * M + j] * sT[(j * NY + r) * NVA + c];
    }
    if (c0 + cc >= NVA) v = (R)0;      // padding columns stay zero
    rv[cc] = v;
    xo[cc] = v;
    if (act) mine[c0 + cc] = v;
  }
  __syncthreads();
  bool converged = false;
  int terms = 0;
  R dprev = (R)0;
  for (int term = 0; term < NS_MAX_TERMS; ++term) {
    ++terms;
    // Y = M_{k+1} U_k on this item's columns.  The row of M is re-read from
    // LDS every term through an opaque offset (not hoisted into registers).
    int moff = (js * NY + r) * NVA;
    asm volatile("" : "+v"(moff));
    R acc[NS_CPT];
#pragma unroll
    for (int cc = 0; cc < NS_CPT; ++cc) acc[cc] = (R)0;
    const R* Uk = buf + k * NY * NVAP + c0;
    // a runtime loop over the rows of U_k, two rows per step: each step's
    // loads and FMAs are scheduled together (a fully unrolled product keeps
    // every load live)
#pragma nounroll
    for (int q = 0; q < NY; ++q) {
      const R mq = sT[moff + q];
      R u[NS_CPT];
#pragma unroll
      for (int cc = 0; cc < NS_CPT; ++cc) u[cc] = Uk[q * NVAP + cc];
#pragma unroll
      for (int cc = 0; cc < NS_CPT; ++cc) acc[cc] = fma(mq, u[cc], acc[cc]);
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < NS_CPT; ++cc)
      if (act) mine[c0 + cc] = acc[cc];
    __syncthreads();
    // U_new = R + h (A^ x I) Y
    R dmax = (R)0, xmax = (R)0;
#pragma unroll
    for (int cc = 0; cc < NS_CPT; ++cc) {
      R v = rv[cc];
#pragma unroll
      for (int q = 0; q < NSK; ++q) v = fma(wk[q], buf[(q * NY + r) * NVAP + c0 + cc], v);
      acc[cc] = v;
      if (act && c0 + cc < NVA) {
        dmax = nmax(fabs(v - xo[cc]), dmax);
        xmax = nmax(fabs(v), xmax);
      }
      xo[cc] = v;
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < NS_CPT; ++cc)
      if (act) mine[c0 + cc] = acc[cc];
    dmax = wave_max(dmax);
    xmax = wave_max(xmax);
    if (lane == 0) {
      sred[wave] = dmax;
      sred[NWV + wave] = xmax;
    }
    __syncthreads();
    R dm = sred[0], xm = sred[NWV];
#pragma unroll
    for (int w = 1; w < NWV; ++w) {
      dm = nmax(sred[w], dm);
      xm = nmax(sred[NWV + w], xm);
    }
    if (ns_stop<R>(dm, dprev, xm)) {
      converged = true;
      break;
    }
    dprev = dm;
    __syncthreads();   // sred is rewritten by the next term
  }
  if (tid == 0) {
    if (!converged) queue_gauss_jordan(i, fallback, counters);
    atomicAdd(counters + 1, (unsigned long long)terms);   // counters[1]: series terms
  }
  if (bad) atomicOr(&sbad, bad);
  __syncthreads();
  if (tid == 0 && sbad) atomicOr(status, sbad);
  if (!converged) return;
  // outputs: U and the BABD row [G | H_active | b] from the last stage block
  double* rowi = rows + (size_t)i * NY * CRW;
  if (act) {
#pragma unroll
    for (int cc = 0; cc < NS_CPT; ++cc) {
      const int c = c0 + cc;
      if (c >= NVA) continue;
      const double v = (double)xo[cc];
      recU[(k * NY + r) * NVA + c] = xo[cc];
      if (k == NSK - 1) {
        if (c < NY) rowi[r * CRW + c] = -v;
        else if (c < NY + NPA) rowi[r * CRW + 2 * NY + pact(c - NY)] = -v;
        else rowi[r * CRW + CRW - 1] = v - fb[(size_t)i * NY + r];
      }
    }
  }
  // C = I and the inactive H columns: colloc_rows_constant, once per mesh
}

// ------------------------------------- series condensation on matrix cores
// The series of condense_ns with its products Y_k = M_{k+1} U_k on
// v_mfma_f32_16x16x4f32 (float32, the FP32 direction).  One wave per stage
// block k (NSK waves per interval); wave k holds, in registers,
//   M_{k+1} as MFMA A operands  (MF_RT row tiles x MF_KS k-steps, 1 value each)
//   R_k, U_k and Y_k in the accumulator layout of MF_RT x MF_CT 16x16 tiles
//   (4 values per lane and tile).
// Accumulator layout: lane l, value v of tile (ri, ci) is row
// 16 ri + 4 v + l / 16, column 16 ci + l % 16 of the block.  With this row
// numbering, value v of row tile ri is exactly the B operand of k-step
// 4 ri + v (lane l: row 4 ks + l / 16, column l % 16), so U_k feeds the
// next product from the registers it is formed in.  The A operand of row
// tile ri holds, in lane l, row 16 ri + 4 (l % 4) + (l % 16) / 4 of M (the
// output row the MFMA writes to value (l % 16) % 4 of lane group
// (l % 16) / 4) and column 4 ks + l / 16.  Rows and columns past NY / NVA
// are zero, so the padded entries of every tile stay zero.
// A term: the products, Y_k through LDS to every wave, then
//   U_k <- R_k + sum_q h A^[k][q] Y_q
// and the workgroup max of the change (two barriers per term).
typedef float mf_f4 __attribute__((ext_vector_type(4)));
constexpr int MF_RT = (NY + 15) / 16;           // row tiles of a stage block
constexpr int MF_CT = (NVA + 15) / 16;          // column tiles
constexpr int MF_KS = (NY + 3) / 4;             // k-steps of a product
constexpr int MF_T = MF_RT * MF_CT;
constexpr int MF_THREADS = NSK * WAVE;
constexpr int MF_LDS0 = M * NY * NVA + NSK * SE_SCR_MF;   // sT and the stage eliminations' scratch
constexpr int MF_LDS = (MF_LDS0 > NSK * MF_T * WAVE * 4) ? MF_LDS0 : NSK * MF_T * WAVE * 4;
constexpr bool MF_OK = NSK >= 1 && NSK <= 16 && MF_T * 12 + MF_RT * MF_KS <= 176 && MF_LDS * 4 <= 64 * 1024;
// Three workgroups fit a CU's LDS; three waves per SIMD leave 168 registers
// a lane.  Where the tiles' registers (12 per accumulator tile, one per
// A operand) leave room for addressing below that, the kernel is held to it
// (19.5 -> 16.8 ms for the n_x = 20 family's 298 663-interval launch);
// otherwise the compiler's choice (two waves per SIMD) stands.
constexpr int MF_WAVES = MF_T * 12 + MF_RT * MF_KS <= 140 ? 3 : 1;

__device__ __forceinline__ void condense_mf32(int item, Batch bt, const double* __restrict__ t,
                                              const float* __restrict__ jac, const double* __restrict__ fa,
                                              const double* __restrict__ fb, double* __restrict__ rows,
                                              float* __restrict__ rec, int* __restrict__ status,
                                              int* __restrict__ fallback, unsigned long long* __restrict__ counters) {
  // sT (stage eliminations) and, once M and R are in registers, the
  // exchange buffer of the products
  __shared__ __attribute__((aligned(16))) float smem_mf[MF_LDS];
  __shared__ float sred[2 * NSK];
  __shared__ int sbad;
  float* sT = smem_mf;
  mf_f4* ybuf = reinterpret_cast<mf_f4*>(smem_mf);
  const int i = listed_interval(bt, item);
  const int tid = threadIdx.x;
  const int lane = lane_id(), k = wave_id();
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  if (tid == 0) sbad = 0;
  const float h = (float)(t[i + ib + 1] - t[i + ib]);
  float* recU = rec + (size_t)i * REC;
  float* recT = recU + NS * NVA;
  float* recG = recT + M * NY * NVA;
  const int bad = stage_eliminate_mf<NSK, float>(i, jac, fa, sT, recT, recG, sT + M * NY * NVA);
  __syncthreads();

  const int js = k + 1;
  const int g = lane >> 4, c16 = lane & 15;
  // M_{k+1} as A operands
  float am[MF_RT][MF_KS];
  {
    const int rl = 4 * (c16 & 3) + (c16 >> 2);
#pragma unroll
    for (int ri = 0; ri < MF_RT; ++ri)
#pragma unroll
      for (int ks = 0; ks < MF_KS; ++ks) {
        const int row = 16 * ri + rl, col = 4 * ks + g;
        am[ri][ks] = (row < NY && col < NY) ? sT[(js * NY + row) * NVA + col] : 0.0f;
      }
  }
  float wk[NSK];   // h A^[k][q]
#pragma unroll
  for (int q = 0; q < NSK; ++q) wk[q] = h * (float)c_rkA[js * M + q + 1];
  // R_k = h A[k+1][0] T_0 + (c < NY ? I : h sum_{j>=1} A[k+1][j] T_j); the first iterate
  mf_f4 rv[MF_T], u[MF_T];
#pragma unroll
  for (int ri = 0; ri < MF_RT; ++ri)
#pragma unroll
    for (int ci = 0; ci < MF_CT; ++ci)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * ri + 4 * v + g, col = 16 * ci + c16;
        float val = 0.0f;
        if (row < NY && col < NVA) {
          val = h * (float)c_rkA[js * M] * sT[row * NVA + col];
          if (col < NY) {
            if (col == row) val += 1.0f;
          } else {
#pragma unroll
            for (int j = 1; j < M; ++j) val += h * (float)c_rkA[js * M + j] * sT[(j * NY + row) * NVA + col];
          }
        }
        rv[ri * MF_CT + ci][v] = val;
        u[ri * MF_CT + ci][v] = val;
      }
  __syncthreads();   // sT is overwritten by the exchange buffer from here on

  bool converged = false;
  int terms = 0;
  float dprev = 0.0f;
  for (int term = 0; term < NS_MAX_TERMS; ++term) {
    ++terms;
    // one row of output tiles at a time (MF_CT independent accumulation
    // chains), each stored to the exchange buffer when complete, so only
    // one row's accumulators are live
#pragma unroll
    for (int ri = 0; ri < MF_RT; ++ri) {
      mf_f4 y[MF_CT];
#pragma unroll
      for (int ci = 0; ci < MF_CT; ++ci) y[ci] = mf_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < MF_KS; ++ks)
#pragma unroll
        for (int ci = 0; ci < MF_CT; ++ci)
          y[ci] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[ri][ks], u[(ks >> 2) * MF_CT + ci][ks & 3], y[ci], 0, 0, 0);
#pragma unroll
      for (int ci = 0; ci < MF_CT; ++ci) ybuf[(k * MF_T + ri * MF_CT + ci) * WAVE + lane] = y[ci];
    }
    __syncthreads();
    // U_k <- R_k + sum_q h A^[k][q] Y_q (padded entries stay zero)
    float dmax = 0.0f, xmax = 0.0f;
#pragma unroll
    for (int tt = 0; tt < MF_T; ++tt) {
      mf_f4 acc = rv[tt];
#pragma unroll
      for (int q = 0; q < NSK; ++q) {
        const mf_f4 yq = ybuf[(q * MF_T + tt) * WAVE + lane];
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = fmaf(wk[q], yq[v], acc[v]);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        dmax = nmax(fabsf(acc[v] - u[tt][v]), dmax);
        xmax = nmax(fabsf(acc[v]), xmax);
      }
      u[tt] = acc;
    }
    dmax = wave_max(dmax);
    xmax = wave_max(xmax);
    if (lane == 0) {
      sred[k] = dmax;
      sred[NSK + k] = xmax;
    }
    __syncthreads();   // also: every wave has read ybuf before the next term writes it
    float dm = sred[0], xm = sred[NSK];
#pragma unroll
    for (int w = 1; w < NSK; ++w) {
      dm = nmax(sred[w], dm);
      xm = nmax(sred[NSK + w], xm);
    }
    if (ns_stop<float>(dm, dprev, xm)) {
      converged = true;
      break;
    }
    dprev = dm;
  }
  if (tid == 0) {
    if (!converged) queue_gauss_jordan(i, fallback, counters);
    atomicAdd(counters + 1, (unsigned long long)terms);
  }
  if (bad) atomicOr(&sbad, bad);
  __syncthreads();
  if (tid == 0 && sbad) atomicOr(status, sbad);
  if (!converged) return;
  // outputs: U_k and the BABD row [G | H_active | b] from the last stage block
  double* rowi = rows + (size_t)i * NY * CRW;
  const bool last = k == NSK - 1;
#pragma unroll
  for (int ri = 0; ri < MF_RT; ++ri)
#pragma unroll
    for (int ci = 0; ci < MF_CT; ++ci)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * ri + 4 * v + g, col = 16 * ci + c16;
        if (row >= NY || col >= NVA) continue;
        const float x = u[ri * MF_CT + ci][v];
        recU[(k * NY + row) * NVA + col] = x;
        if (last) {
          const double dv = (double)x;
          if (col < NY) rowi[row * CRW + col] = -dv;
          else if (col < NY + NPA) rowi[row * CRW + 2 * NY + pact(col - NY)] = -dv;
          else rowi[row * CRW + CRW - 1] = dv - fb[(size_t)i * NY + row];
        }
      }
  // C = I and the inactive H columns: colloc_rows_constant, once per mesh
}

// ------------------------------------------------------ cyclic reduction
// One workgroup per pair of adjacent BABD rows (level l):
//   row a: G_a x_L + C_a x_M + H_a p = b_a,   row b: G_b x_M + C_b x_R + H_b p = b_b.
// Householder QR of [C_a; G_b] (2NY x NY) eliminates x_M (the same 2-block
// step as partition_factorization_kernel, solve_babd_system.py:184-246):
// the top NY rows [R | Gp | Ep | Jp | d] are the record used to recover x_M,
// the bottom NY rows are the combined row [G | C | H | b] of level l+1.
// Each column of the 2NY x RECW block is held by CR_PARTS lanes of one wave
// (CR_ROWS rows each), one lane in each 16-lane row of the wave, so a pair
// has many waves and short per-lane chains; column sums over the parts and
// the pivot broadcast are permlane16/32 swaps.  The Householder vector
// is broadcast through LDS, where rows above the current pivot are kept zero
// so every loop has a fixed trip count.  An odd last row is carried.
// Workgroup g makes output row g of the level: oinst[g] is its instance,
// seg[oinst[g]] the instance's place in this level's buffers.  TI / TO: the
// stored type of the input rows / of the output rows and records (float for
// the FP32 cyclic reduction; the arithmetic is float64 either way).
template <class TI, class TO>
__device__ __forceinline__ void cr_pair(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                        const int* __restrict__ act, const TI* __restrict__ rows_in,
                                        TO* __restrict__ rows_out, TO* __restrict__ recs) {
  constexpr int R2 = 2 * NY;
  __shared__ double sv[2][CR_PARTS * CR_ROWS];
  __shared__ double sbeta[2];
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
  if (k == npairs) {  // carry the odd last row
    const TI* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    TO* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = tid; e < NY * CRW; e += blockDim.x) dst[e] = (TO)src[e];
    return;
  }
  for (int e = tid; e < 2 * CR_PARTS * CR_ROWS; e += blockDim.x) (&sv[0][0])[e] = 0.0;
  const int lane = lane_id();
  const int part = lane / CR_COLS;                 // which CR_ROWS-row slice
  const int cidx = wave_id() * CR_COLS + (lane % CR_COLS);
  const bool live = cidx < RECW_L;
  // the carried column's place in the record ([R | Gp | Ep | Jp | d]; Jp
  // absent when HZ) and in the output row (column - NY)
  const int pcol = cidx < 3 * NY + NPH ? cidx : RECW - 1;
  const TI* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const TI* rb = ra + NY * CRW;
  // column layout: [x_M (NY) | x_L (NY) | x_R (NY) | p (NP) | rhs]
  double a[CR_ROWS];
#pragma unroll
  for (int i = 0; i < CR_ROWS; ++i) {
    const int r = part * CR_ROWS + i;
    double v = 0.0;
    if (live && r < R2) {
      const bool top = r < NY;
      const TI* src = top ? ra + r * CRW : rb + (r - NY) * CRW;
      if (cidx < NY) v = top ? src[NY + cidx] : src[cidx];
      else if (cidx < 2 * NY) v = top ? src[cidx - NY] : 0.0;
      else if (cidx < 3 * NY) v = top ? 0.0 : src[cidx - NY];
      else v = src[pcol - NY];
    }
    a[i] = v;
  }
  // partial sum of squares over this lane's rows at or below the pivot row
  double tail = 0.0;
#pragma unroll
  for (int i = 0; i < CR_ROWS; ++i) tail = fma(a[i], a[i], tail);
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  bool stopped = false;
  static_for<0, NY>([&](auto kc) {
    constexpr int kk = decltype(kc)::value;
    constexpr int buf = kk & 1;
    constexpr int pk = kk / CR_ROWS, ik = kk % CR_ROWS;
    // column norm below the pivot and the pivot entry, for every column
    const double tfull = sum_rows(tail);
    const double akk = bcast_row<pk>(a[ik]);
    if (cidx == kk) {
      const double nrm = sqrt(tfull);
      const double g = akk >= 0.0 ? -nrm : nrm;
      const double v0 = akk - g;
      const double vs = tfull - akk * akk + v0 * v0;
      const bool reflect = !(vs < eps);
#pragma unroll
      for (int i = 0; i < CR_ROWS; ++i) {
        const int r = part * CR_ROWS + i;
        double v = 0.0;
        if (r == kk) v = v0;
        else if (r > kk && r < R2) v = a[i];
        sv[buf][r] = v;                      // rows < kk (stale from step kk-2) become zero
        if (reflect) {
          if (r == kk) a[i] = g;
          else if (r > kk) a[i] = 0.0;
        }
      }
      if (part == 0) sbeta[buf] = reflect ? -2.0 * fast_rcp(vs) : 0.0;
    }
    __syncthreads();
    const double beta = sbeta[buf];
    if (beta == 0.0) stopped = true;   // the reference stops reflecting at |v|^2 < eps
    // dot products and squared tails in four independent chains (FP64 FMA
    // latency is long)
    double w = 0.0;
    if (!stopped && cidx > kk && live) {
      double wc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < CR_ROWS; ++i) wc[i & 3] = fma(sv[buf][part * CR_ROWS + i], a[i], wc[i & 3]);
      w = (wc[0] + wc[1]) + (wc[2] + wc[3]);
    }
    w = sum_rows(w);
    if (!stopped && cidx > kk && live) {
      w *= beta;
#pragma unroll
      for (int i = 0; i < CR_ROWS; ++i) a[i] = fma(w, sv[buf][part * CR_ROWS + i], a[i]);
    }
    {
      double tc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < CR_ROWS; ++i) {
        const int r = part * CR_ROWS + i;
        tc[i & 3] = (r > kk) ? fma(a[i], a[i], tc[i & 3]) : tc[i & 3];
      }
      tail = (tc[0] + tc[1]) + (tc[2] + tc[3]);
    }
  });
  if (live) {
    TO* rec = recs + (size_t)k * NY * RECW;
    TO* out = rows_out + (size_t)k * NY * CRW;
#pragma unroll
    for (int i = 0; i < CR_ROWS; ++i) {
      const int r = part * CR_ROWS + i;
      if (r < NY) rec[r * RECW + pcol] = (TO)a[i];
      else if (r < R2 && cidx >= NY) out[(r - NY) * CRW + pcol - NY] = (TO)a[i];
    }
  }
}

// Pairing by substitution, for levels whose rows all have C = I (the
// condensation gives C_i = I, and this step keeps it): row a
// G_a x_L + x_M + H_a p = b_a gives x_M = b_a - G_a x_L - H_a p, and row b
// G_b x_M + C_b x_R + H_b p = b_b becomes
//   (-G_b G_a) x_L + C_b x_R + (H_b - G_b H_a) p = b_b - G_b b_a.
// The record is [R = I | Gp = G_a | Ep = 0 | Jp = H_a | d = b_a], which
// cr_back solves like a Householder record.  One product G_b [G_a | H_a | b_a]
// (NY x NY times NY x (NY+NP+1)) per pair instead of NY dependent reflector
// steps.  Used only on the finest levels (a pair spans a small fraction of
// the horizon, so the transition products stay near the identity); the
// coarse levels keep the Householder step.
// Every thread issues all its loads of G_b, [G_a | H_a | b_a] and [H_b | b_b]
// before the first store to LDS (fixed trip counts); thread (rg, cg) then computes a
// CS_TR x CS_TC tile of the product from LDS, the tiles go back to LDS,
// and the output row is written row by row, with consecutive threads on
// consecutive columns.
constexpr int CS_XC = NY + NPH + 1;                   // columns of [G_a | H_a | b_a] (no H_a when HZ)
constexpr int CS_TR = 2, CS_TC = 4;                   // a thread's product tile
constexpr int CS_RG = (NY + CS_TR - 1) / CS_TR;       // row groups
constexpr int CS_CG = (CS_XC + CS_TC - 1) / CS_TC;    // column groups
constexpr int CS_THREADS = 256;
constexpr int CS_GR = CS_RG * CS_TR;                  // padded rows of G_b
constexpr int CS_XP = CS_CG * CS_TC;                  // padded columns of [G_a | H_a | b_a]
constexpr int CS_LG = (NY * NY + CS_THREADS - 1) / CS_THREADS;      // loads per thread: G_b
constexpr int CS_LX = (NY * CS_XC + CS_THREADS - 1) / CS_THREADS;   // [G_a | H_a | b_a]
constexpr int CS_RC = CS_XC - NY;                                    // columns of [H_b | b_b]
constexpr int CS_LB = (NY * CS_RC + CS_THREADS - 1) / CS_THREADS;   // [H_b | b_b]
// TI / TO: stored type of the input / output rows; the product is computed
// in TO (float32 arithmetic for the FP32 cyclic reduction).
template <class TI, class TO>
__device__ __forceinline__ void cr_subst(const int* __restrict__ oinst, const CRSeg* __restrict__ seg,
                                         const int* __restrict__ act, const TI* __restrict__ rows_in,
                                         TO* __restrict__ rows_out, TO* __restrict__ recs) {
  __shared__ __attribute__((aligned(16))) TO sGT[NY][CS_GR];   // G_b transposed (padded rows zero)
  __shared__ __attribute__((aligned(16))) TO sX[NY][CS_XP];    // [G_a | H_a | b_a], then the product
  __shared__ TO sR[NY][CS_RC];                                 // [H_b | b_b]
  const int ib = oinst[blockIdx.x];
  if (act && !act[ib]) return;
  const CRSeg sg = seg[ib];
  if (!sg.subst) return;
  const int k = blockIdx.x - sg.out_off;
  const int nrows = sg.nrows;
  rows_in += (size_t)sg.in_off * NY * CRW;
  rows_out += (size_t)sg.out_off * NY * CRW;
  (void)recs;   // the substitution leaves no record
  const int npairs = nrows >> 1;
  const int tid = threadIdx.x;
  // C = I on every row of a substitution level, so it is not read here;
  // it is written (as I) only where the next level is not one (subst == 1)
  const bool wc = sg.subst == 1;
  if (k == npairs) {  // carry the odd last row
    const TI* src = rows_in + (size_t)(nrows - 1) * NY * CRW;
    TO* dst = rows_out + (size_t)npairs * NY * CRW;
    for (int e = tid; e < NY * CRW; e += CS_THREADS) {
      const int r = e / CRW, c = e - (e / CRW) * CRW;
      if (c < NY || c == CRW - 1 || (!HZ && c >= 2 * NY)) dst[e] = (TO)src[e];
      else if (wc && c < 2 * NY) dst[e] = (c - NY == r) ? (TO)1 : (TO)0;
    }
    return;
  }
  const TI* ra = rows_in + (size_t)(2 * k) * NY * CRW;
  const TI* rb = ra + NY * CRW;
  TO* out = rows_out + (size_t)k * NY * CRW;
  // column c of [G_a | H_a | b_a] in a row: G_a at c, H_a after C_a, b_a last
  auto xcol = [](int c) { return c < NY ? c : (c < NY + NPH ? NY + c : CRW - 1); };
  TI vg[CS_LG], vx[CS_LX], vb[CS_LB];
#pragma unroll
  for (int q = 0; q < CS_LG; ++q) {
    const int e = tid + q * CS_THREADS;
    const int r = e / NY, c = e - (e / NY) * NY;
    vg[q] = e < NY * NY ? rb[r * CRW + c] : (TI)0;
  }
#pragma unroll
  for (int q = 0; q < CS_LX; ++q) {
    const int e = tid + q * CS_THREADS;
    const int r = e / CS_XC, c = e - (e / CS_XC) * CS_XC;
    vx[q] = e < NY * CS_XC ? ra[r * CRW + xcol(c)] : (TI)0;
  }
#pragma unroll
  for (int q = 0; q < CS_LB; ++q) {
    const int e = tid + q * CS_THREADS;
    const int r = e / CS_RC, c = NY + e - (e / CS_RC) * CS_RC;
    vb[q] = e < NY * CS_RC ? rb[r * CRW + xcol(c)] : (TI)0;
  }
#pragma unroll
  for (int q = 0; q < CS_LG; ++q) {
    const int e = tid + q * CS_THREADS;
    const int r = e / NY, c = e - (e / NY) * NY;
    if (e < NY * NY) sGT[c][r] = (TO)vg[q];
  }
#pragma unroll
  for (int q = 0; q < CS_LX; ++q) {
    const int e = tid + q * CS_THREADS;
    const int r = e / CS_XC, c = e - (e / CS_XC) * CS_XC;
    if (e < NY * CS_XC) sX[r][c] = (TO)vx[q];
  }
#pragma unroll
  for (int q = 0; q < CS_LB; ++q) {
    const int e = tid + q * CS_THREADS;
    if (e < NY * CS_RC) sR[e / CS_RC][e - (e / CS_RC) * CS_RC] = (TO)vb[q];
  }
  // the padding: rows NY.. of G_b, columns CS_XC.. of [G_a | H_a | b_a]
  if constexpr (CS_GR > NY)
    for (int e = tid; e < NY * (CS_GR - NY); e += CS_THREADS) sGT[e / (CS_GR - NY)][NY + e % (CS_GR - NY)] = (TO)0;
  if constexpr (CS_XP > CS_XC)
    for (int e = tid; e < NY * (CS_XP - CS_XC); e += CS_THREADS)
      sX[e / (CS_XP - CS_XC)][CS_XC + e % (CS_XP - CS_XC)] = (TO)0;
  // no record: back substitution reads row a itself (x_M = b_a - G_a x_L - H_a p),
  // which stays in place in this level's input rows
  __syncthreads();
  constexpr int NT = CS_RG * CS_CG;
  constexpr int NTL = (NT + CS_THREADS - 1) / CS_THREADS;
  TO acc[NTL][CS_TR][CS_TC];
#pragma unroll
  for (int u = 0; u < NTL; ++u) {
    const int t = tid + u * CS_THREADS;
    const int rg = t < NT ? t / CS_CG : 0, cg = t < NT ? t - (t / CS_CG) * CS_CG : 0;
#pragma unroll
    for (int a = 0; a < CS_TR; ++a)
#pragma unroll
      for (int c = 0; c < CS_TC; ++c) acc[u][a][c] = 0;
#pragma unroll 4
    for (int q = 0; q < NY; ++q) {
      TO g[CS_TR], x[CS_TC];
#pragma unroll
      for (int a = 0; a < CS_TR; ++a) g[a] = sGT[q][rg * CS_TR + a];
#pragma unroll
      for (int c = 0; c < CS_TC; ++c) x[c] = sX[q][cg * CS_TC + c];
#pragma unroll
      for (int a = 0; a < CS_TR; ++a)
#pragma unroll
        for (int c = 0; c < CS_TC; ++c) acc[u][a][c] = fma(g[a], x[c], acc[u][a][c]);
    }
  }
  __syncthreads();   // every thread has read sX
#pragma unroll
  for (int u = 0; u < NTL; ++u) {
    const int t = tid + u * CS_THREADS;
    if (t >= NT) continue;
    const int rg = t / CS_CG, cg = t - (t / CS_CG) * CS_CG;
#pragma unroll
    for (int a = 0; a < CS_TR; ++a) {
      const int r = rg * CS_TR + a;
      if (r < NY) {
#pragma unroll
        for (int c = 0; c < CS_TC; ++c) sX[r][cg * CS_TC + c] = acc[u][a][c];
      }
    }
  }
  __syncthreads();
  // [G' | H' | b'] = [-P_G | H_b - P_H | b_b - P_b], P = G_b [G_a | H_a | b_a]
  for (int e = tid; e < NY * CS_XC; e += CS_THREADS) {
    const int r = e / CS_XC, c = e - (e / CS_XC) * CS_XC;
    const TO v = sX[r][c];
    if (c < NY) {
      out[r * CRW + c] = -v;                                   // G'
    } else {
      out[r * CRW + xcol(c)] = sR[r][c - NY] - v;              // H', b'
    }
  }
  if (wc) {
    for (int e = tid; e < NY * NY; e += CS_THREADS) {      // C' = C_b = I
      const int r = e / NY, c = e - (e / NY) * NY;
      out[r * CRW + NY + c] = (r == c) ? (TO)1 : (TO)0;
    }
  }
}

// Boundary system on the last remaining row G x_0 + C x_N + H p = b and the
// boundary conditions B_0 x_0 + B_n x_N + V_n p = -r_bc: unknowns
// [x_N | x_0 | p] as in qr_decomposition's final block (solve_babd_system.py:364-381).
// Workgroup b solves instance b, whose last row is row b of `rows`.
template <int FW, class TRW>
__device__ __forceinline__ void cr_final(Batch bt, const TRW* __restrict__ rows, const double* __restrict__ Dr,
                         const double* __restrict__ rbc, double* __restrict__ dy, double* __restrict__ dp,
                         int* __restrict__ status) {
  using Solver = GJ<NFIN, NFIN + 1, FW>;
  __shared__ double colbuf[Solver::COLBUF];
  __shared__ int pivrow[NFIN];
  __shared__ double pivval[NFIN];
  __shared__ int invperm[Solver::RQ * WAVE];
  const int ib = blockIdx.x;
  if (skipped(bt, ib)) return;
  const TRW* row = rows + (size_t)ib * NY * CRW;
  Dr += (size_t)ib * DRSZ;
  rbc += (size_t)ib * NRB;
  double* x0 = dy + (size_t)bt.noff[ib] * NY;
  double* xn = dy + (size_t)(bt.noff[ib + 1] - 1) * NY;
  dp += (size_t)ib * NP;
  Solver gj;
  constexpr int DRW = 2 * NY + NP;
  for (int k = threadIdx.x; k < Solver::RQ * WAVE; k += blockDim.x) invperm[k] = -1;
#pragma unroll
  for (int q = 0; q < Solver::RQ; ++q) {
    const int r = Solver::row(q);
#pragma unroll
    for (int cq = 0; cq < Solver::CQ; ++cq) {
      const int cc = Solver::col(cq);
      double v = 0.0;
      if (r < NY) {
        if (cc < NY) v = (double)row[r * CRW + NY + cc];                 // C
        else if (cc < 2 * NY) v = (double)row[r * CRW + cc - NY];        // G
        else if (cc < NFIN) v = HZ ? 0.0 : (double)row[r * CRW + cc];    // H
        else if (cc == NFIN) v = (double)row[r * CRW + CRW - 1];         // b
      } else if (r < NFIN) {
        const int rr = r - NY;
        if (cc < NY) v = Dr[rr * DRW + NY + cc];                 // B_n
        else if (cc < 2 * NY) v = Dr[rr * DRW + cc - NY];        // B_0
        else if (cc < NFIN) v = Dr[rr * DRW + cc];               // V_n
        else if (cc == NFIN) v = -rbc[rr];
      }
      gj.a[q][cq] = v;
    }
  }
  const int bad = gj.eliminate(colbuf, pivrow, pivval);
  __syncthreads();
  for (int k = threadIdx.x; k < NFIN; k += blockDim.x) invperm[pivrow[k]] = k;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Solver::RQ; ++q) {
    const int r = Solver::row(q);
#pragma unroll
    for (int cq = 0; cq < Solver::CQ; ++cq) {
      if (Solver::col(cq) != NFIN || r >= NFIN) continue;
      const int kk = invperm[r];
      if (kk < 0 || kk >= NFIN) continue;
      const double v = gj.a[q][cq] / pivval[kk];
      if (kk < NY) xn[kk] = v;
      else if (kk < 2 * NY) x0[kk - NY] = v;
      else dp[kk - 2 * NY] = v;
    }
  }
  if (bad && threadIdx.x == 0) atomicOr(status, 4);
}

// Back substitution of one cyclic-reduction level: one wave per record,
//   x_M = R^-1 (d - Gp x_L - Ep x_R - Jp p)   (as partition_backward_substitution_kernel,
// solve_babd_system.py:537-572, with the triangular solve done directly).
// Launched with the level's output-row grid of cr_pair; carried rows have
// no record and return.  TC / TRW: stored type of the records / of the
// level's input rows (arithmetic in float64).
template <class TC, class TRW>
__device__ __forceinline__ void cr_back(int level, Batch bt, const int* __restrict__ oinst,
                        const CRSeg* __restrict__ seg, const TC* __restrict__ recs,
                        const TRW* __restrict__ rows_in, double* __restrict__ dy,
                        const double* __restrict__ dp) {
  const int ib = oinst[blockIdx.x];
  if (skipped(bt, ib)) return;
  const CRSeg sg = seg[ib];
  const int k = blockIdx.x - sg.out_off;
  if (k >= (sg.nrows >> 1)) return;
  const int lane = threadIdx.x;
  const int base = bt.noff[ib];
  const int nint = bt.noff[ib + 1] - base - 1;
  const int mid = base + ((2 * k + 1) << level);
  const int left = base + ((2 * k) << level);
  const int right = base + min((2 * k + 2) << level, nint);
  const double* xl = dy + (size_t)left * NY;
  const double* xr = dy + (size_t)right * NY;
  dp += (size_t)ib * NP;
  if (sg.subst) {
    // substitution level: row a of the pair, [G_a | I | H_a | b_a] (this
    // level's input, kept in place), gives x_M = b_a - G_a x_L - H_a p
    const TRW* ra = rows_in + (size_t)(sg.in_off + 2 * k) * NY * CRW + (lane < NY ? lane : 0) * CRW;
    double rhs = (double)ra[CRW - 1];
#pragma unroll 8
    for (int c = 0; c < NY; ++c) rhs = fma(-(double)ra[c], xl[c], rhs);
#pragma unroll 8
    for (int c = 0; c < NPH; ++c) rhs = fma(-(double)ra[2 * NY + c], dp[c], rhs);
    if (lane < NY) dy[(size_t)mid * NY + lane] = rhs;
    return;
  }
  const TC* rec = recs + (size_t)(sg.rec_off + k) * NY * RECW;
  // lane r holds row r of R and the right-hand side d - Gp x_L - Ep x_R - Jp p
  const TC* rw = rec + (lane < NY ? lane : 0) * RECW;
  double Rr[NY];
#pragma unroll
  for (int c = 0; c < NY; ++c) Rr[c] = (double)rw[c];
  double rhs = (double)rw[RECW - 1];
#pragma unroll 8
  for (int c = 0; c < NY; ++c) rhs = fma(-(double)rw[NY + c], xl[c], rhs);
#pragma unroll 8
  for (int c = 0; c < NY; ++c) rhs = fma(-(double)rw[2 * NY + c], xr[c], rhs);
#pragma unroll 8
  for (int c = 0; c < NPH; ++c) rhs = fma(-(double)rw[3 * NY + c], dp[c], rhs);
  double x = 0.0;
#pragma unroll
  for (int r = NY - 1; r >= 0; --r) {
    const double xv = readlane_d(rhs / Rr[r], r);
    if (lane == r) x = xv;
    if (lane < r) rhs = fma(-Rr[r], xv, rhs);
  }
  if (lane < NY) dy[(size_t)mid * NY + lane] = x;
}

#endif  // !COLLOC_LARGE

// ---------------------------------------------------------------- recover
// One workgroup per interval (recover_delta_k_kernel, solver:1619-1650):
//   u_0 = dy_i, u_j = U_j v,  dydot_j = T_j [u_j; dp; 1],  dz_j = -G_j [u_j; dp; 1].
template <class TR>
__device__ __forceinline__ void recover(int nint, Batch bt, const TR* __restrict__ rec,
                        const double* __restrict__ dy, const double* __restrict__ dp,
                        double* __restrict__ dydot, double* __restrict__ dz) {
  __shared__ double sw[M][NVA];
  const int i = listed_interval(bt, blockIdx.x);
  const int ib = bt.iinst[i];
  if (skipped(bt, ib)) return;
  dy += (size_t)ib * NY;    // left node of interval i is i + ib
  dp += (size_t)ib * NP;
  const int tid = threadIdx.x;
  const TR* recU = rec + (size_t)i * REC;
  const TR* recT = recU + NS * NVA;
  const TR* recG = recT + M * NY * NVA;
  for (int e = tid; e < M * NVA; e += blockDim.x) {
    const int j = e / NVA, q = e - (e / NVA) * NVA;
    double v = 0.0;
    if (q < NY) v = (j == 0) ? dy[(size_t)i * NY + q] : 0.0;
    else if (q < NY + NPA) v = dp[pact(q - NY)];
    else v = 1.0;
    sw[j][q] = v;
  }
  __syncthreads();
  // u_j rows: one thread per record row (a row's NVA entries are read in
  // sequence; the wave's rows share cache lines across the iterations), the
  // vector [dy_i; dp; 1] is an LDS broadcast
  for (int rr = tid; rr < NS; rr += blockDim.x) {
    const TR* rowp = recU + (size_t)rr * NVA;
    double s0 = 0.0, s1 = 0.0;
    int q = 0;
    for (; q + 1 < NVA; q += 2) {
      s0 = fma((double)rowp[q], sw[0][q], s0);
      s1 = fma((double)rowp[q + 1], sw[0][q + 1], s1);
    }
    if (q < NVA) s0 = fma((double)rowp[q], sw[0][q], s0);
    sw[rr / NY + 1][rr % NY] = s0 + s1;
  }
  __syncthreads();
  for (int rr = tid; rr < M * (NY + NZ); rr += blockDim.x) {
    const bool isz = rr >= M * NY;
    const int j = isz ? (rr - M * NY) / NZ1 : rr / NY;
    const TR* rowp = isz ? recG + (size_t)(rr - M * NY) * NVA : recT + (size_t)rr * NVA;
    double s0 = 0.0, s1 = 0.0;
    int q = 0;
    for (; q + 1 < NVA; q += 2) {
      s0 = fma((double)rowp[q], sw[j][q], s0);
      s1 = fma((double)rowp[q + 1], sw[j][q + 1], s1);
    }
    if (q < NVA) s0 = fma((double)rowp[q], sw[j][q], s0);
    const double s = s0 + s1;
    if (isz) {
      const int a = (rr - M * NY) - j * NZ;
      dz[((size_t)i * M + j) * NZ + a] = -s;
    } else {
      const int a = rr - j * NY;
      dydot[((size_t)i * M + j) * NY + a] = s;
    }
  }
}

// ------------------------------------------------------------------- axpy
// out = x + a_b d over the packed state [y | ydot | ztil | p], with a_b =
// steps[b] the step of the element's instance.  Elements of instances with
// a_b = 0 are not touched: the line search reads a trial state only for the
// instances it is still searching for, and the accepted step leaves the
// other instances' state as it is (so a non-finite direction of an instance
// that takes no step cannot leak).
struct Seg4 { long long off[4]; long long len[4]; };
__device__ __forceinline__ void axpy4(Seg4 s, Batch bt, const double* x, const double* __restrict__ d,
                                      const double* __restrict__ steps, double* out) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (gid < s.len[q]) {
      int b;
      if (q == 0) b = bt.ninst[gid / NY];
      else if (q == 1) b = bt.iinst[gid / (M * NY)];
      else if (q == 2) b = bt.iinst[gid / (M * NZ1)];
      else b = (int)(gid / NP1);
      const double a = steps[b];
      const long long e = s.off[q] + gid;
      if (a != 0.0) out[e] = fma(a, d[e], x[e]);
      return;
    }
    gid -= s.len[q];
  }
}

// The same update over the intervals of a launch list (bt.ilist set): one
// workgroup per listed interval i of instance b updates its stage blocks
// (y-dot, z-tilde) and its left node; the instance's last interval also its
// right node and p.  Instances with steps[b] == 0 are left unchanged.
__device__ __forceinline__ void axpy_listed(Seg4 s, Batch bt, const double* x, const double* __restrict__ d,
                                            const double* __restrict__ steps, double* out) {
  if ((int)blockIdx.x >= (bt.nlistp ? *bt.nlistp : bt.nlist)) return;
  const int i = bt.ilist[blockIdx.x];
  const int b = bt.iinst[i];
  const double a = steps[b];
  if (a == 0.0) return;
  const bool last = i + b + 2 == bt.noff[b + 1];
  constexpr int NYD = M * NY, NZD = M * NZ;
  const int n = NYD + NZD + NY + (last ? NY + NP : 0);
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    long long k;
    if (e < NYD) k = s.off[1] + (long long)i * NYD + e;
    else if (e < NYD + NZD) k = s.off[2] + (long long)i * NZD + (e - NYD);
    else if (e < NYD + NZD + 2 * NY) k = s.off[0] + (long long)(i + b) * NY + (e - NYD - NZD);   // left, right node
    else k = s.off[3] + (long long)b * NP + (e - NYD - NZD - 2 * NY);
    out[k] = fma(a, d[k], x[k]);
  }
}

// -------------------------------------------------------- linear residual
// rho = F + J delta of the Newton system J delta = -F for a direction delta
// = (dy, dydot, dz, dp), in float64, with the stored stage Jacobians (float32
// for the FP32 direction) and Dr: the linearisation of compute_f_q
// (solver:478-559) and of r at the current state,
//   stage j of interval i:  rho_a = f_a + [Df; Dg] [dyt_j; dz_j; dp] - [dydot_j; 0],
//                           dyt_j = dy_i + h sum_k A_jk dydot_k
//   continuity:             rho_b = f_b + dy_{i+1} - dy_i - h sum_k b_k dydot_k
//   boundary:               rho_r = r_bc + Dr [dy_0; dy_N; dp].
// One thread per row: the listed intervals' M (NY+NZ) stage rows, then their
// NY continuity rows, then the B instances' NY+NP boundary rows.  WRITE:
// rho replaces f_a, f_b, r_bc (the right-hand side of a refinement solve);
// norm[b] takes max |rho| over instance b's rows.
template <class TJ, bool WRITE>
__device__ __forceinline__ void linres(int nint, Batch bt, const double* __restrict__ t, const TJ* __restrict__ jac,
                                       const double* __restrict__ dy, const double* __restrict__ dyd,
                                       const double* __restrict__ dz, const double* __restrict__ dp,
                                       double* __restrict__ fa, double* __restrict__ fb, double* __restrict__ rbc,
                                       const double* __restrict__ Dr, unsigned long long* __restrict__ norm) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nli = launch_intervals(bt, nint);
  const long long nsr = (long long)nli * M * NYZ;   // stage rows
  const long long ncr = (long long)nli * NY;        // continuity rows
  double v = 0.0;
  int b = -1;
  if (gid < nsr) {
    const int k = (int)(gid / NYZ), r = (int)(gid - (long long)k * NYZ);
    const int i = listed_interval(bt, k / M), j = k - (k / M) * M;
    b = bt.iinst[i];
    if (skipped(bt, b)) {
      b = -1;
    } else {
      const int st = i * M + j;
      const int nd = i + b;
      const double h = t[nd + 1] - t[nd];
      const TJ* row = jac + (size_t)st * JSTAGE + (size_t)r * NJC;
      const double* ydi = dyd + (size_t)i * M * NY;
      double s = fa[(size_t)st * NYZ + r];
      for (int c = 0; c < NY; ++c) {
        double q = 0.0;
#pragma unroll
        for (int kk = 0; kk < M; ++kk) q = fma(c_rkA[j * M + kk], ydi[kk * NY + c], q);
        s = fma((double)row[c], fma(h, q, dy[(size_t)nd * NY + c]), s);
      }
      for (int c = 0; c < NZ; ++c) s = fma((double)row[NY + c], dz[(size_t)st * NZ + c], s);
      for (int c = 0; c < NP; ++c) s = fma((double)row[NY + NZ + c], dp[(size_t)b * NP + c], s);
      if (r < NY) s -= ydi[j * NY + r];
      if (WRITE) fa[(size_t)st * NYZ + r] = s;
      v = s;
    }
  } else if (gid < nsr + ncr) {
    const int k = (int)((gid - nsr) / NY), l = (int)(gid - nsr - (long long)k * NY);
    const int i = listed_interval(bt, k);
    b = bt.iinst[i];
    if (skipped(bt, b)) {
      b = -1;
    } else {
      const int nd = i + b;
      const double h = t[nd + 1] - t[nd];
      double q = 0.0;
#pragma unroll
      for (int kk = 0; kk < M; ++kk) q = fma(c_rkb[kk], dyd[((size_t)i * M + kk) * NY + l], q);
      const double s = fb[(size_t)i * NY + l] + dy[(size_t)(nd + 1) * NY + l] - dy[(size_t)nd * NY + l] - h * q;
      if (WRITE) fb[(size_t)i * NY + l] = s;
      v = s;
    }
  } else if (gid < nsr + ncr + (long long)bt.B * NRB) {
    const int k = (int)((gid - nsr - ncr) / NRB), rr = (int)(gid - nsr - ncr - (long long)k * NRB);
    b = k;
    if (skipped(bt, b)) {
      b = -1;
    } else {
      const double* D = Dr + (size_t)b * DRSZ + (size_t)rr * (2 * NY + NP);
      const double* y0 = dy + (size_t)bt.noff[b] * NY;
      const double* y1 = dy + (size_t)(bt.noff[b + 1] - 1) * NY;
      double s = rbc[(size_t)b * NRB + rr];
      for (int c = 0; c < NY; ++c) s = fma(D[c], y0[c], s);
      for (int c = 0; c < NY; ++c) s = fma(D[NY + c], y1[c], s);
      for (int c = 0; c < NP; ++c) s = fma(D[2 * NY + c], dp[(size_t)b * NP + c], s);
      if (WRITE) rbc[(size_t)b * NRB + rr] = s;
      v = s;
    }
  }
  seg_max_to(fabs(v), b, norm);
}

// out = a + b over n state entries (b == nullptr: out = a).
__device__ __forceinline__ void state_add(long long n, const double* __restrict__ a, const double* __restrict__ b,
                                          double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = b ? a[e] + b[e] : a[e];
}

#if !COLLOC_TABLE_FORM
// -------------------------------------------------------- initial input
// form_initial_input_kernel (solver:316-356).
__device__ __forceinline__ void init_input(int nint, Batch bt, const double* __restrict__ y0,
                           const double* __restrict__ z0, const double* __restrict__ p,
                           const double* __restrict__ c, double* __restrict__ ydot, double* __restrict__ ztil) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nint * M) return;
  const int i = gid / M, j = gid - (gid / M) * M;
  const int ib = bt.iinst[i];
  const int nd = i + ib;
  const double cj = c_rkc[j];
  double yt[NY], zt[NZ1], fo[NY];
#pragma unroll
  for (int l = 0; l < NY; ++l) yt[l] = (1.0 - cj) * y0[(size_t)nd * NY + l] + cj * y0[(size_t)(nd + 1) * NY + l];
#pragma unroll
  for (int l = 0; l < NZ; ++l) {
    zt[l] = (1.0 - cj) * z0[(size_t)nd * NZ + l] + cj * z0[(size_t)(nd + 1) * NZ + l];
    ztil[(size_t)gid * NZ + l] = zt[l];
  }
  prob::f(yt, zt, p + (size_t)ib * NP, c + (size_t)ib * NC, fo);
#pragma unroll
  for (int l = 0; l < NY; ++l) ydot[(size_t)gid * NY + l] = fo[l];
}

// ------------------------------------------------------- segment residual
// Per-interval acceptance residual at the m+1 Gauss points.
//  relative == 0: compute_segment_residual_kernel (solver:1833-1892) with
//    residual_type 0 (max), 1 (sum of |.|) or 2 (sum of squares), and with
//    scale_by_time the scaling rho *= dt applied inside the component loop,
//    as the reference does; for residual_type 0 the g part takes |h| of the
//    same component index (solver:1880), components beyond ny contribute 0.
//  relative == 1: compute_segment_residual_kernel2 (solver:1959-2013): the
//    largest |h_k| / (1 + max_gp |ydot_k|) and |g_k| / (1 + max_gp |z_k|).
// Entry nint is ||r||_inf / tol (solver:1811-1817).
__device__ __forceinline__ void gauss_point(int gp, double h, const double* __restrict__ yi,
                                            const double* __restrict__ yd, const double* __restrict__ zz,
                                            double* ydt, double* yt, double* zt) {
#pragma unroll
  for (int l = 0; l < NY; ++l) {
    double s = 0.0;
    double yy = yi[l];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      s += c_gL[gp * M + k] * yd[k * NY + l];
      yy += h * c_gI[gp * M + k] * yd[k * NY + l];
    }
    ydt[l] = s;
    yt[l] = yy;
  }
#pragma unroll
  for (int l = 0; l < NZ; ++l) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) s += c_gL[gp * M + k] * zz[k * NZ + l];
    zt[l] = s;
  }
}

__device__ __forceinline__ void segres(int nint, Batch bt, const double* __restrict__ t,
                       const double* __restrict__ y, const double* __restrict__ ydot,
                       const double* __restrict__ ztil, const double* __restrict__ p,
                       const double* __restrict__ c, double alpha, double tol,
                       int rtype, int scale_time, int relative,
                       double* __restrict__ residual, unsigned long long* __restrict__ maxres) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double val = 0.0;
  int ib = -1;
  if (i < nint) {
    ib = bt.iinst[i];
    const int nd = i + ib;
    p += (size_t)ib * NP;
    c += (size_t)ib * NC;
    const double h = t[nd + 1] - t[nd];
    const double* yi = y + (size_t)nd * NY;
    const double* yd = ydot + (size_t)i * M * NY;
    const double* zz = ztil + (size_t)i * M * NZ;
    double rho_h = 0.0, rho_g = 0.0;
    if (!relative) {
      for (int gp = 0; gp < M + 1; ++gp) {
        double ydt[NY], yt[NY], zt[NZ1], fo[NY], go[NZ1];
        gauss_point(gp, h, yi, yd, zz, ydt, yt, zt);
        prob::f(yt, zt, p, c, fo);
        const double w = c_gw[gp];
#pragma unroll
        for (int l = 0; l < NY; ++l) {
          fo[l] -= ydt[l];
          const double e = fabs(fo[l]);
          if (rtype == 2) rho_h += w * fo[l] * fo[l];
          else if (rtype == 1) rho_h += w * e;
          else rho_h = fmax(rho_h, w * e);
          if (scale_time) rho_h *= h;
        }
        if (NZ > 0) {
          prob::g(yt, zt, p, c, alpha, go);
#pragma unroll
          for (int l = 0; l < NZ; ++l) {
            if (rtype == 2) rho_g += w * go[l] * go[l];
            else if (rtype == 1) rho_g += w * fabs(go[l]);
            else rho_g = fmax(rho_g, l < NY ? w * fabs(fo[l < NY ? l : 0]) : 0.0);
            if (scale_time) rho_g *= h;
          }
        }
      }
      if (rtype == 2) val = sqrt(rho_h + rho_g) / tol;
      else if (rtype == 1) val = (fabs(rho_h) + fabs(rho_g)) / tol;
      else val = fmax(rho_h, rho_g) / tol;
    } else {
      double ydm[NY], zm[NZ1];
#pragma unroll
      for (int l = 0; l < NY; ++l) ydm[l] = 0.0;
#pragma unroll
      for (int l = 0; l < NZ; ++l) zm[l] = 0.0;
      for (int gp = 0; gp < M + 1; ++gp) {
        double ydt[NY], yt[NY], zt[NZ1];
        gauss_point(gp, h, yi, yd, zz, ydt, yt, zt);
#pragma unroll
        for (int l = 0; l < NY; ++l) ydm[l] = fmax(ydm[l], fabs(ydt[l]));
#pragma unroll
        for (int l = 0; l < NZ; ++l) zm[l] = fmax(zm[l], fabs(zt[l]));
      }
      for (int gp = 0; gp < M + 1; ++gp) {
        double ydt[NY], yt[NY], zt[NZ1], fo[NY], go[NZ1];
        gauss_point(gp, h, yi, yd, zz, ydt, yt, zt);
        prob::f(yt, zt, p, c, fo);
#pragma unroll
        for (int l = 0; l < NY; ++l) rho_h = fmax(rho_h, fabs(fo[l] - ydt[l]) / (1.0 + ydm[l]));
        if (NZ > 0) {
          prob::g(yt, zt, p, c, alpha, go);
#pragma unroll
          for (int l = 0; l < NZ; ++l) rho_g = fmax(rho_g, fabs(go[l]) / (1.0 + zm[l]));
        }
      }
      val = fmax(rho_h, rho_g) / tol;
    }
    residual[nd] = val;
  } else if (i < nint + bt.B) {
    ib = i - nint;
    const int last = bt.noff[ib + 1] - 1;
    double ro[NY + NP];
    prob::r(y + (size_t)bt.noff[ib] * NY, y + (size_t)last * NY, p + (size_t)ib * NP, c + (size_t)ib * NC, ro);
    double m = 0.0;
#pragma unroll
    for (int l = 0; l < NY + NP; ++l) m = nmax(fabs(ro[l]), m);
    val = m / tol;
    residual[last] = val;
  }
  seg_max_to(val, ib, maxres);
}

#endif  // !COLLOC_TABLE_FORM

// ---------------------------------------------------------- recover z
// recover_solution_kernel (solver:2049-2075).
__device__ __forceinline__ void recover_z(int nnodes, Batch bt, const double* __restrict__ ztil,
                                          double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  const int ib = bt.ninst[i];
  const bool last = i == bt.noff[ib + 1] - 1;
  const int iv = (last ? i - 1 : i) - ib;   // interval whose stages give node i
#pragma unroll
  for (int l = 0; l < NZ; ++l) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) s += (last ? c_L1[j] : c_L0[j]) * ztil[((size_t)iv * M + j) * NZ + l];
    z[(size_t)i * NZ + l] = s;
  }
}

// ------------------------------------------- device-resident Newton loop
// The per-instance decisions of the damped Newton loop (solver:109-174) as
// kernels, so that one iteration of every instance of a batch runs from
// device memory without a host round trip (replayed as a HIP graph).
// nrm[b]: ||f_q||_inf of instance b's iterate; taken[b]: its accepted step
// of this iteration (0: none yet); trials[b]: failed trial steps;
// done / gave: the iteration at which it converged / the progress test
// stopped it (-1: not); hist[b][hl]: norm history; ring[b][window + 1]: the
// last norms for the progress test; *it: the iteration; cnt: [0] instances
// still iterating after the check, [1] still searching after a trial, [2]
// intervals in the launch list; *par: the solve's settings (in device
// memory, so that a captured iteration serves any of them).
struct NtParams {
  double tol, min_ratio;
  int hl, window, patience, max_iter, max_trials, pad;
};
struct NtState {
  double* nrm;
  double* taken;
  double* next;                  // norm of the accepted trial step (the next iterate's)
  int* trials;
  int* done;
  int* gave;
  double* hist;
  double* ring;
  int* it;
  unsigned long long* cnt;
  int* nlist;
  int* lofs;
  const NtParams* par;
};

// After the residual of the iterate (norms in slot[b]; slot null: the
// iterate is the accepted trial of the previous iteration, whose norm the
// line search kept in next[b]): record it, stop the instances that converged
// (||f_q|| < tol, before the step, solver:116-120) or that the progress test
// stops, and start the line search of the others at step 1.
__device__ __forceinline__ void nt_check(int B, NtState st, const unsigned long long* __restrict__ slot, int* act,
                                         int* srch, double* steps) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (!act[b]) {
    srch[b] = 0;
    steps[b] = 0.0;
    return;
  }
  const int it = *st.it;
  const NtParams P = *st.par;
  const double cur = slot ? __longlong_as_double((long long)slot[b]) : st.next[b];
  st.nrm[b] = cur;
  if (it < P.hl) st.hist[(size_t)b * P.hl + it] = cur;
  const int rs = P.patience > 0 ? P.window + 1 : 1;
  bool on = true;
  if (cur < P.tol) {
    on = false;
    st.done[b] = it;
  } else if (P.patience > 0 && it >= P.patience && it >= P.window &&
             !(cur * P.min_ratio <= st.ring[(size_t)b * rs + (it - P.window) % rs])) {
    on = false;
    st.gave[b] = it;
  }
  st.ring[(size_t)b * rs + it % rs] = cur;
  act[b] = on ? 1 : 0;
  srch[b] = on ? 1 : 0;
  steps[b] = on ? 1.0 : 0.0;
  st.taken[b] = 0.0;
  st.trials[b] = 0;
  if (on) {
    atomicAdd(st.cnt, 1ull);
    atomicAdd(st.cnt + 1, 1ull);
  }
}

// Launch list of the intervals of the instances still iterating: one
// workgroup scans the instances' interval counts (lofs[b]: the first slot
// of instance b), then every interval of an active instance takes its slot.
constexpr int NT_SCAN = 1024;
__device__ __forceinline__ void nt_list_scan(Batch bt, NtState st) {
  __shared__ int part[NT_SCAN];
  const int tid = threadIdx.x;
  const int per = (bt.B + NT_SCAN - 1) / NT_SCAN;
  const int b0 = tid * per;
  int sum = 0;
  for (int b = b0; b < b0 + per && b < bt.B; ++b)
    if (bt.act[b]) sum += bt.noff[b + 1] - bt.noff[b] - 1;
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < NT_SCAN; off <<= 1) {   // inclusive scan (Hillis-Steele)
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int b = b0; b < b0 + per && b < bt.B; ++b) {
    st.lofs[b] = run;
    if (bt.act[b]) run += bt.noff[b + 1] - bt.noff[b] - 1;
  }
  if (tid == NT_SCAN - 1) {
    *st.nlist = part[tid];
    st.cnt[2] = (unsigned long long)part[tid];
  }
}

__device__ __forceinline__ void nt_list_fill(int nint, Batch bt, NtState st, int* list) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nint) return;
  const int b = bt.iinst[i];
  if (bt.act[b]) list[st.lofs[b] + (i - (bt.noff[b] - b))] = i;
}

// After a trial step's residual (norms in slot[b]): the first step that
// strictly decreases the norm is taken (solver:155-174); otherwise the step
// is halved, and after max_trials failures the search ends without a step.
__device__ __forceinline__ void nt_line_search(int B, NtState st, const unsigned long long* __restrict__ slot,
                                               int* srch, double* steps) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !srch[b]) return;
  const double cur = __longlong_as_double((long long)slot[b]);
  if (cur < st.nrm[b]) {
    st.taken[b] = steps[b];
    st.next[b] = cur;
    srch[b] = 0;
    steps[b] = 0.0;
    return;
  }
  const int tr = st.trials[b] + 1;
  st.trials[b] = tr;
  if (tr >= st.par->max_trials) {
    srch[b] = 0;
    steps[b] = 0.0;
    return;
  }
  steps[b] *= 0.5;
  atomicAdd(st.cnt + 1, 1ull);
}

// End of the line search: the accepted step becomes the state update's
// step; an instance that took none keeps its state, so every later
// iteration would repeat this one: it stops here, its history continued at
// the same norm (the host loop's stall rule).
__device__ __forceinline__ void nt_accept(int B, NtState st, int* act, double* steps) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (!act[b]) {
    steps[b] = 0.0;
    return;
  }
  const double tk = st.taken[b];
  steps[b] = tk;
  if (tk == 0.0) {
    act[b] = 0;
    const int it = *st.it;
    const int hl = st.par->hl, mi = st.par->max_iter;
    const int end = mi < hl ? mi : hl;
    for (int k = it + 1; k < end; ++k) st.hist[(size_t)b * hl + k] = st.nrm[b];
  }
}

#if !COLLOC_LARGE
// The series condensation kernels run one work item (a listed interval) per
// workgroup (persist = 0), or as a persistent grid (persist = 1): each
// workgroup takes the next item from a ticket counter (*ticket, cleared by
// the runtime before the launch) until the launch list is exhausted, so a
// grid sized to the machine serves any list length, including one known
// only on the device (a captured Newton iteration).  The body of an item
// ends with its last LDS read before any barrier-free exit, and the next
// item starts after the two barriers around the ticket.
template <class F>
__device__ __forceinline__ void series_items(int nint, const Batch& bt, int persist, unsigned long long* ticket,
                                             F&& body) {
  if (!persist) {
    if (beyond_list(bt, nint)) return;
    body((int)blockIdx.x);
    return;
  }
  __shared__ int sitem;
  const int n = launch_intervals(bt, nint);
  for (;;) {
    if (threadIdx.x == 0) sitem = (int)atomicAdd(ticket, 1ull);
    __syncthreads();
    const int item = sitem;
    __syncthreads();
    if (item >= n) break;
    body(item);
  }
}
// A kernel whose blocks do not fit leaves the item to Gauss-Jordan.
__device__ __forceinline__ void queue_all(int item, const Batch& bt, int* fallback, unsigned long long* counters) {
  if (threadIdx.x == 0) {
    const int i = listed_interval(bt, item);
    if (!(bt.act && !bt.act[bt.iinst[i]])) queue_gauss_jordan(i, fallback, counters);
  }
}
#endif  // !COLLOC_LARGE

}  // namespace colloc

#if COLLOC_TABLE_FORM || COLLOC_JAC_GEMM
#include "colloc_tables.h"
#endif
#if COLLOC_LARGE
#include "colloc_large.h"
#endif

// ------------------------------------------------------------ entry points
extern "C" {
__global__ void colloc_fill_u64(unsigned long long* p, unsigned long long v, int n) { colloc::fill_u64(p, v, n); }

__global__ __launch_bounds__(256) void colloc_nt_check(int B, colloc::NtState st, const unsigned long long* slot,
    int* act, int* srch, double* steps) {
  colloc::nt_check(B, st, slot, act, srch, steps);
}
__global__ __launch_bounds__(colloc::NT_SCAN) void colloc_nt_list_scan(colloc::Batch bt, colloc::NtState st) {
  colloc::nt_list_scan(bt, st);
}
__global__ __launch_bounds__(256) void colloc_nt_list_fill(int nint, colloc::Batch bt, colloc::NtState st, int* list) {
  colloc::nt_list_fill(nint, bt, st, list);
}
__global__ __launch_bounds__(256) void colloc_nt_line_search(int B, colloc::NtState st,
    const unsigned long long* slot, int* srch, double* steps) {
  colloc::nt_line_search(B, st, slot, srch, steps);
}
__global__ __launch_bounds__(256) void colloc_nt_accept(int B, colloc::NtState st, int* act, double* steps) {
  colloc::nt_accept(B, st, act, steps);
}
__global__ void colloc_nt_tick(colloc::NtState st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *st.it += 1;
}

// The table form launches one workgroup per TAB_G points (stages, then
// instances), the straight-line form one thread per point.
#if COLLOC_TABLE_FORM
#define COLLOC_CB(name) name##_tab
constexpr int kTabPoints = colloc::TAB_G;
#else
#define COLLOC_CB(name) name
constexpr int kTabPoints = 0;
#endif
__global__ __launch_bounds__(256) void colloc_residual_full(int nint, colloc::Batch bt, const double* t,
    const double* y, const double* ydot, const double* ztil, const double* p, const double* c, double alpha,
    double* ytil, double* fa, double* fb, double* rbc, unsigned long long* norm) {
  colloc::COLLOC_CB(residual)<true>(nint, bt, t, y, ydot, ztil, p, c, alpha, ytil, fa, fb, rbc, norm);
}

__global__ __launch_bounds__(256) void colloc_residual_norm(int nint, colloc::Batch bt, const double* t,
    const double* y, const double* ydot, const double* ztil, const double* p, const double* c, double alpha,
    double* ytil, double* fa, double* fb, double* rbc, unsigned long long* norm) {
  colloc::COLLOC_CB(residual)<false>(nint, bt, t, y, ydot, ztil, p, c, alpha, ytil, fa, fb, rbc, norm);
}

__global__ __launch_bounds__(256) void colloc_jacobian(int nint, colloc::Batch bt, const double* y,
    const double* ytil, const double* ztil, const double* p, const double* c, double alpha, double* jac, double* Dr) {
  colloc::COLLOC_CB(jacobian)<double>(nint, bt, y, ytil, ztil, p, c, alpha, jac, Dr);
}
// stage Jacobians stored in float32 (the FP32 direction)
__global__ __launch_bounds__(256) void colloc_jacobian32(int nint, colloc::Batch bt, const double* y,
    const double* ytil, const double* ztil, const double* p, const double* c, double alpha, float* jac, double* Dr) {
  colloc::COLLOC_CB(jacobian)<float>(nint, bt, y, ytil, ztil, p, c, alpha, jac, Dr);
}
#if COLLOC_JAC_GEMM
// the same on matrix cores (colloc_tables.h jacobian_gemm); colloc_jac_gemm
// holds its stages per workgroup for the host's grid
__global__ __launch_bounds__(256) void colloc_jacobian32_gemm(int nint, colloc::Batch bt, const double* y,
    const double* ytil, const double* ztil, const double* p, const double* c, double alpha, float* jac, double* Dr) {
  colloc::jacobian_gemm(nint, bt, y, ytil, ztil, p, c, alpha, jac, Dr);
}
__device__ int colloc_jac_gemm = prob::JG_POINTS;
#endif
__global__ __launch_bounds__(256) void colloc_jacobian_static(long long nst, double* jac) {
  colloc::jacobian_static<double>(nst, jac);
}
__global__ __launch_bounds__(256) void colloc_jacobian_static32(long long nst, float* jac) {
  colloc::jacobian_static<float>(nst, jac);
}

#if !COLLOC_LARGE
__global__ __launch_bounds__(Cfg::CW * 64) void colloc_condense(const int* iinst, const int* queue,
    const unsigned long long* nq, const double* t, const double* jac, const double* fa, const double* fb,
    double* rows, double* rec, int* status) {
  colloc::condense<Cfg::CW, double>(iinst, queue, nq, t, jac, fa, fb, rows, rec, status);
}
// the float64 Gauss-Jordan fallback of the FP32 direction (float32 stage Jacobians)
__global__ __launch_bounds__(Cfg::CW * 64) void colloc_condense32(const int* iinst, const int* queue,
    const unsigned long long* nq, const double* t, const float* jac, const double* fa, const double* fb,
    double* rows, float* rec, int* status) {
  colloc::condense<Cfg::CW, float>(iinst, queue, nq, t, jac, fa, fb, rows, rec, status);
}

__global__ __launch_bounds__(colloc::NS_THREADS) void colloc_condense_ns(int nint, colloc::Batch bt,
    const double* t, const double* jac, const double* fa, const double* fb, double* rows, double* rec, int* status,
    int* fallback, unsigned long long* counters) {
  if (colloc::beyond_list(bt, nint)) return;
  if constexpr (colloc::NS_OK) colloc::condense_ns<double>(blockIdx.x, bt, t, jac, fa, fb, rows, rec, status, fallback, counters);
  else colloc::queue_all(blockIdx.x, bt, fallback, counters);
}
// The persistent-grid forms (a work queue over the launch list) are their
// own kernels, so that the loop does not cost the one-workgroup-per-interval
// forms registers.
__global__ __launch_bounds__(colloc::NS_THREADS) void colloc_condense_ns_q(int nint, colloc::Batch bt,
    const double* t, const double* jac, const double* fa, const double* fb, double* rows, double* rec, int* status,
    int* fallback, unsigned long long* counters) {
  colloc::series_items(nint, bt, 1, counters + 3, [&](int item) {
    if constexpr (colloc::NS_OK) colloc::condense_ns<double>(item, bt, t, jac, fa, fb, rows, rec, status, fallback, counters);
    else colloc::queue_all(item, bt, fallback, counters);
  });
}

// The constant blocks of the condensation's BABD rows (the cyclic
// reduction's level-0 rows): C = I and, where the stage functions depend on
// only some parameters, the zero H columns of the others.  No later level
// writes into level 0's rows, so the runtime writes these once per mesh and
// the condensation kernels write only G, the active H columns and b.
// One workgroup per interval.
__global__ __launch_bounds__(256) void colloc_rows_constant(int nint, double* rows) {
  const int i = blockIdx.x;
  if (i >= nint) return;
  double* rowi = rows + (size_t)i * colloc::NY * colloc::CRW;
  for (int e = threadIdx.x; e < colloc::NY * colloc::NY; e += blockDim.x) {
    const int a = e / colloc::NY, q = e - (e / colloc::NY) * colloc::NY;
    rowi[a * colloc::CRW + colloc::NY + q] = (a == q) ? 1.0 : 0.0;
  }
  if constexpr (!colloc::HZ && colloc::NPA < colloc::NP) {
    for (int e = threadIdx.x; e < colloc::NY * colloc::NP; e += blockDim.x) {
      const int a = e / colloc::NP, j = e - (e / colloc::NP) * colloc::NP;
      if (!colloc::p_active(j)) rowi[a * colloc::CRW + 2 * colloc::NY + j] = 0.0;
    }
  }
}

// The FP32-direction condensation with the series products on matrix cores:
// one wave per stage block (the runtime launches it when colloc_caps[0] says
// the problem's blocks fit, else colloc_condense_ns32).
__global__ __launch_bounds__(colloc::MF_THREADS) __attribute__((amdgpu_waves_per_eu(colloc::MF_WAVES)))
void colloc_condense_mf32(int nint, colloc::Batch bt,
    const double* t, const float* jac, const double* fa, const double* fb, double* rows, float* rec, int* status,
    int* fallback, unsigned long long* counters) {
  if (colloc::beyond_list(bt, nint)) return;
  if constexpr (colloc::MF_OK) colloc::condense_mf32(blockIdx.x, bt, t, jac, fa, fb, rows, rec, status, fallback, counters);
  else colloc::queue_all(blockIdx.x, bt, fallback, counters);
}
__global__ __launch_bounds__(colloc::MF_THREADS) void colloc_condense_mf32_q(int nint, colloc::Batch bt,
    const double* t, const float* jac, const double* fa, const double* fb, double* rows, float* rec, int* status,
    int* fallback, unsigned long long* counters) {
  colloc::series_items(nint, bt, 1, counters + 3, [&](int item) {
    if constexpr (colloc::MF_OK) colloc::condense_mf32(item, bt, t, jac, fa, fb, rows, rec, status, fallback, counters);
    else colloc::queue_all(item, bt, fallback, counters);
  });
}
// capabilities of this problem's kernels: [0] colloc_condense_mf32 runs the
// matrix-core series (else it leaves every interval to Gauss-Jordan);
// [1] the block size of colloc_cr_pair; [2] the BABD rows carry no H block
// (no stage function depends on a parameter: H is identically zero);
// [3] workgroups per 256 stages of colloc_jacobian (threads per stage)
// [4] points per workgroup of the table-form callback kernels (0: one
// thread per point, straight-line callbacks); [5] the large-block kernels
// (colloc_large.h) replace the register-resident ones
__device__ int colloc_caps[8] = {colloc::MF_OK ? 1 : 0, colloc::CR_THREADS, colloc::HZ ? 1 : 0, prob::JAC_PARTS,
                                 kTabPoints, 0, 0, 0};

// The FP32-direction condensation: two 512-thread workgroups per CU (half the
// LDS and registers of the float64 kernel).
__global__ __launch_bounds__(colloc::NS_THREADS, 4) void colloc_condense_ns32(int nint, colloc::Batch bt,
    const double* t, const float* jac, const double* fa, const double* fb, double* rows, float* rec, int* status,
    int* fallback, unsigned long long* counters) {
  if (colloc::beyond_list(bt, nint)) return;
  if constexpr (colloc::ns_ok<float>()) colloc::condense_ns<float>(blockIdx.x, bt, t, jac, fa, fb, rows, rec, status, fallback, counters);
  else colloc::queue_all(blockIdx.x, bt, fallback, counters);
}
__global__ __launch_bounds__(colloc::NS_THREADS, 4) void colloc_condense_ns32_q(int nint, colloc::Batch bt,
    const double* t, const float* jac, const double* fa, const double* fb, double* rows, float* rec, int* status,
    int* fallback, unsigned long long* counters) {
  colloc::series_items(nint, bt, 1, counters + 3, [&](int item) {
    if constexpr (colloc::ns_ok<float>()) colloc::condense_ns<float>(item, bt, t, jac, fa, fb, rows, rec, status, fallback, counters);
    else colloc::queue_all(item, bt, fallback, counters);
  });
}

__global__ __launch_bounds__(colloc::CR_THREADS) void colloc_cr_pair(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, double* rows_out, double* recs) {
  colloc::cr_pair<double, double>(oinst, seg, act, rows_in, rows_out, recs);
}
// the FP32 cyclic reduction: its first level (float64 rows of the
// condensation in, float32 out) and the others (float32 in and out)
__global__ __launch_bounds__(colloc::CR_THREADS) void colloc_cr_pair_df(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, float* rows_out, float* recs) {
  colloc::cr_pair<double, float>(oinst, seg, act, rows_in, rows_out, recs);
}
__global__ __launch_bounds__(colloc::CR_THREADS) void colloc_cr_pair_ff(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const float* rows_in, float* rows_out, float* recs) {
  colloc::cr_pair<float, float>(oinst, seg, act, rows_in, rows_out, recs);
}

__global__ __launch_bounds__(colloc::CS_THREADS) void colloc_cr_subst(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, double* rows_out, double* recs) {
  colloc::cr_subst<double, double>(oinst, seg, act, rows_in, rows_out, recs);
}
__global__ __launch_bounds__(colloc::CS_THREADS) void colloc_cr_subst_df(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const double* rows_in, float* rows_out, float* recs) {
  colloc::cr_subst<double, float>(oinst, seg, act, rows_in, rows_out, recs);
}
__global__ __launch_bounds__(colloc::CS_THREADS) void colloc_cr_subst_ff(const int* oinst, const colloc::CRSeg* seg,
    const int* act, const float* rows_in, float* rows_out, float* recs) {
  colloc::cr_subst<float, float>(oinst, seg, act, rows_in, rows_out, recs);
}

__global__ __launch_bounds__(Cfg::FW * 64) void colloc_cr_final(colloc::Batch bt, const double* rows,
    const double* Dr, const double* rbc, double* dy, double* dp, int* status) {
  colloc::cr_final<Cfg::FW, double>(bt, rows, Dr, rbc, dy, dp, status);
}
__global__ __launch_bounds__(Cfg::FW * 64) void colloc_cr_final_f(colloc::Batch bt, const float* rows,
    const double* Dr, const double* rbc, double* dy, double* dp, int* status) {
  colloc::cr_final<Cfg::FW, float>(bt, rows, Dr, rbc, dy, dp, status);
}

__global__ __launch_bounds__(64) void colloc_cr_back(int level, colloc::Batch bt, const int* oinst,
    const colloc::CRSeg* seg, const double* recs, const double* rows_in, double* dy, const double* dp) {
  colloc::cr_back<double, double>(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
// FP32 cyclic reduction: float32 records with the first level's float64 input rows, or float32 both
__global__ __launch_bounds__(64) void colloc_cr_back_fd(int level, colloc::Batch bt, const int* oinst,
    const colloc::CRSeg* seg, const float* recs, const double* rows_in, double* dy, const double* dp) {
  colloc::cr_back<float, double>(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
__global__ __launch_bounds__(64) void colloc_cr_back_ff(int level, colloc::Batch bt, const int* oinst,
    const colloc::CRSeg* seg, const float* recs, const float* rows_in, double* dy, const double* dp) {
  colloc::cr_back<float, float>(level, bt, oinst, seg, recs, rows_in, dy, dp);
}
#endif  // !COLLOC_LARGE

__global__ __launch_bounds__(256) void colloc_recover(int nint, colloc::Batch bt, const double* rec,
    const double* dy, const double* dp, double* dydot, double* dz) {
  if (colloc::beyond_list(bt, nint)) return;
  colloc::recover<double>(nint, bt, rec, dy, dp, dydot, dz);
}
// condensation records stored in float32 (the FP32 direction)
__global__ __launch_bounds__(256) void colloc_recover32(int nint, colloc::Batch bt, const float* rec,
    const double* dy, const double* dp, double* dydot, double* dz) {
  if (colloc::beyond_list(bt, nint)) return;
  colloc::recover<float>(nint, bt, rec, dy, dp, dydot, dz);
}

__global__ __launch_bounds__(256) void colloc_linres(int nint, colloc::Batch bt, const double* t, const double* jac,
    const double* dy, const double* dyd, const double* dz, const double* dp, double* fa, double* fb, double* rbc,
    const double* Dr, unsigned long long* norm, int write) {
  if (write) colloc::linres<double, true>(nint, bt, t, jac, dy, dyd, dz, dp, fa, fb, rbc, Dr, norm);
  else colloc::linres<double, false>(nint, bt, t, jac, dy, dyd, dz, dp, fa, fb, rbc, Dr, norm);
}
// stage Jacobians stored in float32 (the FP32 direction)
__global__ __launch_bounds__(256) void colloc_linres32(int nint, colloc::Batch bt, const double* t, const float* jac,
    const double* dy, const double* dyd, const double* dz, const double* dp, double* fa, double* fb, double* rbc,
    const double* Dr, unsigned long long* norm, int write) {
  if (write) colloc::linres<float, true>(nint, bt, t, jac, dy, dyd, dz, dp, fa, fb, rbc, Dr, norm);
  else colloc::linres<float, false>(nint, bt, t, jac, dy, dyd, dz, dp, fa, fb, rbc, Dr, norm);
}
__global__ __launch_bounds__(256) void colloc_state_add(long long n, const double* a, const double* b, double* out) {
  colloc::state_add(n, a, b, out);
}

__global__ __launch_bounds__(256) void colloc_axpy4(colloc::Seg4 s, colloc::Batch bt, const double* x,
    const double* d, const double* steps, double* out) {
  colloc::axpy4(s, bt, x, d, steps, out);
}

__global__ __launch_bounds__(64) void colloc_axpy_listed(colloc::Seg4 s, colloc::Batch bt, const double* x,
    const double* d, const double* steps, double* out) {
  colloc::axpy_listed(s, bt, x, d, steps, out);
}

__global__ __launch_bounds__(256) void colloc_init_input(int nint, colloc::Batch bt, const double* y0,
    const double* z0, const double* p, const double* c, double* ydot, double* ztil) {
  colloc::COLLOC_CB(init_input)(nint, bt, y0, z0, p, c, ydot, ztil);
}

__global__ __launch_bounds__(256) void colloc_segres(int nint, colloc::Batch bt, const double* t, const double* y,
    const double* ydot, const double* ztil, const double* p, const double* c, double alpha, double tol,
    int rtype, int scale_time, int relative, double* residual, unsigned long long* maxres) {
  colloc::COLLOC_CB(segres)(nint, bt, t, y, ydot, ztil, p, c, alpha, tol, rtype, scale_time, relative, residual, maxres);
}

__global__ __launch_bounds__(256) void colloc_recover_z(int nnodes, colloc::Batch bt, const double* ztil, double* z) {
  colloc::recover_z(nnodes, bt, ztil, z);
}
}
