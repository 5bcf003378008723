// The callbacks of a problem in the polynomial table form (large problems,
// whose straight-line code would be too large): an evaluator of
// prob::*_prog() at TAB_G points per workgroup, and the kernels that call
// the callbacks (residual, stage Jacobians, initial input, segment
// residual), with the same arguments and results as their straight-line
// counterparts in colloc_kernels.h.  A straight-line unit uses the
// evaluator for its matrix-core float32 Jacobian (COLLOC_JAC_GEMM).
//
// A workgroup holds, in LDS, the atoms of its points as ext[atom][G] (the
// caller fills the nb base atoms: y | z | p | instance constants | alpha;
// the evaluator the nodes) and the monomials as mv[monomial][G].  A node or
// output row is one lane of a wave (a 64-row slab per wave); its terms are
// read coalesced (term-major slabs) and each is one fused multiply-add per
// point with the monomial's G values from LDS.
#pragma once

namespace colloc {

constexpr int TAB_G = 8;          // points (stages, instances) per workgroup
constexpr int TAB_THREADS = 256;

template <class T>
__device__ __forceinline__ T ipow(T x, int k) {
  T r = (T)1;
  const int n = k < 0 ? -k : k;
  for (int q = 0; q < n; ++q) r *= x;
  return k < 0 ? (T)1 / r : r;
}

// the node functions, numbered as polyform.KINDS
template <class T>
__device__ __forceinline__ T node_value(int kind, T a, double r) {
  switch (kind) {
    case 0: return tanh(a);
    case 1: return exp(a);
    case 2: return sin(a);
    case 3: return cos(a);
    case 4: return tan(a);
    case 5: return sinh(a);
    case 6: return cosh(a);
    case 7: return sqrt(a);
    case 8: return log(a);
    case 9: return asin(a);
    case 10: return acos(a);
    case 11: return atan(a);
    case 12: return fabs(a);
    case 13: return a > (T)0 ? (T)1 : (a < (T)0 ? (T)-1 : (T)0);
    default: return pow(a, (T)r);
  }
}

// Row `lane` of slab s at the G points: acc[g] = sum_t coef_t mv[mono_t][g].
template <class T, int G>
__device__ __forceinline__ void slab_eval(const TabProg& P, int s, const T* __restrict__ mv, T (&acc)[G]) {
  const int lane = lane_id();
#pragma unroll
  for (int g = 0; g < G; ++g) acc[g] = (T)0;
  const int t1 = P.slab_term_off[s + 1];
  for (int t = P.slab_term_off[s] + lane; t < t1; t += WAVE) {
    const T c = (T)P.coef[t];
    const T* mp = mv + (size_t)P.mono[t] * G;
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = fma(c, mp[g], acc[g]);
  }
}

// The monomials and nodes of a program at the workgroup's G points, level
// by level (ext's base atoms filled by the caller, visible to every thread).
template <class T, int G>
__device__ __forceinline__ void tab_prepare(const TabProg& P, T* __restrict__ ext, T* __restrict__ mv) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nw = nthr / WAVE, w = wave_id(), lane = lane_id();
  for (int L = 0; L < P.nlev; ++L) {
    const int m0 = P.mono_lev_off[L], m1 = P.mono_lev_off[L + 1];
    for (int e = tid; e < (m1 - m0) * G; e += nthr) {
      const int k = m0 + e / G, g = e % G;
      T v = (T)1;
      for (int q = P.mono_fac_off[k]; q < P.mono_fac_off[k + 1]; ++q)
        v *= ipow(ext[(size_t)P.mono_fac_atom[q] * G + g], (int)P.mono_fac_pow[q]);
      mv[(size_t)k * G + g] = v;
    }
    __syncthreads();
    if (L + 1 < P.nlev) {
      for (int s = P.node_slab_off[L] + w; s < P.node_slab_off[L + 1]; s += nw) {
        T acc[G];
        slab_eval<T, G>(P, s, mv, acc);
        const int r = P.slab_rows[s * WAVE + lane];
        if (r >= 0) {
#pragma unroll
          for (int g = 0; g < G; ++g) ext[(size_t)(P.nb + r) * G + g] = node_value<T>(P.node_kind[r], acc[g], P.node_r[r]);
        }
      }
      __syncthreads();
    }
  }
}

// The output rows: out(destination, acc) for each row of each output slab.
template <class T, int G, class F>
__device__ __forceinline__ void tab_outputs(const TabProg& P, const T* __restrict__ mv, F&& out) {
  const int nw = blockDim.x / WAVE;
  for (int s = P.out0 + wave_id(); s < P.nslab; s += nw) {
    T acc[G];
    slab_eval<T, G>(P, s, mv, acc);
    const int d = P.slab_rows[s * WAVE + threadIdx.x % WAVE];
    if (d >= 0) out(d, acc);
  }
}

// The base atoms that are not the point's y: z | p | instance constants | alpha.
template <class T, int G>
__device__ __forceinline__ void tab_base_rest(T* ext, int g, const double* z, const double* p, const double* c,
                                              double alpha, bool valid) {
  for (int a = NY + threadIdx.x / G; a < prob::TB_NB; a += blockDim.x / G) {
    double v = 0.0;
    if (valid) {
      if (a < NY + NZ) v = z[a - NY];
      else if (a < NY + NZ + NP) v = p[a - NY - NZ];
      else if (a < NY + NZ + NP + NC) v = c[a - NY - NZ - NP];
      else v = alpha;
    }
    ext[(size_t)a * G + g] = (T)v;
  }
}

#if COLLOC_JAC_GEMM
// ------------------------------------------- float32 Jacobian on matrix cores
// For a straight-line unit whose problem also has the polynomial form: the
// state-dependent entries of a stage block are C mv, C the constant matrix
// of the jac program's coefficients over its monomials and mv the
// monomials at the stage (tab_prepare, in LDS).  C's nonzero 16 x 4 blocks
// are stored row tile by row tile (prob::jg_blk_off, jg_blk_ks) as
// v_mfma_f32_16x16x4f32 A operands (prob::jg_coef; lane l: row 16 rt +
// l % 16, column 4 ks + l / 16).  Workgroup k < ceil(Nst / JG_POINTS)
// takes stages JG_POINTS k .. + JG_POINTS - 1 and its waves take the row
// tiles in turn, multiplying only the nonzero blocks (the monomials as B
// operands from LDS), each row tile's accumulators stored to the rows'
// destinations (prob::jg_dest) of every stage; the workgroups after them
// write Dr (straight-line prob::Dr, one thread per instance).  The static
// entries are the straight-line unit's (jacobian_static).
constexpr int JG_ST = prob::JG_POINTS / 16;     // 16-stage tiles per workgroup
static_assert(prob::JG_POINTS % 16 == 0, "stage tiles of 16");
__device__ __forceinline__ void jacobian_gemm(int nint, Batch bt, const double* __restrict__ y,
                                              const double* __restrict__ ytil, const double* __restrict__ ztil,
                                              const double* __restrict__ p, const double* __restrict__ c,
                                              double alpha, float* __restrict__ jac, double* __restrict__ Dr) {
  typedef float jg_f4 __attribute__((ext_vector_type(4)));
  constexpr int G = prob::JG_POINTS;
  __shared__ __attribute__((aligned(16))) float smem[(prob::JAC_NATOMS + prob::JAC_NMONO) * G];
  __shared__ long long sst[G];
  const int nst = launch_intervals(bt, nint) * M;
  const int nwg = (nst + G - 1) / G;
  if ((int)blockIdx.x >= nwg) {
    const int ib = ((int)blockIdx.x - nwg) * (int)blockDim.x + (int)threadIdx.x;
    if (ib < bt.B && !skipped(bt, ib))
      prob::Dr(y + (size_t)bt.noff[ib] * NY, y + (size_t)(bt.noff[ib + 1] - 1) * NY, p + (size_t)ib * NP,
               c + (size_t)ib * NC, Dr + (size_t)ib * DRSZ);
    return;
  }
  float* ext = smem;
  float* mv = smem + (size_t)prob::JAC_NATOMS * G;
  const int g = threadIdx.x % G;
  int st = 0, b = -1;
  const int gid = blockIdx.x * G + g;
  if (gid < nst) {
    const int i = listed_interval(bt, gid / M);
    st = i * M + (gid - (gid / M) * M);
    b = bt.iinst[i];
    if (skipped(bt, b)) b = -1;
  }
  const bool valid = b >= 0;
  for (int a = threadIdx.x / G; a < NY; a += blockDim.x / G)
    ext[(size_t)a * G + g] = valid ? (float)ytil[(size_t)st * NY + a] : 0.0f;
  tab_base_rest<float, G>(ext, g, ztil + (size_t)(valid ? st : 0) * NZ, p + (size_t)(valid ? b : 0) * NP,
                          c + (size_t)(valid ? b : 0) * NC, alpha, valid);
  if (threadIdx.x < G) sst[threadIdx.x] = valid ? (long long)st : -1;
  __syncthreads();
  tab_prepare<float, G>(prob::jac_prog(), ext, mv);   // ends with the monomials in LDS, synchronised
  const int lane = lane_id(), w = wave_id(), nw = blockDim.x / WAVE;
  long long sc[JG_ST];
#pragma unroll
  for (int tt = 0; tt < JG_ST; ++tt) sc[tt] = sst[16 * tt + (lane & 15)];
  for (int rt = w; rt < prob::JG_RT; rt += nw) {
    // the row tile's nonzero blocks: A operands and k-steps loaded first,
    // the B operands (monomials 4 ks .. 4 ks + 3 at the 16 stages of a tile)
    // from LDS
    const int j0 = prob::jg_blk_off[rt], nb = prob::jg_blk_off[rt + 1] - j0;
    float a[prob::JG_KMAX];
    int kq[prob::JG_KMAX];
#pragma unroll
    for (int q = 0; q < prob::JG_KMAX; ++q) {
      if (q < nb) {
        a[q] = prob::jg_coef[(size_t)(j0 + q) * WAVE + lane];
        kq[q] = prob::jg_blk_ks[j0 + q];
      }
    }
    jg_f4 acc[JG_ST];
#pragma unroll
    for (int tt = 0; tt < JG_ST; ++tt) acc[tt] = jg_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < prob::JG_KMAX; ++q) {
      if (q < nb) {
        const int k = 4 * kq[q] + (lane >> 4);
#pragma unroll
        for (int tt = 0; tt < JG_ST; ++tt) {
          const float bv = k < prob::JAC_NMONO ? mv[(size_t)k * G + 16 * tt + (lane & 15)] : 0.0f;
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q], bv, acc[tt], 0, 0, 0);
        }
      }
    }
    // value v of lane l: row 16 rt + 4 (l / 16) + v, stage 16 tt + l % 16
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int d = prob::jg_dest[16 * rt + 4 * (lane >> 4) + v];
      if (d < 0) continue;
#pragma unroll
      for (int tt = 0; tt < JG_ST; ++tt)
        if (sc[tt] >= 0) jac[(size_t)sc[tt] * JSTAGE + d] = acc[tt][v];
    }
  }
}
#endif  // COLLOC_JAC_GEMM

#if COLLOC_TABLE_FORM
constexpr int TAB_MAX2(int a, int b) { return a > b ? a : b; }
constexpr int TAB_FG_LDS = (prob::FG_NATOMS + prob::FG_NMONO + NY + NZ) * TAB_G;
constexpr int TAB_R_LDS = (TAB_MAX2(prob::RI_NATOMS + prob::RI_NMONO, prob::RF_NATOMS + prob::RF_NMONO)) * TAB_G;
constexpr int TAB_DR_LDS = (TAB_MAX2(prob::DRI_NATOMS + prob::DRI_NMONO, prob::DRF_NATOMS + prob::DRF_NMONO)) * TAB_G;

// r (or Dr) of the instances b0 .. b0+G-1 (b < B, not skipped): the initial
// rows from y at the first node, the final rows from y at the last node.
// out(b, g, destination, value).
template <int G, class F>
__device__ __forceinline__ void tab_boundary(const TabProg& PI, const TabProg& PF, int b0, const Batch& bt,
                                             const double* __restrict__ y, const double* __restrict__ p,
                                             const double* __restrict__ c, double* smem, F&& out) {
  for (int side = 0; side < 2; ++side) {
    const TabProg& P = side == 0 ? PI : PF;
    double* ext = smem;
    double* mv = smem + (size_t)P.natoms * G;
    const int g = threadIdx.x % G;
    const int b = b0 + g;
    const bool valid = b < bt.B && !skipped(bt, b);
    const double* yb = valid ? y + (size_t)(side == 0 ? bt.noff[b] : bt.noff[b + 1] - 1) * NY : y;
    for (int a = threadIdx.x / G; a < NY; a += blockDim.x / G) ext[(size_t)a * G + g] = valid ? yb[a] : 0.0;
    for (int a = NY + threadIdx.x / G; a < prob::TB_NB; a += blockDim.x / G) {
      double v = 0.0;   // z and alpha do not enter r
      if (valid && a >= NY + NZ && a < NY + NZ + NP) v = p[(size_t)b * NP + a - NY - NZ];
      else if (valid && a >= NY + NZ + NP && a < NY + NZ + NP + NC) v = c[(size_t)b * NC + a - NY - NZ - NP];
      ext[(size_t)a * G + g] = v;
    }
    __syncthreads();
    tab_prepare<double, G>(P, ext, mv);
    tab_outputs<double, G>(P, mv, [&](int d, const double (&acc)[G]) {
#pragma unroll
      for (int q = 0; q < G; ++q)
        if (b0 + q < bt.B && !skipped(bt, b0 + q)) out(b0 + q, d, acc[q]);
    });
    __syncthreads();
  }
}

// ---------------------------------------------------------------- residual
// The residual (the table-form restatement of residual<FULL> in
// colloc_kernels.h: collocation_update_kernel + compute_f_q_kernel1/2,
// solver:478-559): workgroup k < ceil(Nst / G) evaluates the stages
// G k .. G k + G - 1 of the launch, the workgroups after it the boundary
// conditions of G instances each.
template <bool FULL>
__device__ __forceinline__ void residual_tab(int nint, Batch bt, const double* __restrict__ t,
                                             const double* __restrict__ y, const double* __restrict__ ydot,
                                             const double* __restrict__ ztil, const double* __restrict__ p,
                                             const double* __restrict__ c, double alpha,
                                             double* __restrict__ ytil, double* __restrict__ fa,
                                             double* __restrict__ fb, double* __restrict__ rbc,
                                             unsigned long long* __restrict__ norm) {
  constexpr int G = TAB_G;
  __shared__ __attribute__((aligned(16))) double smem[TAB_MAX2(TAB_FG_LDS, TAB_R_LDS)];
  const int nst = launch_intervals(bt, nint) * M;
  const int nwg = (nst + G - 1) / G;
  const int g = threadIdx.x % G;
  if ((int)blockIdx.x >= nwg) {   // boundary conditions
    const int b0 = ((int)blockIdx.x - nwg) * G;
    tab_boundary<G>(prob::ri_prog(), prob::rf_prog(), b0, bt, y, p, c, smem, [&](int b, int d, double v) {
      if (FULL) rbc[(size_t)b * NRB + d] = v;
      atomicMax(norm + b, (unsigned long long)__double_as_longlong(fabs(v)));
    });
    return;
  }
  double* ext = smem;
  double* mv = ext + (size_t)prob::FG_NATOMS * G;
  double* fgv = mv + (size_t)prob::FG_NMONO * G;
  const int gid = blockIdx.x * G + g;
  int i = 0, j = 0, st = 0, b = -1;
  if (gid < nst) {
    i = listed_interval(bt, gid / M);
    j = gid - (gid / M) * M;
    st = i * M + j;
    b = bt.iinst[i];
    if (skipped(bt, b)) b = -1;
  }
  const bool valid = b >= 0;
  const int nd = i + (valid ? b : 0);
  const double h = valid ? t[nd + 1] - t[nd] : 0.0;
  const double* yi = y + (size_t)nd * NY;
  const double* yd = ydot + (size_t)i * M * NY;
  // base atoms: y-tilde_j = y_i + h sum_k A_jk ydot_k, z-tilde_j, p, c, alpha
  for (int a = threadIdx.x / G; a < NY; a += blockDim.x / G) {
    double v = 0.0;
    if (valid) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) s += c_rkA[j * M + k] * yd[k * NY + a];
      v = yi[a] + h * s;
    }
    ext[(size_t)a * G + g] = v;
  }
  tab_base_rest<double, G>(ext, g, ztil + (size_t)st * NZ, p + (size_t)(valid ? b : 0) * NP,
                           c + (size_t)(valid ? b : 0) * NC, alpha, valid);
  __syncthreads();
  const TabProg P = prob::fg_prog();
  tab_prepare<double, G>(P, ext, mv);
  tab_outputs<double, G>(P, mv, [&](int d, const double (&acc)[G]) {
#pragma unroll
    for (int q = 0; q < G; ++q) fgv[(size_t)d * G + q] = acc[q];
  });
  __syncthreads();
  double amax = 0.0;
  if (valid) {
    double* fai = fa + (size_t)st * NYZ;
    for (int l = threadIdx.x / G; l < NYZ; l += blockDim.x / G) {
      const double v = l < NY ? fgv[(size_t)l * G + g] - yd[j * NY + l] : fgv[(size_t)l * G + g];
      amax = nmax(fabs(v), amax);
      if (FULL) fai[l] = v;
    }
    if (FULL) {
      for (int l = threadIdx.x / G; l < NY; l += blockDim.x / G) ytil[(size_t)st * NY + l] = ext[(size_t)l * G + g];
    }
    if (j == M - 1) {
      for (int l = threadIdx.x / G; l < NY; l += blockDim.x / G) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) s += c_rkb[k] * yd[k * NY + l];
        const double v = y[(size_t)(nd + 1) * NY + l] - yi[l] - h * s;
        amax = nmax(fabs(v), amax);
        if (FULL) fb[(size_t)i * NY + l] = v;
      }
    }
  }
  seg_max_to(amax, valid ? b : -1, norm);
}

// ---------------------------------------------------------------- jacobian
// The state-dependent entries of the stage Jacobian blocks (the static ones
// are in place, jacobian_static), evaluated in TJ (float for the FP32
// direction), G stages per workgroup; the workgroups after the stages
// write Dr of G instances each (construct_jacobian_kernel, solver:680-797).
template <class TJ>
__device__ __forceinline__ void jacobian_tab(int nint, Batch bt, const double* __restrict__ y,
                                             const double* __restrict__ ytil, const double* __restrict__ ztil,
                                             const double* __restrict__ p, const double* __restrict__ c,
                                             double alpha, TJ* __restrict__ jac, double* __restrict__ Dr) {
  constexpr int G = TAB_G;
  constexpr int JL = (prob::JAC_NATOMS + prob::JAC_NMONO) * G;
  constexpr int JLD = (JL * (int)sizeof(TJ) + 7) / 8;
  __shared__ __attribute__((aligned(16))) double smem[TAB_MAX2(JLD, TAB_DR_LDS)];
  const int nst = launch_intervals(bt, nint) * M;
  const int nwg = (nst + G - 1) / G;
  const int g = threadIdx.x % G;
  if ((int)blockIdx.x >= nwg) {
    const int b0 = ((int)blockIdx.x - nwg) * G;
    tab_boundary<G>(prob::dri_prog(), prob::drf_prog(), b0, bt, y, p, c, smem,
                    [&](int b, int d, double v) { Dr[(size_t)b * DRSZ + d] = v; });
    return;
  }
  TJ* ext = reinterpret_cast<TJ*>(smem);
  TJ* mv = ext + (size_t)prob::JAC_NATOMS * G;
  int st = 0, b = -1;
  const int gid = blockIdx.x * G + g;
  if (gid < nst) {
    const int i = listed_interval(bt, gid / M);
    st = i * M + (gid - (gid / M) * M);
    b = bt.iinst[i];
    if (skipped(bt, b)) b = -1;
  }
  const bool valid = b >= 0;
  for (int a = threadIdx.x / G; a < NY; a += blockDim.x / G)
    ext[(size_t)a * G + g] = valid ? (TJ)ytil[(size_t)st * NY + a] : (TJ)0;
  tab_base_rest<TJ, G>(ext, g, ztil + (size_t)st * NZ, p + (size_t)(valid ? b : 0) * NP,
                       c + (size_t)(valid ? b : 0) * NC, alpha, valid);
  __syncthreads();
  // the stage index of every point, for the stores
  __shared__ long long sst[G];
  if (threadIdx.x < G) sst[threadIdx.x] = valid ? (long long)st : -1;
  const TabProg P = prob::jac_prog();
  tab_prepare<TJ, G>(P, ext, mv);
  tab_outputs<TJ, G>(P, mv, [&](int d, const TJ (&acc)[G]) {
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (sst[q] >= 0) jac[(size_t)sst[q] * JSTAGE + d] = acc[q];
  });
}

// ------------------------------------------------------- initial input
// form_initial_input_kernel (solver:316-356), G stages per workgroup.
__device__ __forceinline__ void init_input_tab(int nint, Batch bt, const double* __restrict__ y0,
                                               const double* __restrict__ z0, const double* __restrict__ p,
                                               const double* __restrict__ c, double* __restrict__ ydot,
                                               double* __restrict__ ztil) {
  constexpr int G = TAB_G;
  __shared__ __attribute__((aligned(16))) double smem[TAB_FG_LDS];
  double* ext = smem;
  double* mv = ext + (size_t)prob::FG_NATOMS * G;
  double* fgv = mv + (size_t)prob::FG_NMONO * G;
  const int g = threadIdx.x % G;
  const int gid = blockIdx.x * G + g;
  const bool valid = gid < nint * M;
  const int i = valid ? gid / M : 0, j = valid ? gid - (gid / M) * M : 0;
  const int ib = valid ? bt.iinst[i] : 0;
  const int nd = i + ib;
  const double cj = c_rkc[j];
  for (int a = threadIdx.x / G; a < NY + NZ; a += blockDim.x / G) {
    double v = 0.0;
    if (valid) {
      if (a < NY) {
        v = (1.0 - cj) * y0[(size_t)nd * NY + a] + cj * y0[(size_t)(nd + 1) * NY + a];
      } else {
        v = (1.0 - cj) * z0[(size_t)nd * NZ + a - NY] + cj * z0[(size_t)(nd + 1) * NZ + a - NY];
        ztil[(size_t)gid * NZ + a - NY] = v;
      }
    }
    ext[(size_t)a * G + g] = v;
  }
  for (int a = NY + NZ + threadIdx.x / G; a < prob::TB_NB; a += blockDim.x / G) {
    double v = 0.0;
    if (valid) {
      if (a < NY + NZ + NP) v = p[(size_t)ib * NP + a - NY - NZ];
      else if (a < NY + NZ + NP + NC) v = c[(size_t)ib * NC + a - NY - NZ - NP];
    }
    ext[(size_t)a * G + g] = v;   // alpha: the reference evaluates f only (no alpha)
  }
  __syncthreads();
  const TabProg P = prob::fg_prog();
  tab_prepare<double, G>(P, ext, mv);
  tab_outputs<double, G>(P, mv, [&](int d, const double (&acc)[G]) {
#pragma unroll
    for (int q = 0; q < G; ++q) fgv[(size_t)d * G + q] = acc[q];
  });
  __syncthreads();
  if (valid)
    for (int l = threadIdx.x / G; l < NY; l += blockDim.x / G) ydot[(size_t)gid * NY + l] = fgv[(size_t)l * G + g];
}

// ------------------------------------------------------- segment residual
// The table-form segres (see segres in colloc_kernels.h for the variants
// and the reference lines): workgroup k < Nint evaluates f and g of
// interval k at its M+1 Gauss points, one point per column of the
// workgroup's tables, and one thread accumulates them in the reference's
// order; the workgroups after the intervals take the boundary residual of
// TAB_G instances each.
constexpr int SEG_G = M + 1;
constexpr int TAB_SEG_LDS = (prob::FG_NATOMS + prob::FG_NMONO + NY + NZ + NY) * SEG_G;

__device__ __forceinline__ void segres_tab(int nint, Batch bt, const double* __restrict__ t,
                                           const double* __restrict__ y, const double* __restrict__ ydot,
                                           const double* __restrict__ ztil, const double* __restrict__ p,
                                           const double* __restrict__ c, double alpha, double tol, int rtype,
                                           int scale_time, int relative, double* __restrict__ residual,
                                           unsigned long long* __restrict__ maxres) {
  constexpr int G = SEG_G;
  __shared__ __attribute__((aligned(16))) double smem[TAB_MAX2(TAB_SEG_LDS, TAB_R_LDS)];
  if ((int)blockIdx.x >= nint) {
    const int b0 = ((int)blockIdx.x - nint) * TAB_G;
    __shared__ unsigned long long rmax[TAB_G];
    if (threadIdx.x < TAB_G) rmax[threadIdx.x] = 0;
    __syncthreads();
    tab_boundary<TAB_G>(prob::ri_prog(), prob::rf_prog(), b0, bt, y, p, c, smem, [&](int b, int d, double v) {
      atomicMax(rmax + (b - b0), (unsigned long long)__double_as_longlong(fabs(v)));
    });
    __syncthreads();
    if (threadIdx.x < TAB_G) {
      const int b = b0 + threadIdx.x;
      if (b < bt.B) {
        const double val = __longlong_as_double((long long)rmax[threadIdx.x]) / tol;
        residual[bt.noff[b + 1] - 1] = val;
        atomicMax(maxres + b, (unsigned long long)__double_as_longlong(fabs(val)));
      }
    }
    return;
  }
  double* ext = smem;
  double* mv = ext + (size_t)prob::FG_NATOMS * G;
  double* fgv = mv + (size_t)prob::FG_NMONO * G;
  double* ydt = fgv + (size_t)(NY + NZ) * G;
  const int i = blockIdx.x;
  const int ib = bt.iinst[i];
  const int nd = i + ib;
  p += (size_t)ib * NP;
  c += (size_t)ib * NC;
  const double h = t[nd + 1] - t[nd];
  const double* yi = y + (size_t)nd * NY;
  const double* yd = ydot + (size_t)i * M * NY;
  const double* zz = ztil + (size_t)i * M * NZ;
  // every thread: Gauss point gp = tid % G (the tables' column)
  const int gp = threadIdx.x % G;
  const bool col = (int)threadIdx.x < (blockDim.x / G) * G;   // threads past the last full column group idle
  if (col) {
    for (int l = threadIdx.x / G; l < NY; l += blockDim.x / G) {
      double s = 0.0, yy = yi[l];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        s += c_gL[gp * M + k] * yd[k * NY + l];
        yy += h * c_gI[gp * M + k] * yd[k * NY + l];
      }
      ydt[(size_t)l * G + gp] = s;
      ext[(size_t)l * G + gp] = yy;
    }
    for (int l = threadIdx.x / G; l < NZ; l += blockDim.x / G) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) s += c_gL[gp * M + k] * zz[k * NZ + l];
      ext[(size_t)(NY + l) * G + gp] = s;
    }
    for (int a = NY + NZ + threadIdx.x / G; a < prob::TB_NB; a += blockDim.x / G) {
      double v;
      if (a < NY + NZ + NP) v = p[a - NY - NZ];
      else if (a < NY + NZ + NP + NC) v = c[a - NY - NZ - NP];
      else v = alpha;
      ext[(size_t)a * G + gp] = v;
    }
  }
  __syncthreads();
  const TabProg P = prob::fg_prog();
  tab_prepare<double, G>(P, ext, mv);
  tab_outputs<double, G>(P, mv, [&](int d, const double (&acc)[G]) {
#pragma unroll
    for (int q = 0; q < G; ++q) fgv[(size_t)d * G + q] = acc[q];
  });
  __syncthreads();
  if (threadIdx.x != 0) return;
  // the accumulation of segres (colloc_kernels.h), in its order
  double val = 0.0, rho_h = 0.0, rho_g = 0.0;
  if (!relative) {
    for (int q = 0; q < M + 1; ++q) {
      const double w = c_gw[q];
      for (int l = 0; l < NY; ++l) {
        const double fo = fgv[(size_t)l * G + q] - ydt[(size_t)l * G + q];
        const double e = fabs(fo);
        if (rtype == 2) rho_h += w * fo * fo;
        else if (rtype == 1) rho_h += w * e;
        else rho_h = fmax(rho_h, w * e);
        if (scale_time) rho_h *= h;
      }
      if (NZ > 0) {
        for (int l = 0; l < NZ; ++l) {
          const double go = fgv[(size_t)(NY + l) * G + q];
          if (rtype == 2) rho_g += w * go * go;
          else if (rtype == 1) rho_g += w * fabs(go);
          else rho_g = fmax(rho_g, l < NY ? w * fabs(fgv[(size_t)l * G + q] - ydt[(size_t)l * G + q]) : 0.0);
          if (scale_time) rho_g *= h;
        }
      }
    }
    if (rtype == 2) val = sqrt(rho_h + rho_g) / tol;
    else if (rtype == 1) val = (fabs(rho_h) + fabs(rho_g)) / tol;
    else val = fmax(rho_h, rho_g) / tol;
  } else {
    for (int l = 0; l < NY; ++l) {
      double ydm = 0.0;
      for (int q = 0; q < M + 1; ++q) ydm = fmax(ydm, fabs(ydt[(size_t)l * G + q]));
      for (int q = 0; q < M + 1; ++q)
        rho_h = fmax(rho_h, fabs(fgv[(size_t)l * G + q] - ydt[(size_t)l * G + q]) / (1.0 + ydm));
    }
    for (int l = 0; l < NZ; ++l) {
      double zm = 0.0;
      for (int q = 0; q < M + 1; ++q) zm = fmax(zm, fabs(ext[(size_t)(NY + l) * G + q]));
      for (int q = 0; q < M + 1; ++q) rho_g = fmax(rho_g, fabs(fgv[(size_t)(NY + l) * G + q]) / (1.0 + zm));
    }
    val = fmax(rho_h, rho_g) / tol;
  }
  residual[nd] = val;
  atomicMax(maxres + ib, (unsigned long long)__double_as_longlong(fabs(val)));
}
#endif  // COLLOC_TABLE_FORM

}  // namespace colloc
