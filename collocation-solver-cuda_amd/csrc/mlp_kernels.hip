// libcolloc_mlp.so -- kernels of the learned warm start (gfx950).
//
// A multilayer perceptron maps the parameters of an OCP instance to its
// solution on a fixed grid of normalised times.  Its weights and
// activations are bf16, products accumulate in fp32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16), the optimiser keeps fp32 master weights.
// Every entry point is asynchronous on the stream it is given and takes
// plain device pointers; the C ABI is declared in include/colloc_mlp.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "colloc_mlp.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define LAUNCHCHK()                                                                          \
  do {                                                                                       \
    hipError_t _e = hipGetLastError();                                                       \
    if (_e != hipSuccess) return fail(CM_E_HIP, "launch failed: %s", hipGetErrorString(_e)); \
  } while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 with round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// ------------------------------------------------------------------ GEMM
// C[m, n] = sum_k A[m, k] B[n, k] for row-major bf16 A (M x K) and B (N x K),
// fp32 accumulation; K is a multiple of 32.  A 256-thread workgroup owns a
// 64 x 64 tile of C; each of its four waves a 32 x 32 quarter, as 2 x 2
// blocks of 16 x 16 computed by v_mfma_f32_16x16x32_bf16.  A and B tiles of
// 64 x 32 go through LDS, rows padded to 40 bf16 (80 bytes) so that the
// 16-byte fragment reads of 16 consecutive rows fall in different banks.
// MFMA 16x16x32 operands: lane l holds row (l mod 16) of A / column
// (l mod 16) of B at k = 8 (l div 16) .. +7; the result holds column
// (l mod 16), rows 4 (l div 16) .. +3.
constexpr int BM = 64, BN = 64, BK = 32, LDSROW = BK + 8;

__global__ __launch_bounds__(256) void mlp_gemm_nt(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ B, int ldb,
                                                   const float* __restrict__ bias, const float* __restrict__ aux,
                                                   int ldaux, int act, float* __restrict__ Cf, int ldcf,
                                                   uint16_t* __restrict__ Cb, int ldcb) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[BM * LDSROW];
  __shared__ __attribute__((aligned(16))) uint16_t sB[BN * LDSROW];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // global -> LDS: thread t moves 8 bf16 of row t / 4 at k offset 8 (t mod 4)
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  for (int k0 = 0; k0 < K; k0 += BK) {
    const int gm = m0 + lr, gn = n0 + lr;
    const uint4 va = gm < M ? *reinterpret_cast<const uint4*>(A + (size_t)gm * lda + k0 + lk) : zero4;
    const uint4 vb = gn < N ? *reinterpret_cast<const uint4*>(B + (size_t)gn * ldb + k0 + lk) : zero4;
    *reinterpret_cast<uint4*>(sA + lr * LDSROW + lk) = va;
    *reinterpret_cast<uint4*>(sB + lr * LDSROW + lk) = vb;
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    bf16x8 fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint4 a = *reinterpret_cast<const uint4*>(sA + (wm + i * 16 + fr) * LDSROW + fk);
      const uint4 b = *reinterpret_cast<const uint4*>(sB + (wn + i * 16 + fr) * LDSROW + fk);
      fa[i] = __builtin_bit_cast(bf16x8, a);
      fb[i] = __builtin_bit_cast(bf16x8, b);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // epilogue
  const int cn = lane & 15, cr = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + j * 16 + cn;
      if (n >= N) continue;
      const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + cr + r;
        if (m >= M) continue;
        float v = acc[i][j][r] + bn;
        if (act == CM_ACT_GELU_BWD) v *= gelu_grad(aux[(size_t)m * ldaux + n]);
        if (Cf) Cf[(size_t)m * ldcf + n] = v;
        if (Cb) Cb[(size_t)m * ldcb + n] = f2bf(act == CM_ACT_GELU ? gelu(v) : v);
      }
    }
}

// out[c, r] = in[r, c] (bf16), through a 32 x 32 LDS tile.
__global__ __launch_bounds__(256) void mlp_transpose(int R, int C, const uint16_t* __restrict__ in, int ldi,
                                                     uint16_t* __restrict__ out, int ldo) {
  __shared__ uint16_t tile[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    tile[y][tx] = (r < R && c < C) ? in[(size_t)r * ldi + c] : (uint16_t)0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int c = c0 + y, r = r0 + tx;
    if (c < C && r < R) out[(size_t)c * ldo + r] = tile[tx][y];
  }
}

// out[n] = sum_m X[m, n] (fp32), one thread per column.
__global__ __launch_bounds__(256) void mlp_col_sum(int M, int N, const float* __restrict__ X, int ldx,
                                                   float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += X[(size_t)m * ldx + n];
  out[n] = s;
}

// Mean-squared-error gradient over the first n_real columns:
// d = Y - T, loss += sum d^2, dY = scale * 2 d (0 in the padding columns).
__global__ __launch_bounds__(256) void mlp_mse_grad(int M, int N, int n_real, const float* __restrict__ Y, int ldy,
                                                    const float* __restrict__ T, int ldt, float scale,
                                                    float* __restrict__ dYf, uint16_t* __restrict__ dYb, int ldd,
                                                    double* __restrict__ loss) {
  __shared__ double red[256];
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (e < (long long)M * N) {
    const int m = (int)(e / N), n = (int)(e - (e / N) * N);
    float g = 0.0f;
    if (n < n_real) {
      const float d = Y[(size_t)m * ldy + n] - T[(size_t)m * ldt + n];
      sq = (double)d * d;
      g = scale * 2.0f * d;
    }
    if (dYf) dYf[(size_t)m * ldd + n] = g;
    if (dYb) dYb[(size_t)m * ldd + n] = f2bf(g);
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

// Adam on fp32 master weights w[rows][cols] with gradient g; refreshes the
// bf16 copy wb (row-major) and, when given, its transpose wbt[cols][rows].
__global__ __launch_bounds__(256) void mlp_adam(long long n, float* __restrict__ w, const float* __restrict__ g,
                                                float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                                                float b2, float eps, float bc1, float bc2,
                                                uint16_t* __restrict__ wb, uint16_t* __restrict__ wbt, int rows,
                                                int cols) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float a = b1 * m1[e] + (1.0f - b1) * gi;
  const float b = b2 * m2[e] + (1.0f - b2) * gi * gi;
  m1[e] = a;
  m2[e] = b;
  const float wn = w[e] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  w[e] = wn;
  const uint16_t h = f2bf(wn);
  if (wb) wb[e] = h;
  if (wbt) {
    const int r = (int)(e / cols), c = (int)(e - (e / cols) * cols);
    wbt[(size_t)c * rows + r] = h;
  }
}

__global__ __launch_bounds__(256) void mlp_cast_bf16(long long n, const float* __restrict__ x,
                                                     uint16_t* __restrict__ y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e]);
}

// Solution grid -> mesh: G[b] holds, per grid node k < ng (normalised time
// k / (ng - 1)) and component c < nc, the network's standardised output
// G[b][k * nc + c] (row stride ldg); value = mean[k nc + c] + sd[k nc + c] G.
// Node n of instance b (nodes noff[b] .. noff[b+1]-1, times t) gets the
// cubic Lagrange interpolant through the four grid nodes around its
// normalised time; components c < ny go to Y[n][c], the rest to Z[n][c-ny],
// each added to the base value Yb / Zb of the node when those are given.
__global__ __launch_bounds__(256) void mlp_grid_to_mesh(int B, const int* __restrict__ noff,
                                                        const double* __restrict__ t, int ng, int nc, int ny,
                                                        const float* __restrict__ G, int ldg,
                                                        const float* __restrict__ mean, const float* __restrict__ sd,
                                                        const double* Yb, const double* Zb, double* Y,
                                                        double* Z) {   // Yb may alias Y, Zb Z
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= noff[B]) return;
  int lo = 0, hi = B - 1;   // instance of node n: noff[b] <= n < noff[b+1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (noff[mid] <= n) lo = mid;
    else hi = mid - 1;
  }
  const int b = lo;
  const double t0 = t[noff[b]], t1 = t[noff[b + 1] - 1];
  const double s = (t[n] - t0) / (t1 - t0) * (ng - 1);
  int k0 = (int)floor(s) - 1;
  k0 = k0 < 0 ? 0 : (k0 > ng - 4 ? ng - 4 : k0);
  double w[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double v = 1.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q != a) v *= (s - (k0 + q)) / (double)(a - q);
    w[a] = v;
  }
  const float* g = G + (size_t)b * ldg;
  const int nz = nc - ny;
  for (int c = 0; c < nc; ++c) {
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int idx = (k0 + a) * nc + c;
      v += w[a] * ((double)mean[idx] + (double)sd[idx] * (double)g[idx]);
    }
    if (c < ny) Y[(size_t)n * ny + c] = v + (Yb ? Yb[(size_t)n * ny + c] : 0.0);
    else Z[(size_t)n * nz + (c - ny)] = v + (Zb ? Zb[(size_t)n * nz + (c - ny)] : 0.0);
  }
}

inline unsigned blocks(long long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

extern "C" {

const char* cm_last_error(void) { return g_err.c_str(); }

int cm_gemm_nt(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb, const float* bias,
               const float* aux, int ldaux, int act, float* Cf, int ldcf, uint16_t* Cb, int ldcb, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || (!Cf && !Cb)) return fail(CM_E_ARG, "bad gemm arguments");
  if (K % 32 || lda % 8 || ldb % 8 || lda < K || ldb < K)
    return fail(CM_E_ARG, "gemm: K (%d) must be a multiple of 32, lda/ldb multiples of 8 and >= K", K);
  if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15)
    return fail(CM_E_ARG, "gemm: A and B must be 16-byte aligned");
  if (act < CM_ACT_NONE || act > CM_ACT_GELU_BWD) return fail(CM_E_ARG, "gemm: bad act %d", act);
  if (act == CM_ACT_GELU_BWD && !aux) return fail(CM_E_ARG, "gemm: GELU backward needs aux");
  dim3 grid(blocks(N, BN), blocks(M, BM));
  hipLaunchKernelGGL(mlp_gemm_nt, grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb, bias, aux,
                     ldaux, act, Cf, ldcf, Cb, ldcb);
  LAUNCHCHK();
  return CM_OK;
}

int cm_transpose_bf16(int R, int C, const uint16_t* in, int ldi, uint16_t* out, int ldo, void* stream) {
  if (R <= 0 || C <= 0 || !in || !out || ldi < C || ldo < R) return fail(CM_E_ARG, "bad transpose arguments");
  dim3 grid(blocks(C, 32), blocks(R, 32));
  hipLaunchKernelGGL(mlp_transpose, grid, dim3(256), 0, (hipStream_t)stream, R, C, in, ldi, out, ldo);
  LAUNCHCHK();
  return CM_OK;
}

int cm_col_sum(int M, int N, const float* X, int ldx, float* out, void* stream) {
  if (M <= 0 || N <= 0 || !X || !out || ldx < N) return fail(CM_E_ARG, "bad col_sum arguments");
  hipLaunchKernelGGL(mlp_col_sum, dim3(blocks(N, 256)), dim3(256), 0, (hipStream_t)stream, M, N, X, ldx, out);
  LAUNCHCHK();
  return CM_OK;
}

int cm_mse_grad(int M, int N, int n_real, const float* Y, int ldy, const float* T, int ldt, float scale, float* dYf,
                uint16_t* dYb, int ldd, double* loss, void* stream) {
  if (M <= 0 || N <= 0 || n_real < 0 || n_real > N || !Y || !T || ldy < N || ldt < n_real || ldd < N)
    return fail(CM_E_ARG, "bad mse_grad arguments");
  hipLaunchKernelGGL(mlp_mse_grad, dim3(blocks((long long)M * N, 256)), dim3(256), 0, (hipStream_t)stream, M, N,
                     n_real, Y, ldy, T, ldt, scale, dYf, dYb, ldd, loss);
  LAUNCHCHK();
  return CM_OK;
}

int cm_adam(long long n, float* w, const float* g, float* m1, float* m2, float lr, float b1, float b2, float eps,
            int step, uint16_t* wb, uint16_t* wbt, int rows, int cols, void* stream) {
  if (n <= 0 || !w || !g || !m1 || !m2 || step < 1) return fail(CM_E_ARG, "bad adam arguments");
  if (wbt && (long long)rows * cols != n) return fail(CM_E_ARG, "adam: rows x cols != n");
  const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(mlp_adam, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n, w, g, m1, m2, lr, b1, b2,
                     eps, bc1, bc2, wb, wbt, rows, cols);
  LAUNCHCHK();
  return CM_OK;
}

int cm_cast_bf16(long long n, const float* x, uint16_t* y, void* stream) {
  if (n <= 0 || !x || !y) return fail(CM_E_ARG, "bad cast arguments");
  hipLaunchKernelGGL(mlp_cast_bf16, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n, x, y);
  LAUNCHCHK();
  return CM_OK;
}

int cm_grid_to_mesh(int B, const int* noff, int n_nodes, const double* t, int ng, int nc, int ny, const float* G,
                    int ldg, const float* mean, const float* sd, const double* Yb, const double* Zb, double* Y,
                    double* Z, void* stream) {
  if (B <= 0 || n_nodes <= 0 || ng < 4 || nc <= 0 || ny < 0 || ny > nc || ldg < ng * nc || !noff || !t || !G ||
      !mean || !sd || !Y || (ny < nc && !Z))
    return fail(CM_E_ARG, "bad grid_to_mesh arguments");
  hipLaunchKernelGGL(mlp_grid_to_mesh, dim3(blocks(n_nodes, 256)), dim3(256), 0, (hipStream_t)stream, B, noff, t, ng,
                     nc, ny, G, ldg, mean, sd, Yb, Zb, Y, Z);
  LAUNCHCHK();
  return CM_OK;
}

}  // extern "C"
