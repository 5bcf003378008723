/*
 * colloc_mlp.h -- C ABI of the learned warm start (libcolloc_mlp.so).
 *
 * The reference has no learned component: its Newton solve always starts
 * from the problem's initial guess (all ones, or an InputFile / estimate,
 * OCP2ABVP.py:355-447).  These entry points train and evaluate a multilayer
 * perceptron that maps instance parameters to a solution on a fixed grid of
 * normalised times, and interpolate that solution onto a solver mesh, as the
 * initial guess handed to cs_form_initial_input (colloc.h).
 *
 * Conventions
 *  - All pointers are device pointers; bf16 arrays are uint16_t bit
 *    patterns (round to nearest even from fp32); matrices are row-major
 *    with the given leading dimensions (in elements).
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = the
 *    default stream) and returns 0 on success or a CM_E_* code;
 *    cm_last_error() describes the last failure of the calling thread.
 */
#ifndef COLLOC_MLP_H
#define COLLOC_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CM_OK 0
#define CM_E_HIP 1
#define CM_E_ARG 2

#define CM_ACT_NONE 0      /* v = acc + bias                                   */
#define CM_ACT_GELU 1      /* Cf = acc + bias (pre-activation), Cb = gelu(Cf)  */
#define CM_ACT_GELU_BWD 2  /* v = (acc + bias) * gelu'(aux)                    */

const char* cm_last_error(void);

/* acc[m][n] = sum_k A[m][k] B[n][k] with bf16 A (M x K), B (N x K) and fp32
 * accumulation on the matrix cores; K a multiple of 32, lda / ldb multiples
 * of 8, A and B 16-byte aligned.  v = acc + bias[n] (bias may be NULL),
 * then per act (above); Cf[m][n] = v (fp32, may be NULL) and Cb[m][n] =
 * bf16(v, or gelu(v) for CM_ACT_GELU) (may be NULL). */
int cm_gemm_nt(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
               const float* bias, const float* aux, int ldaux, int act,
               float* Cf, int ldcf, uint16_t* Cb, int ldcb, void* stream);

/* out[c][r] = in[r][c] for an R x C bf16 matrix. */
int cm_transpose_bf16(int R, int C, const uint16_t* in, int ldi, uint16_t* out, int ldo, void* stream);

/* out[n] = sum_m X[m][n] (fp32; bias gradients). */
int cm_col_sum(int M, int N, const float* X, int ldx, float* out, void* stream);

/* Squared-error gradient over the first n_real of N columns: d = Y - T,
 * *loss += sum d^2 (double, may be NULL), dY = scale * 2 d, 0 in the
 * columns n_real..N-1; written as fp32 dYf and / or bf16 dYb. */
int cm_mse_grad(int M, int N, int n_real, const float* Y, int ldy, const float* T, int ldt, float scale,
                float* dYf, uint16_t* dYb, int ldd, double* loss, void* stream);

/* One Adam step (bias-corrected for step >= 1) on n fp32 master weights w
 * with gradient g and moments m1, m2; refreshes the bf16 copy wb (may be
 * NULL) and the transposed bf16 copy wbt[cols][rows] of w[rows][cols] (may
 * be NULL). */
int cm_adam(long long n, float* w, const float* g, float* m1, float* m2, float lr, float b1, float b2,
            float eps, int step, uint16_t* wb, uint16_t* wbt, int rows, int cols, void* stream);

/* y[i] = bf16(x[i]). */
int cm_cast_bf16(long long n, const float* x, uint16_t* y, void* stream);

/* Interpolate per-instance grid solutions onto meshes.  G[b] (row stride
 * ldg) holds standardised values at ng grid nodes of normalised time
 * k / (ng - 1), nc components each (node-major: k * nc + c), un-standardised
 * as mean + sd * G.  Node n of instance b (noff[b] <= n < noff[b+1], times
 * t[n], n_nodes = noff[B]) gets the cubic Lagrange interpolant through the
 * four nearest grid nodes: Y[n][c] for c < ny, Z[n][c - ny] otherwise
 * (fp64, the layouts of cs_form_initial_input's y0 and z0), added to the
 * base values Yb[n][c] / Zb[n][c - ny] when those are given (may be NULL;
 * Yb may be Y and Zb may be Z). */
int cm_grid_to_mesh(int B, const int* noff, int n_nodes, const double* t, int ng, int nc, int ny,
                    const float* G, int ldg, const float* mean, const float* sd, const double* Yb,
                    const double* Zb, double* Y, double* Z, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* COLLOC_MLP_H */
