/*
 * colloc.h -- C ABI of the MI355X collocation BVP-DAE solver (libcolloc.so).
 *
 * The reference solver (UW-OCP/Collocation-Solver-CUDA) exposes its hot path
 * as a set of synchronous, numpy-in/numpy-out stage functions in
 * collocation_solver_parallel_2d_shared.py ("solver") and
 * solve_babd_system.py, plus a generated problem module whose callbacks
 * _abvp_f/_g/_Df/_Dg/_r/_Dr define the BVP-DAE (OCP2ABVP.py:203-327).
 * Each entry point below replaces one of those functions; the comment on each
 * cites the reference function (file:line) it stands in for.
 *
 * Conventions
 *  - Plain C types only; host arrays are row-major float64 in the reference's
 *    layouts: t[N]; y[N][ny]; y_dot[(N-1)*m][ny] and z_tilde[(N-1)*m][nz]
 *    with row i*m+j for stage j of interval i (solver:278-283); p[np].
 *  - Batches: a handle holds B >= 1 instances of one problem (same callbacks,
 *    own mesh, constants, state and parameters), laid end to end.  With
 *    N_b nodes per instance, the node arrays (t, y, z, the segment residual)
 *    are the instances' arrays concatenated (N = sum N_b rows), the stage
 *    arrays likewise (sum (N_b - 1) * m rows), and p, r_bc, Dr and the
 *    constants hold one block per instance.  Scalars that the reference
 *    returns per solve (residual norm, iteration count, maximum segment
 *    residual) are arrays of B values.  cs_set_mesh makes a batch of one, in
 *    which every layout is exactly the reference's.
 *  - All solver state lives in device memory owned by the handle; host
 *    arrays are copied in and out on the handle's stream, and every call
 *    returns after the stream has drained.
 *  - Return value 0 = success, otherwise one of the CS_E_* codes below;
 *    cs_last_error() gives a message for the calling thread.
 *  - A handle is not re-entrant: one host thread per handle.
 *  - The problem-specific device code (the generated callbacks plus all
 *    kernels specialised to the problem's sizes) is a HIP source text that
 *    the caller generates; cs_jit_compile() compiles it with hiprtc and
 *    cs_load_module() loads the resulting gfx950 code object.
 */
#ifndef COLLOC_H
#define COLLOC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_OK 0
#define CS_E_HIP 1          /* a HIP runtime call failed                     */
#define CS_E_SHAPE 2        /* sizes do not match the handle / module          */
#define CS_E_SINGULAR 3     /* a zero or non-finite pivot was met              */
#define CS_E_ARG 4          /* invalid argument                               */
#define CS_E_COMPILE 5      /* hiprtc compilation failed                      */
#define CS_E_NOMODULE 6     /* no problem module loaded                       */
#define CS_E_UNSUPPORTED 7  /* shape outside what the kernels support          */

typedef struct cs_handle cs_handle;

typedef struct {
  int ny;      /* size_y  (OCP2ABVP.py:344)                         */
  int nz;      /* size_z                                            */
  int np;      /* size_p                                            */
  int m;       /* collocation stages (solver -m flag, solver:2381)   */
  int n_max;   /* largest number of time nodes N the handle holds
                  (summed over the instances of a batch)              */
  int nc;      /* number of per-instance constants (0 if none)        */
  int batch;   /* most instances of a batch (0 or 1: single instance) */
} cs_dims;

/* Error text of the last failed call on this thread. */
const char* cs_last_error(void);

/* Library version string. */
const char* cs_version(void);

/* Compile a HIP source text for gfx950 with hiprtc.  include_dir is added to
 * the include path.  On success *code (malloc'ed, free with cs_free) holds the
 * code object of *size bytes.  *log (may be NULL) receives the compiler log
 * (malloc'ed).  Stands in for Numba's JIT of the generated bvp_problem module
 * and the @cuda.jit kernels (solver:10-17). */
int cs_jit_compile(const char* source, const char* name, const char* include_dir,
                   void** code, size_t* size, char** log);
void cs_free(void* ptr);

/* Create a handle owning all device workspace for problems of the given
 * sizes, and one HIP stream.  Replaces the per-call cuda.device_array
 * allocations of every stage wrapper (e.g. solver:294-307, 411-445).  Any
 * ny >= 1: modules of problems with ny <= 64 carry the register-resident
 * condensation and cyclic reduction, larger ones the large-block kernels
 * (float64 cyclic reduction, host-driven Newton loop; an interval whose
 * series condensation does not converge is reported as CS_E_SINGULAR). */
int cs_create(const cs_dims* dims, cs_handle** out);
void cs_destroy(cs_handle* h);

/* Load a code object produced by cs_jit_compile for this handle's sizes. */
int cs_load_module(cs_handle* h, const void* code, size_t size);

/* Time mesh (solver:53-55 t_span) of a single instance.  2 <= N <= n_max. */
int cs_set_mesh(cs_handle* h, const double* t, int N);

/* Meshes of a batch of B instances (1 <= B <= dims.batch): instance b has
 * nodes[b] >= 2 time nodes; t holds the meshes concatenated.  The batched
 * counterpart of the per-solve t_span (solver:53-55); the reference solves
 * one instance per process.  Invalidates the saved initial input. */
int cs_set_batch_mesh(cs_handle* h, int B, const int* nodes, const double* t);

/* Per-instance constants of the generated problem (nc values), given to
 * every instance of the batch. */
int cs_set_constants(cs_handle* h, const double* c, int nc);

/* Constants of each instance of a batch: c[B][nc]. */
int cs_set_batch_constants(cs_handle* h, const double* c, int B, int nc);

/* Instances and total nodes of the current mesh. */
int cs_batch_info(cs_handle* h, int* B, int* N);

/* Collocation state: y[N][ny], y_dot, z_tilde, p[B][np] (solver:101 outputs). */
int cs_set_state(cs_handle* h, const double* y, const double* y_dot,
                 const double* z_tilde, const double* p);
int cs_get_state(cs_handle* h, double* y, double* y_dot, double* z_tilde, double* p);

/* form_initial_input_parallel (solver:287-356): y = y0, p = p0, stage values
 * interpolated linearly from node values and y_dot = f(y_stage, z_tilde, p). */
int cs_form_initial_input(cs_handle* h, const double* y0, const double* z0, const double* p0);

/* cs_form_initial_input with y0, z0 and p0 in device memory (for example
 * an initial guess computed on the GPU, colloc_mlp.h).  The handle's
 * stream waits for nothing else: the caller makes sure the producer of the
 * arrays has finished (e.g. hipStreamSynchronize / hipDeviceSynchronize). */
int cs_form_initial_input_device(cs_handle* h, const double* y0, const double* z0, const double* p0);

/* Re-form the collocation state from the node values, mesh and parameters
 * last passed to cs_form_initial_input (kept in device memory), without any
 * host transfer.  Same result as repeating that call; used to restart a
 * solve from its initial guess (e.g. repeated or batched solves). */
int cs_restart(cs_handle* h);

/* compute_f_q_parallel (solver:399-466): residual of the current state.
 * norm_inf[b] = ||f_q||_inf of instance b (the host reduction of solver:116).
 * Any of the output arrays may be NULL; otherwise y_tilde[(N-B)*m][ny],
 * f_a[N-B][m*(ny+nz)], f_b[N-B][ny], r_bc[B][ny+np]. */
int cs_compute_f_q(cs_handle* h, double alpha, double* norm_inf, double* y_tilde,
                   double* f_a, double* f_b, double* r_bc);

/* cs_compute_f_q that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  It evaluates the same
 * residual into the handle's f_a / f_b / y_tilde / r_bc buffers and writes
 * norm_inf[b] (host array of B values), with the per-instance max-norm slots
 * in a fixed 64 KiB device array of the library instead of the handle's
 * workspace: at most 8192 instances (a larger batch is rejected with
 * CS_E_SHAPE).  All handles on one GPU share that array, which keeps the last
 * call's contents (each call clears the slots it uses first): calls that
 * overlap on one GPU, from several threads, race. */
int cs_compute_f_q_noalloc(cs_handle* h, double alpha, double* norm_inf);

/* construct_jacobian_parallel (solver:611-797): stage Jacobians at the
 * stage values of the last cs_compute_f_q call.  Kept on the device; when
 * non-NULL, Df[(N-B)*m][ny][ny+nz+np], Dg[(N-B)*m][nz][ny+nz+np] and
 * Dr[B][ny+np][2*ny+np] (= [B_0 | B_n | V_n], solver:655-665) are copied out. */
int cs_construct_jacobian(cs_handle* h, double alpha, double* Df, double* Dg, double* Dr);

/* reduce_jacobian_parallel (solver:840-1001): per-interval condensation to
 * A_i dy_i + C_i dy_{i+1} + H_i dp = b_i with C_i = I.  Outputs (may be
 * NULL): A[N-1][ny][ny], H[N-1][ny][np], b[N-1][ny]. */
int cs_reduce_jacobian(cs_handle* h, double* A, double* H, double* b);

/* partition_factorization_parallel + construct_babd_mshoot + qr_decomposition
 * + backward_substitution + recover_babd_solution +
 * partition_backward_substitution_parallel (solve_babd_system.py:93-572,
 * driven at solver:134-147).  Outputs (may be NULL): delta_y[N][ny],
 * delta_p[B][np]. */
int cs_solve_babd(cs_handle* h, double* delta_y, double* delta_p);

/* Arithmetic of the Newton direction's condensation (reduce_jacobian,
 * solver:840-1001): 64 (default; float64 throughout, as the reference) or 32
 * (the stage Jacobians evaluated and stored in float32, and the
 * stage eliminations and the series solve of the stage systems in float32:
 * an inexact-Newton direction).  Residuals, the BABD solve, the recovery,
 * the line-search merit and the convergence test stay float64; intervals
 * whose float32 series does not converge are condensed in float64.  Takes
 * effect at the next Jacobian; cs_construct_jacobian returns the stored
 * Jacobians widened to float64.  No counterpart in the reference. */
int cs_set_direction_precision(cs_handle* h, int bits);

/* Newton loops (cs_newton, cs_newton_progress, cs_newton_stall) run device
 * resident when on != 0: the per-instance decisions of solver:109-174 (the
 * convergence test, the progress / stall tests, the line search's step
 * choice and acceptance) are made by kernels, the launch list of the
 * iterating instances' intervals is built on the device, and each iteration
 * is a replay of HIP graphs captured once per mesh and settings (the host
 * reads three counters per iteration instead of synchronising for every
 * residual).  The results are those of the host-driven loop (on = 0, the
 * default).  CS_E_UNSUPPORTED for a module of large-block kernels (ny > 64).
 * No counterpart in the reference. */
int cs_set_newton_graph(cs_handle* h, int on);

/* Iterative refinement of the Newton direction: after the direction delta
 * (any precision), `steps` times: rho = F + J delta in float64 with the
 * stored stage Jacobians and Dr, the correction e from the same direction
 * kernels with rho as the right-hand side (J e = -rho), delta <- delta + e.
 * For the FP32 direction this takes delta towards the float64 solution of
 * the stored system.  0 (the default) to 8.  The north star's "FP32
 * direction with iterative refinement"; no counterpart in the reference. */
int cs_set_refinement(cs_handle* h, int steps);

/* Stored type of the cyclic reduction's rows and records (the partitioned
 * QR of solve_babd_system.py:93-258 and its back substitution,
 * :495-572): 64 (default) or 32.  With 32 every level after the first
 * stores float32 rows and records and the substitution pairing computes in
 * float32; the Householder pairing, the boundary-block solve (the reduced
 * system of qr_decomposition, :338-440) and the back substitution compute
 * in float64 (with large-block kernels, ny > 64, the substitution products
 * run on float32 matrix cores).  Part of the FP32 direction (with
 * cs_set_direction_precision and cs_set_refinement).  No counterpart in the
 * reference. */
int cs_set_babd_precision(cs_handle* h, int bits);

/* max |rho| per instance (B doubles) of rho = F + J delta, the linearised
 * collocation system (the stages of compute_f_q_parallel, solver:478-559,
 * and r at the boundary) at the current state, for the direction the last
 * solve_babd / recover_delta_k left, with the Jacobians of the last
 * construct_jacobian and f_a, f_b, r_bc of the last compute_f_q.  Nothing
 * is overwritten.  No counterpart in the reference (a check of the
 * direction's accuracy). */
int cs_linear_residual(cs_handle* h, double* norm_out);

/* `steps` refinement steps (as cs_set_refinement) on the current direction
 * of the stage functions (after solve_babd / recover_delta_k, with the
 * f_a, f_b, r_bc and Jacobians it was computed from).  f_a, f_b and r_bc
 * are left holding the last linear residual: call compute_f_q again before
 * using them.  No counterpart in the reference. */
int cs_refine_direction(cs_handle* h, int steps);

/* The current Newton direction as it stands (after solve_babd,
 * recover_delta_k or cs_refine_direction), without recomputing any part:
 * delta_y [N][ny], delta_y_dot and delta_z_tilde [Nint*m][.], delta_p
 * [B][np]; null pointers skip a part.  No counterpart in the reference. */
int cs_get_direction(cs_handle* h, double* delta_y, double* delta_y_dot, double* delta_z_tilde, double* delta_p);

/* Kernel of the series condensation in the FP32 direction: 0 (default) the
 * products of the series on matrix cores (v_mfma_f32_16x16x4f32, one wave
 * per stage block) where the problem's stage blocks fit (ny <= 48 and
 * ny + active parameters + 1 <= 48, up to 16 stage blocks), 1 the vector-ALU
 * series kernel.  Both solve the same stage systems of reduce_jacobian
 * (solver:887-1001) to the same tolerance.  No counterpart in the reference. */
int cs_set_condense_kernel(cs_handle* h, int kind);

/* Launch shapes and kernel choices (no counterpart in the reference; the
 * first two chosen by the autotuner, autotune.py, and results identical
 * for every value):
 *   CS_LAUNCH_CONDENSE_GRID: 0 (default) launches the series condensation
 *     with one workgroup per interval; n > 0 launches a persistent grid of
 *     n workgroups that take the intervals of the launch list from a work
 *     queue (a ticket counter), so a captured Newton iteration needs no grid
 *     sized for every interval;
 *   CS_LAUNCH_RESIDUAL_BLOCK: workgroup size (64, 128, 256) of the residual
 *     kernel of straight-line callbacks (one thread per stage);
 *   CS_LAUNCH_COMPUTE_UNITS (get only): the device's compute units;
 *   CS_LAUNCH_JACOBIAN_GEMM: 1 (default) evaluates the float32 stage
 *     Jacobians (the FP32 direction) on matrix cores where the module has
 *     that kernel (a straight-line unit whose problem also has the
 *     polynomial form: the entries as a constant coefficient matrix times
 *     the monomials of the stage), 0 the straight-line kernel; the two agree
 *     to float32 rounding.  Get: 1 when the matrix-core kernel is in use. */
#define CS_LAUNCH_CONDENSE_GRID 0
#define CS_LAUNCH_RESIDUAL_BLOCK 1
#define CS_LAUNCH_COMPUTE_UNITS 2
#define CS_LAUNCH_JACOBIAN_GEMM 3
int cs_set_launch(cs_handle* h, int which, int value);
int cs_get_launch(cs_handle* h, int which, int* value);

/* Pairing rule of the BABD solve's cyclic reduction.  Every block row that
 * the condensation produces has C_i = I; while a pair of rows spans at most
 * 1/min_spans of its instance's intervals, the middle node is eliminated by
 * substitution (x_M = b_a - G_a x_L - H_a p, one small matrix product per
 * pair, and C stays I), and on the coarser levels by the Householder step of
 * partition_factorization_kernel (solve_babd_system.py:184-246).
 * min_spans = 0: Householder on every level.  Default 64.  The reference
 * always uses Householder QR. */
int cs_set_babd_substitution(cs_handle* h, int min_spans);

/* recover_delta_k_parallel (solver:1580-1650).  Outputs (may be NULL):
 * delta_y_dot[(N-1)*m][ny], delta_z_tilde[(N-1)*m][nz]. */
int cs_recover_delta_k(cs_handle* h, double* delta_y_dot, double* delta_z_tilde);

/* The damped Newton loop with line search (solver:109-174), for every
 * instance of the batch at once: instance b iterates until its own
 * ||f_q||_inf < tol or max_iter iterations, with its own line search.
 * iter_time[b] receives its loop counter at exit (as the reference's
 * iter_time), norm[b] its last residual norm, norms[b][0..hist_len) its norm
 * history (may be NULL; entries past its last iteration are not written).
 * An instance whose line search accepts no step keeps its state, so each of
 * its later iterations would repeat that one: it stops at once with the
 * result they would give (iter_time = max_iter - 1, the same norm, the
 * history continued at that norm). */
int cs_newton(cs_handle* h, double alpha, double tol, int max_iter, int* iter_time,
              double* norm, double* norms, int hist_len);

/* cs_newton with a progress test, for Newton solves started from a guess
 * that should already be in the quadratic basin (a learned warm start): an
 * instance whose ||f_q||_inf, from iteration `patience` on, falls by less
 * than a factor min_ratio (>= 1) in one iteration stops there, not
 * converged, and iter_time[b] is that iteration.  The caller falls back to
 * its plain solve for it.  No counterpart in the reference, which iterates to
 * max_iter. */
int cs_newton_progress(cs_handle* h, double alpha, double tol, int max_iter, int patience, double min_ratio,
                       int* iter_time, double* norm, double* norms, int hist_len);

/* cs_newton with a stall test over a window, for the continuation stages
 * of a fallback solve: from iteration `start` on, an instance whose
 * ||f_q||_inf at iteration k is not below its norm at iteration k - window
 * divided by min_ratio (>= 1) stops there, not converged, and iter_time[b]
 * is k.  cs_newton_progress is (start = patience, window = 1).  Damped
 * iterations that are still descending pass; instances on a plateau stop
 * instead of running to max_iter.  No counterpart in the reference. */
int cs_newton_stall(cs_handle* h, double alpha, double tol, int max_iter, int start, int window, double min_ratio,
                    int* iter_time, double* norm, double* norms, int hist_len);

/* compute_segment_residual_parallel (solver:1770-1892) with residual_type 1
 * and scale_by_time.  residual[N] (may be NULL; the last node of each
 * instance holds its ||r||_inf / tol), max_residual[B]. */
int cs_segment_residual(cs_handle* h, double alpha, double tol, double* residual,
                        double* max_residual);

/* The other acceptance residuals of the reference, selected as its module
 * flags select them (solver:30-33): relative = 0 is
 * compute_segment_residual_parallel (solver:1770-1892) with residual_type
 * 0 (max), 1 (sum of |.|) or 2 (sum of squares, then sqrt) and the
 * scale_by_time flag; relative = 1 is compute_segment_residual_parallel2
 * (solver:1895-2013, the residual2 flag), where residual_type and
 * scale_by_time are ignored.  cs_segment_residual is (1, 1, 0). */
int cs_segment_residual_ex(cs_handle* h, double alpha, double tol, int residual_type, int scale_by_time,
                           int relative, double* residual, double* max_residual);

/* recover_solution_parallel (solver:2034-2075): z at the nodes, z[N][nz]. */
int cs_recover_solution(cs_handle* h, double* z);

/* The current solution at the nodes without a host copy: forms z at the
 * nodes (as cs_recover_solution) in the handle's buffer and returns device
 * pointers to y[N][ny], z[N][nz] and p[B][np] (any may be NULL).  The
 * pointers are borrowed: they stay valid while the handle lives, and the
 * values until the next call that changes the state or the mesh.  For
 * device-side producers of initial guesses (colloc_mlp.h); the reference
 * returns numpy copies only. */
int cs_node_solution_device(cs_handle* h, const double** y, const double** z, const double** p);

/* Kernel timing with HIP events on the handle's stream.  cs_profile(h, 1)
 * starts recording every launch of the kernels listed in CS_KERNEL_*;
 * cs_profile_get returns the summed milliseconds and launch count. */
#define CS_KERNEL_RESIDUAL 0
#define CS_KERNEL_JACOBIAN 1
#define CS_KERNEL_CONDENSE 2
#define CS_KERNEL_BABD 3
#define CS_KERNEL_RECOVER 4
#define CS_KERNEL_COUNT 5
int cs_profile(cs_handle* h, int enable);
int cs_profile_get(cs_handle* h, int kernel, double* total_ms, int* launches);

/* Condensation (reduce_jacobian) runs a series solve of each interval's
 * stage system and falls back to Gauss-Jordan elimination for intervals
 * where the series does not converge (coarse meshes).  Since the handle was
 * created: *fallback_intervals = intervals condensed by the fallback,
 * *series_terms = series terms summed over all intervals, *intervals =
 * intervals condensed (instances that had already converged are skipped).
 * Any pointer may be NULL.  No counterpart in the reference (it always
 * factors W). */
int cs_condense_stats(cs_handle* h, long long* fallback_intervals, long long* series_terms,
                      long long* intervals);

/* Device status word (bit 0: singular stage block, bit 1: singular stage
 * system, bit 2: singular boundary system); cleared by reading. */
int cs_status(cs_handle* h, int* status);

#ifdef __cplusplus
}
#endif

#endif /* COLLOC_H */
