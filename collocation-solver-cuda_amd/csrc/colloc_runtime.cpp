// libcolloc.so -- host runtime of the MI355X collocation solver.
//
// Owns the device workspace of one problem instance, loads the problem's
// hiprtc-compiled code object and sequences the kernels of one damped Newton
// iteration on a single HIP stream.  The C ABI is declared in include/colloc.h.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "colloc.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return fail(CS_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

struct Seg4 {
  long long off[4];
  long long len[4];
};

// Kernel-side descriptors (layouts match struct Batch / CRSeg of the kernels).
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
// Per-instance state of the device-resident Newton loop and its settings
// (structs NtState and NtParams of the kernels).
struct NtParams {
  double tol, min_ratio;
  int hl, window, patience, max_iter, max_trials, pad;
};
struct NtState {
  double* nrm;
  double* taken;
  double* next;
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
// A captured Newton phase and the timed kernel groups inside it: group
// (kernel id, first stamp slot, last stamp slot) of the graph's slots.
constexpr int kStampSlots = 64;   // timestamp slots per graph
struct StampPair {
  int k, s0, s1;
};
struct Graph {
  hipGraphExec_t exec = nullptr;
  int base = 0;                    // the graph's first slot in the handle's stamp buffer
  int used = 0;                    // slots it writes
  std::vector<StampPair> prof;
};
struct CRSeg {
  int in_off, out_off, rec_off, nrows, subst;
};

enum Fn {
  F_FILL, F_RES_FULL, F_RES_NORM, F_JAC, F_JAC32, F_COND, F_COND32, F_COND_NS, F_COND_NS32, F_CR_PAIR, F_CR_SUBST,
  F_CR_FINAL, F_CR_BACK, F_RECOVER, F_RECOVER32, F_AXPY, F_INIT, F_SEGRES, F_RECZ, F_JAC_STATIC, F_JAC_STATIC32,
  F_AXPY_LISTED, F_COND_MF32, F_NT_CHECK, F_NT_SCAN, F_NT_FILL, F_NT_LS, F_NT_ACCEPT, F_NT_TICK, F_LINRES, F_LINRES32,
  F_STATE_ADD, F_CR_PAIR_DF, F_CR_PAIR_FF, F_CR_SUBST_DF, F_CR_SUBST_FF, F_CR_FINAL_F, F_CR_BACK_FD, F_CR_BACK_FF,
  F_L_STAGE, F_L_STAGE32, F_L_SINIT, F_L_SINIT32, F_L_TERM, F_L_TERM32, F_L_SOUT, F_L_SOUT32, F_L_CR_SUBST,
  F_L_CR_PAIR, F_L_CR_FINAL, F_L_CR_BACK, F_COND_NS_Q, F_COND_NS32_Q, F_COND_MF32_Q, F_ROWS_C,
  F_L_CR_SUBST_DF, F_L_CR_SUBST_FF, F_L_CR_PAIR_DF, F_L_CR_PAIR_FF, F_L_CR_FINAL_F, F_L_CR_BACK_FD, F_L_CR_BACK_FF,
  F_JAC32_GEMM,
  F_COUNT
};
const char* kFnNames[F_COUNT] = {
  "colloc_fill_u64", "colloc_residual_full", "colloc_residual_norm", "colloc_jacobian", "colloc_jacobian32",
  "colloc_condense", "colloc_condense32", "colloc_condense_ns", "colloc_condense_ns32", "colloc_cr_pair",
  "colloc_cr_subst",
  "colloc_cr_final", "colloc_cr_back", "colloc_recover", "colloc_recover32", "colloc_axpy4", "colloc_init_input",
  "colloc_segres", "colloc_recover_z", "colloc_jacobian_static", "colloc_jacobian_static32",
  "colloc_axpy_listed", "colloc_condense_mf32", "colloc_nt_check", "colloc_nt_list_scan", "colloc_nt_list_fill",
  "colloc_nt_line_search", "colloc_nt_accept", "colloc_nt_tick", "colloc_linres", "colloc_linres32",
  "colloc_state_add", "colloc_cr_pair_df", "colloc_cr_pair_ff", "colloc_cr_subst_df", "colloc_cr_subst_ff",
  "colloc_cr_final_f", "colloc_cr_back_fd", "colloc_cr_back_ff",
  "colloc_l_stage_elim", "colloc_l_stage_elim32", "colloc_l_series_init", "colloc_l_series_init32",
  "colloc_l_series_term", "colloc_l_series_term32", "colloc_l_series_out", "colloc_l_series_out32",
  "colloc_l_cr_subst", "colloc_l_cr_pair", "colloc_l_cr_final", "colloc_l_cr_back", "colloc_condense_ns_q",
  "colloc_condense_ns32_q", "colloc_condense_mf32_q", "colloc_rows_constant",
  "colloc_l_cr_subst_df", "colloc_l_cr_subst_ff", "colloc_l_cr_pair_df", "colloc_l_cr_pair_ff",
  "colloc_l_cr_final_f", "colloc_l_cr_back_fd", "colloc_l_cr_back_ff", "colloc_jacobian32_gemm"};
// the large-block kernels (ny > 64) take the place of the register-resident
// condensation and cyclic reduction: a module has one set or the other
bool large_kernel(int f) { return (f >= F_L_STAGE && f <= F_L_CR_BACK) || (f >= F_L_CR_SUBST_DF && f <= F_L_CR_BACK_FF); }
// kernels a module may have or not: the matrix-core float32 Jacobian of a
// straight-line unit whose problem also has the polynomial form
bool optional_kernel(int f) { return f == F_JAC32_GEMM; }
bool small_kernel(int f) {
  return (f >= F_COND && f <= F_CR_BACK) || f == F_COND_MF32 || (f >= F_CR_PAIR_DF && f <= F_CR_BACK_FF) ||
         (f >= F_COND_NS_Q && f <= F_ROWS_C);
}
constexpr int kLargeMaxTerms = 48;        // L_MAX_TERMS of the large-block series
constexpr int kLargeSlot = 10;            // L_SLOT: u64 per interval of its convergence slots
constexpr unsigned kLargeThreads = 256;   // L_THREADS
constexpr unsigned kLargeQrThreads = 512;   // L_QR_THREADS
constexpr unsigned kLargeFinThreads = 512;  // L_FIN_THREADS
constexpr unsigned kSubstThreads = 256;   // CS_THREADS of the substitution pairing kernel
constexpr int kDefaultSubstSpans = 64;    // substitution while a pair spans <= 1/64 of the instance's intervals
constexpr unsigned kNsThreads = 512;   // NS_THREADS of the series condensation kernel
// workgroups of the Gauss-Jordan condensation inside a captured Newton
// iteration (it walks the queue the series kernel filled; a few per CU)
constexpr int kGaussJordanGrid = 1024;

// Per-instance residual-norm slots of cs_compute_f_q_noalloc: a fixed device
// array of this library (64 KiB on every GPU it runs on, held until the
// process ends), so that entry point allocates nothing.  Host code reaches
// it only through the address hipGetSymbolAddress returns for the current
// device (g_scratch_dev), never through the array's name.
constexpr int kScratchSlots = 8192;
static __device__ __align__(256) unsigned long long scratch[kScratchSlots];
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device, looked up once per
// device (cs_create calls this with the handle's device current).
int scratch_address(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CS_E_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(CS_E_HIP, "hipGetSymbolAddress returned no address for the scratch slots");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return CS_OK;
}

}  // namespace

struct cs_handle {
  int ny, nz, np, m, n_max, nc, bmax;
  int N = 0;        // nodes of the current mesh, over all instances
  int B = 0;        // instances of the current mesh
  int nint = 0;     // intervals of the current mesh (N - B)
  int init_N = 0;   // mesh size of the saved initial input (y0buf)
  std::vector<int> hnoff;   // node offsets of the instances (B + 1)
  // launch shapes read from the module
  int cw = 0, fw = 0, npa = -1;
  hipStream_t stream = nullptr;
  hipModule_t module = nullptr;
  hipFunction_t fn[F_COUNT] = {};
  // sizes derived from dims
  int nj, ns, crw, recw, jstage, drsz, nrb;
  size_t recn = 0;                    // doubles per interval of the condensation record
  size_t st_y, st_yd, st_z, st_p, st_len;   // packed state offsets: [y | ydot | ztil | p]
  // device buffers
  double* t = nullptr;
  double* state = nullptr;   // current state
  double* trial = nullptr;   // line-search trial state
  double* dir = nullptr;     // Newton direction (same packing)
  double* c = nullptr;
  double* ytil = nullptr;
  double* fa = nullptr;
  double* fb = nullptr;
  double* rbc = nullptr;
  double* jac = nullptr;
  double* Dr = nullptr;
  double* rows0 = nullptr;   // BABD rows of every cyclic-reduction level, level after level
  size_t rows_cap = 0;       // rows0's capacity in rows
  bool rows_c_ok = false;    // level-0 rows hold C = I (and zero inactive H columns) for this mesh
  int jac_gemm = 0;          // stages per workgroup of the module's matrix-core float32 Jacobian (0: none)
  int jac_gemm_on = 1;       // cs_set_launch CS_LAUNCH_JACOBIAN_GEMM
  double* rec = nullptr;
  double* crrec = nullptr;
  double* segres = nullptr;
  double* zbuf = nullptr;
  double* y0buf = nullptr;
  double* steps = nullptr;   // per-instance step of the state update
  int* iinst = nullptr;      // instance of each interval
  int* ninst = nullptr;      // instance of each node
  int* noff = nullptr;       // node offsets (B + 1)
  int* act = nullptr;        // instances still iterating
  int* srch = nullptr;       // instances still line-searching
  int* fbk = nullptr;        // the intervals the series condensation left to Gauss-Jordan (a queue)
  int* alist = nullptr;      // the current launch list of intervals (one of the two buffers below)
  int nalist = 0;            // its length; 0: launch over all intervals
  int* alist_buf = nullptr;  // intervals of the instances still iterating
  int* slist_buf = nullptr;  // intervals of the instances still line-searching
  // cyclic-reduction schedule of the current mesh: per level, the instance
  // of each output row and each instance's CRSeg
  int* cr_oinst = nullptr;
  size_t cr_oinst_cap = 0;
  CRSeg* cr_seg = nullptr;
  size_t cr_seg_cap = 0;
  std::vector<int> lev_nout;          // output rows per level
  std::vector<char> lev_qr, lev_subst;  // level has instances paired by QR / by substitution
  int subst_spans = kDefaultSubstSpans; // see cs_set_babd_substitution (0: Householder everywhere)
  int dir_bits = 64;                    // condensation arithmetic of the Newton direction (64 or 32)
  int jac_bits = 64;                    // stored type of the last stage Jacobians (64 or 32)
  int jac_static_bits = 0;              // stored type whose static Jacobian entries fill the buffer (0: none)
  int rec_bits = 64;                    // stored type of the last condensation records (64 or 32)
  int mf_ok = 0;                        // the module's series condensation on matrix cores fits its sizes
  int cond_kind = 0;                    // see cs_set_condense_kernel
  int cond_grid = 0;                    // series condensation: 0 one workgroup per interval, else a persistent
                                        // grid of this many workgroups (cs_set_launch)
  int res_block = 256;                  // workgroup size of the straight-line residual kernel (cs_set_launch)
  int pair_threads = 0;                 // block size of colloc_cr_pair (from the module)
  int h_zero = 0;                       // the module's BABD rows carry no H (identically zero)
  int jac_parts = 1;                    // threads per stage of the Jacobian kernel (from the module)
  int tab_g = 0;                        // table-form callbacks: points per workgroup (0: straight-line, a thread each)
  int large = 0;                        // the module's large-block condensation and cyclic reduction (ny > 64)
  int l_tiles = 0, l_stiles = 0;        // large-block: workgroups per interval of a series term, per pair of a
                                        // substitution product
  double* l_ubuf = nullptr;             // large-block series: the second U buffer
  unsigned long long* l_slots = nullptr;  // per interval: convergence slots (kLargeSlot u64)
  unsigned long long* l_active = nullptr; // per term: intervals that computed it
  double* l_work = nullptr;             // large-block Householder pairs' work matrices
  size_t l_work_cap = 0;                // in doubles
  double* l_fin = nullptr;              // large-block boundary systems (one per instance)
  // device-resident Newton loop (cs_set_newton_graph)
  bool use_graph = false;
  int refine = 0;                       // refinement steps per Newton direction (cs_set_refinement)
  int babd_bits = 64;                   // stored type of the cyclic reduction's rows and records (cs_set_babd_precision)
  const int* nalistp = nullptr;         // set: the launch list alist is built on the device, its length here
  int* dlist = nullptr;                 // the device-built launch list (intervals of the iterating instances)
  double* nt_nrm = nullptr;
  double* nt_taken = nullptr;
  double* nt_next = nullptr;            // norm of each instance's accepted trial step
  int* nt_ints = nullptr;               // trials | done | gave | lofs (bmax each) | it | nlist
  unsigned long long* nt_cnt = nullptr; // [0] iterating, [1] searching, [2] listed intervals, [3] spare
  NtParams* nt_par = nullptr;           // the settings of the current solve (device)
  NtParams nt_par_host = {};            // their host copy (the source of the upload)
  double* nt_hist = nullptr;
  size_t nt_hist_cap = 0;
  double* nt_ring = nullptr;
  size_t nt_ring_cap = 0;
  unsigned mesh_version = 0;            // bumped whenever the mesh or the BABD schedule changes
  std::vector<double> graph_key;        // what the captured graphs were made for
  Graph g_dir0, g_dir, g_trial, g_acc;
  bool capturing = false;               // a Newton graph is being captured
  Graph* cap_prof = nullptr;            // the graph being captured with profiling on
  unsigned long long* stamps = nullptr; // kStampSlots per graph, written by colloc_stamp inside the graphs
  double stamp_khz = 0.0;               // the device's wall clock
  std::vector<size_t> lev_obase;      // offset of the level in cr_oinst
  std::vector<size_t> lev_sbase;      // offset of the level in cr_seg
  std::vector<size_t> lev_rbase;      // row offset of the level's input rows in rows0 (one more: the last output)
  unsigned long long cond_fallback = 0;   // intervals condensed by Gauss-Jordan since creation
  unsigned long long cond_terms = 0;      // series terms summed over intervals since creation
  unsigned long long cond_intervals = 0;  // intervals condensed since creation
  // [0, bmax): residual norms; [bmax, 2 bmax): segment-residual maxima;
  // [2 bmax, 2 bmax + 4): condensation counters (queue length, series
  // terms, intervals queued for Gauss-Jordan, the persistent grid's ticket)
  unsigned long long* scal = nullptr;
  int* status = nullptr;
  unsigned long long* scratch_slots = nullptr;   // this device's `scratch` (cs_compute_f_q_noalloc)
  // profiling
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[CS_KERNEL_COUNT];
  double prof_ms[CS_KERNEL_COUNT] = {};
  int prof_n[CS_KERNEL_COUNT] = {};
  std::vector<hipEvent_t> pool;

  double* y() const { return state + st_y; }
  double* yd() const { return state + st_yd; }
  double* zt() const { return state + st_z; }
  double* p() const { return state + st_p; }
  Batch batch(const int* mask) const { return Batch{iinst, ninst, noff, mask, B, nullptr, 0, nullptr}; }
  // with the launch list of the instances still iterating, when one is set
  // (a device-built list: every interval is launched, the work units past
  // the list's length return at once)
  Batch batch_listed(const int* mask) const {
    if (nalistp) return Batch{iinst, ninst, noff, mask, B, alist, 0, nalistp};
    return Batch{iinst, ninst, noff, mask, B, nalist ? alist : nullptr, nalist, nullptr};
  }
  int launch_intervals() const { return nalistp ? nint : (nalist ? nalist : nint); }
  bool listed() const { return nalistp || nalist > 0; }
};

namespace {

hipEvent_t take_event(cs_handle* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

// The device's wall clock (constant rate, stamp_khz) into slot p[0]: a
// captured graph's kernel groups are timed between two of these (an
// event record inside a graph does not give elapsed times reliably).
__global__ void colloc_stamp(unsigned long long* p) {
  if (threadIdx.x == 0) p[0] = wall_clock64();
}

int stamp(cs_handle* h, int slot) {
  unsigned long long* p = h->stamps + slot;
  hipLaunchKernelGGL(colloc_stamp, dim3(1), dim3(64), 0, h->stream, p);
  HIPCHK(hipGetLastError());
  return CS_OK;
}

// Times the launches made in its scope (kernel group k): HIP events around
// them, or inside a graph capture two wall-clock stamps read after each
// replay.
struct ProfScope {
  cs_handle* h;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  int s0 = -1;
  ProfScope(cs_handle* h_, int k_) : h(h_), k(k_) {
    if (!h->prof) return;
    if (h->cap_prof) {
      Graph* g = h->cap_prof;
      if (g->used + 2 <= kStampSlots && stamp(h, g->base + g->used) == CS_OK) s0 = g->used++;
      return;
    }
    a = take_event(h);
    b = take_event(h);
    hipEventRecord(a, h->stream);
  }
  ~ProfScope() {
    if (s0 >= 0) {
      Graph* g = h->cap_prof;
      if (stamp(h, g->base + g->used) == CS_OK) g->prof.push_back({k, s0, g->used++});
    } else if (a) {
      hipEventRecord(b, h->stream);
      h->pending[k].push_back({a, b});
    }
  }
};

void drain_prof(cs_handle* h) {
  for (int k = 0; k < CS_KERNEL_COUNT; ++k) {
    for (auto& pr : h->pending[k]) {
      float ms = 0.f;
      hipEventSynchronize(pr.second);
      hipEventElapsedTime(&ms, pr.first, pr.second);
      h->prof_ms[k] += ms;
      h->prof_n[k] += 1;
      h->pool.push_back(pr.first);
      h->pool.push_back(pr.second);
    }
    h->pending[k].clear();
  }
}

int launch(cs_handle* h, Fn f, unsigned grid, unsigned block, void** args) {
  if (!h->module) return fail(CS_E_NOMODULE, "no problem module loaded");
  if (!h->fn[f]) return fail(CS_E_UNSUPPORTED, "kernel %s is not in this problem's module", kFnNames[f]);
  if (grid == 0) return CS_OK;
  HIPCHK(hipModuleLaunchKernel(h->fn[f], grid, 1, 1, block, 1, 1, 0, h->stream, args, nullptr));
  return CS_OK;
}

#define TRY(expr)              \
  do {                         \
    int _r = (expr);           \
    if (_r != CS_OK) return _r; \
  } while (0)

int dmalloc(double** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(double)));
  return CS_OK;
}

int imalloc(int** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(int)));
  return CS_OK;
}

int fill_u64(cs_handle* h, unsigned long long* p, int n) {
  unsigned long long zero = 0ull;
  void* args[] = {&p, &zero, &n};
  return launch(h, F_FILL, (unsigned)((n + 255) / 256), 256, args);
}

// B per-instance maxima (u64 slots holding non-negative doubles) to the host.
int read_slots(cs_handle* h, const unsigned long long* slots, double* out) {
  std::vector<unsigned long long> v(h->B);
  HIPCHK(hipMemcpyAsync(v.data(), slots, h->B * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::memcpy(out, v.data(), h->B * sizeof(double));
  return CS_OK;
}

int need_mesh(cs_handle* h) {
  if (h->N < 2 || h->B < 1) return fail(CS_E_ARG, "mesh not set");
  return CS_OK;
}

int upload_ints(cs_handle* h, int* dev, const std::vector<int>& v) {
  HIPCHK(hipMemcpyAsync(dev, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  return CS_OK;
}

int upload_doubles(cs_handle* h, double* dev, const std::vector<double>& v) {
  HIPCHK(hipMemcpyAsync(dev, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  return CS_OK;
}

// Workgroups of a table-form callback kernel over the stages of nlaunch
// intervals and the B instances' boundary conditions, tab_g points each.
unsigned point_groups(cs_handle* h, int nlaunch) {
  const long long g = h->tab_g;
  return (unsigned)(((long long)nlaunch * h->m + g - 1) / g + (h->B + g - 1) / g);
}

// Residual of a packed state for the instances in mask (null: all); writes
// f_a/f_b/y_tilde/r_bc when full; per-instance norms to norms[B].
int residual_to(cs_handle* h, const double* st, double alpha, bool full, const int* mask,
                unsigned long long* slot) {
  TRY(need_mesh(h));
  ProfScope ps(h, CS_KERNEL_RESIDUAL);
  TRY(fill_u64(h, slot, h->B));
  int nint = h->nint;
  const double* y = st + h->st_y;
  const double* yd = st + h->st_yd;
  const double* zt = st + h->st_z;
  const double* p = st + h->st_p;
  // over the listed intervals when a launch list is set (the instances still
  // iterating, a superset of the line search's), masked by `mask`
  Batch bt = h->batch_listed(mask);
  void* args[] = {&nint, &bt, &h->t, &y, &yd, &zt, &p, &h->c, &alpha, &h->ytil, &h->fa, &h->fb, &h->rbc, &slot};
  const int nlaunch = h->launch_intervals();
  unsigned nthreads = (unsigned)(nlaunch * h->m + h->B);
  const unsigned rb = (unsigned)h->res_block;
  return launch(h, full ? F_RES_FULL : F_RES_NORM, h->tab_g ? point_groups(h, nlaunch) : (nthreads + rb - 1) / rb,
                h->tab_g ? 256u : rb, args);
}

int residual(cs_handle* h, const double* st, double alpha, bool full, const int* mask, double* norms) {
  TRY(residual_to(h, st, alpha, full, mask, h->scal));
  return read_slots(h, h->scal, norms);
}

int jacobian_static(cs_handle* h);

int jacobian(cs_handle* h, double alpha, const int* mask) {
  ProfScope ps(h, CS_KERNEL_JACOBIAN);
  int nint = h->nint;
  double* y = h->y();
  double* p = h->p();
  double* zt = h->zt();
  Batch bt = h->batch_listed(mask);
  void* args[] = {&nint, &bt, &y, &h->ytil, &zt, &p, &h->c, &alpha, &h->jac, &h->Dr};
  unsigned nthreads = (unsigned)(h->launch_intervals() * h->m + h->B);
  // the FP32 direction stores the stage Jacobians in float32 (in the same buffer)
  h->jac_bits = h->dir_bits;
  TRY(jacobian_static(h));
  if (h->jac_bits == 32 && h->jac_gemm > 0 && h->jac_gemm_on) {
    // matrix-core form: a workgroup per jac_gemm stages, then Dr's threads
    const long long nst = (long long)h->launch_intervals() * h->m;
    const unsigned grid = (unsigned)((nst + h->jac_gemm - 1) / h->jac_gemm + (h->B + 255) / 256);
    return launch(h, F_JAC32_GEMM, grid, 256, args);
  }
  const unsigned grid = h->tab_g ? point_groups(h, h->launch_intervals())
                                 : (nthreads + 255) / 256 * (unsigned)h->jac_parts;
  return launch(h, h->jac_bits == 32 ? F_JAC32 : F_JAC, grid, 256, args);
}

int jacobian_static(cs_handle* h) {
  if (h->jac_static_bits != h->jac_bits) {
    // the static entries, once for the whole buffer in this stored type (the
    // other type's layout overwrote them)
    long long nst = (long long)(h->n_max - 1) * h->m;
    void* sargs[] = {&nst, &h->jac};
    const unsigned blocks = (unsigned)std::min<long long>(nst, 16384);
    TRY(launch(h, h->jac_bits == 32 ? F_JAC_STATIC32 : F_JAC_STATIC, blocks, 256, sargs));
    h->jac_static_bits = h->jac_bits;
  }
  return CS_OK;
}

// Condensation: the series kernel handles every interval whose stage system
// converges as a Neumann series (fine meshes) and queues the others; the
// Gauss-Jordan kernel redoes the queued ones.
int condense_large(cs_handle* h, const int* mask, long long ncond);

// C = I and the zero inactive H columns of the level-0 BABD rows
// (colloc_rows_constant): written once per mesh and module, outside any
// captured iteration; the register-resident condensation kernels write only
// G, the active H columns and b.  The large-block kernels write whole rows.
int fill_rows_constant(cs_handle* h) {
  h->rows_c_ok = false;
  if (!h->module || h->large || h->nint <= 0 || !h->fn[F_ROWS_C]) return CS_OK;
  int nint = h->nint;
  void* args[] = {&nint, &h->rows0};
  TRY(launch(h, F_ROWS_C, (unsigned)nint, 256, args));
  h->rows_c_ok = true;
  return CS_OK;
}

int condense(cs_handle* h, const int* mask, long long ncond) {
  if (h->large) return condense_large(h, mask, ncond);
  if (!h->rows_c_ok) TRY(fill_rows_constant(h));
  ProfScope ps(h, CS_KERNEL_CONDENSE);
  h->cond_intervals += (unsigned long long)ncond;
  int nint = h->nint;
  // [queue length, series terms, intervals queued]
  unsigned long long* counters = h->scal + 2 * h->bmax;
  // captured in a graph: the terms and queued intervals add up over the
  // whole Newton solve, and the Gauss-Jordan kernel is always launched, on
  // a fixed grid that walks whatever the queue holds
  const bool in_graph = h->capturing || h->nalistp != nullptr;
  TRY(fill_u64(h, counters, in_graph ? 1 : 3));
  // the persistent grid's ticket (counters[3]), cleared for every launch
  int persist = h->cond_grid > 0 ? 1 : 0;
  if (persist) TRY(fill_u64(h, counters + 3, 1));
  Batch bt = h->batch_listed(mask);
  void* args_ns[] = {&nint, &bt, &h->t, &h->jac, &h->fa, &h->fb, &h->rows0, &h->rec, &h->status, &h->fbk, &counters};
  // the arithmetic follows the stored type of the stage Jacobians, and so
  // does the records' type
  const bool f32 = h->jac_bits == 32;
  h->rec_bits = h->jac_bits;
  // one workgroup per listed interval, or the persistent grid (at most one
  // workgroup per interval)
  const unsigned grid_ns = persist ? (unsigned)std::min(h->cond_grid, h->launch_intervals())
                                   : (unsigned)h->launch_intervals();
  if (f32 && h->mf_ok && h->cond_kind == 0)   // the series products on matrix cores, one wave per stage block
    TRY(launch(h, persist ? F_COND_MF32_Q : F_COND_MF32, grid_ns, (unsigned)(h->m - 1) * 64, args_ns));
  else
    TRY(launch(h, f32 ? (persist ? F_COND_NS32_Q : F_COND_NS32) : (persist ? F_COND_NS_Q : F_COND_NS), grid_ns,
               kNsThreads, args_ns));
  unsigned grid = (unsigned)std::min(h->launch_intervals(), kGaussJordanGrid);
  if (!in_graph) {
    unsigned long long cnt[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(cnt, counters, sizeof cnt, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cond_fallback += cnt[2];
    h->cond_terms += cnt[1];
    if (cnt[0] == 0) return CS_OK;
    grid = (unsigned)cnt[0];
  }
  void* args[] = {&h->iinst, &h->fbk, &counters, &h->t, &h->jac, &h->fa, &h->fb, &h->rows0, &h->rec, &h->status};
  return launch(h, f32 ? F_COND32 : F_COND, grid, (unsigned)(h->cw * 64), args);
}

// Cyclic-reduction schedule of the current mesh (host side, once per mesh):
// at each level every instance pairs its rows, carrying an odd last one,
// until each has a single row; the level's output rows are numbered
// instance by instance, so after the last level instance b's row is row b.
int build_cr_schedule(cs_handle* h) {
  const int B = h->B;
  ++h->mesh_version;
  std::vector<int> n(B), in_off(B);
  for (int b = 0; b < B; ++b) {
    n[b] = h->hnoff[b + 1] - h->hnoff[b] - 1;
    in_off[b] = h->hnoff[b] - b;
  }
  // substitution levels of instance b: the largest L with 2^L pairs-spans,
  // i.e. 2^L * subst_spans <= its interval count
  std::vector<int> nsub(B, 0);
  if (h->subst_spans > 0)
    for (int b = 0; b < B; ++b)
      while (((long long)h->subst_spans << (nsub[b] + 1)) <= n[b]) ++nsub[b];
  std::vector<int> oinst;
  std::vector<CRSeg> segs;
  h->lev_nout.clear();
  h->lev_obase.clear();
  h->lev_sbase.clear();
  h->lev_rbase.assign(1, 0);   // level 0 reads the condensation's rows, one per interval
  size_t nin = 0;
  for (int b = 0; b < B; ++b) nin += (size_t)n[b];
  h->lev_qr.clear();
  h->lev_subst.clear();
  size_t rec_base = 0;
  for (int lev = 0;; ++lev) {
    bool more = false;
    for (int b = 0; b < B; ++b) more |= n[b] > 1;
    if (!more) break;
    h->lev_obase.push_back(oinst.size());
    h->lev_sbase.push_back(segs.size());
    int out = 0;
    char any_qr = 0, any_sub = 0;
    for (int b = 0; b < B; ++b) {
      const int nout = n[b] / 2 + (n[b] & 1);
      // 2: the next level substitutes too (it does not read C, which stays I)
      const int sub = lev < nsub[b] ? (lev + 1 < nsub[b] ? 2 : 1) : 0;
      segs.push_back(CRSeg{in_off[b], out, (int)rec_base, n[b], sub});
      if (sub) any_sub = 1;
      else any_qr = 1;
      rec_base += n[b] / 2;
      for (int k = 0; k < nout; ++k) oinst.push_back(b);
      in_off[b] = out;
      out += nout;
      n[b] = nout;
    }
    h->lev_nout.push_back(out);
    h->lev_qr.push_back(any_qr);
    h->lev_subst.push_back(any_sub);
    // every level's rows stay in place: the substitution levels' back
    // substitution reads its input rows instead of a record
    h->lev_rbase.push_back(h->lev_rbase.back() + nin);
    nin = (size_t)out;
  }
  if (h->lev_rbase.back() + nin > h->rows_cap)
    return fail(CS_E_SHAPE, "cyclic reduction needs %zu BABD rows, the handle holds %zu",
                h->lev_rbase.back() + nin, h->rows_cap);
  if (oinst.size() > h->cr_oinst_cap) {
    if (h->cr_oinst) hipFree(h->cr_oinst);
    h->cr_oinst = nullptr;
    h->cr_oinst_cap = 0;
    TRY(imalloc(&h->cr_oinst, oinst.size()));
    h->cr_oinst_cap = oinst.size();
  }
  if (segs.size() > h->cr_seg_cap) {
    if (h->cr_seg) hipFree(h->cr_seg);
    h->cr_seg = nullptr;
    h->cr_seg_cap = 0;
    HIPCHK(hipMalloc((void**)&h->cr_seg, segs.size() * sizeof(CRSeg)));
    h->cr_seg_cap = segs.size();
  }
  if (!oinst.empty()) TRY(upload_ints(h, h->cr_oinst, oinst));
  if (!segs.empty())
    HIPCHK(hipMemcpyAsync(h->cr_seg, segs.data(), segs.size() * sizeof(CRSeg), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return CS_OK;
}

// The FP32 cyclic reduction (babd_bits 32) stores every level's output rows
// and records as float at the same element offsets as the float64 one, in
// the same buffers: each float access lies at or below the byte address of
// the float64 access to the same element.  Level 0 reads the condensation's
// float64 rows; cr_final and cr_back read the float32 levels.
int babd_large(cs_handle* h, const int* mask);

int babd(cs_handle* h, const int* mask) {
  if (h->large) return babd_large(h, mask);
  ProfScope ps(h, CS_KERNEL_BABD);
  double* dy = h->dir + h->st_y;
  double* dp = h->dir + h->st_p;
  Batch bt = h->batch(mask);
  const size_t rowsz = (size_t)h->ny * h->crw;
  const unsigned pair_block = (unsigned)h->pair_threads;   // 4 lanes per record column
  const int nlev = (int)h->lev_nout.size();
  const bool f32 = h->babd_bits == 32;
  for (int l = 0; l < nlev; ++l) {
    const int* oinst = h->cr_oinst + h->lev_obase[l];
    const CRSeg* seg = h->cr_seg + h->lev_sbase[l];
    const int* act = mask;
    double* rin = h->rows0 + h->lev_rbase[l] * rowsz;
    double* rout = h->rows0 + h->lev_rbase[l + 1] * rowsz;
    void* args[] = {&oinst, &seg, &act, &rin, &rout, &h->crrec};
    const Fn fs = !f32 ? F_CR_SUBST : (l == 0 ? F_CR_SUBST_DF : F_CR_SUBST_FF);
    const Fn fp = !f32 ? F_CR_PAIR : (l == 0 ? F_CR_PAIR_DF : F_CR_PAIR_FF);
    if (h->lev_subst[l]) TRY(launch(h, fs, (unsigned)h->lev_nout[l], kSubstThreads, args));
    if (h->lev_qr[l]) TRY(launch(h, fp, (unsigned)h->lev_nout[l], pair_block, args));
  }
  {
    double* rin = h->rows0 + h->lev_rbase[nlev] * rowsz;
    void* args[] = {&bt, &rin, &h->Dr, &h->rbc, &dy, &dp, &h->status};
    TRY(launch(h, f32 && nlev > 0 ? F_CR_FINAL_F : F_CR_FINAL, (unsigned)h->B, (unsigned)(h->fw * 64), args));
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int* oinst = h->cr_oinst + h->lev_obase[l];
    const CRSeg* seg = h->cr_seg + h->lev_sbase[l];
    double* rin = h->rows0 + h->lev_rbase[l] * rowsz;
    void* args[] = {&l, &bt, &oinst, &seg, &h->crrec, &rin, &dy, &dp};
    const Fn fb = !f32 ? F_CR_BACK : (l == 0 ? F_CR_BACK_FD : F_CR_BACK_FF);
    TRY(launch(h, fb, (unsigned)h->lev_nout[l], 64, args));
  }
  return CS_OK;
}

// Large-block condensation (ny > 64, colloc_large.h): the stage
// eliminations, R, then series terms until every listed interval has
// converged (the host reads, after each term, how many intervals computed
// it), then the records and BABD rows.  An interval whose series has not
// converged after kLargeMaxTerms terms is an error: the large-block path
// has no Gauss-Jordan fallback.
int condense_large(cs_handle* h, const int* mask, long long ncond) {
  ProfScope ps(h, CS_KERNEL_CONDENSE);
  if (h->capturing || h->nalistp) return fail(CS_E_UNSUPPORTED, "the large-block condensation is host-driven");
  h->cond_intervals += (unsigned long long)ncond;
  int nint = h->nint;
  unsigned long long* counters = h->scal + 2 * h->bmax;
  TRY(fill_u64(h, counters, 3));
  TRY(fill_u64(h, h->l_active, kLargeMaxTerms + 1));
  Batch bt = h->batch_listed(mask);
  const bool f32 = h->jac_bits == 32;
  h->rec_bits = h->jac_bits;
  const unsigned nl = (unsigned)h->launch_intervals();
  {
    void* args[] = {&nint, &bt, &h->jac, &h->fa, &h->rec, &h->status};
    TRY(launch(h, f32 ? F_L_STAGE32 : F_L_STAGE, nl * (unsigned)h->m, kLargeThreads, args));
  }
  {
    void* args[] = {&nint, &bt, &h->t, &h->rec, &h->l_slots};
    TRY(launch(h, f32 ? F_L_SINIT32 : F_L_SINIT, nl * (unsigned)(h->m - 1), kLargeThreads, args));
  }
  const long long rstride = (long long)h->recn;
  const long long ustride = (long long)h->ns * (h->ny + h->npa + 1);
  int nterms = 0;
  for (int term = 1; term <= kLargeMaxTerms; ++term) {
    // U(term - 1) is in the record for even term - 1, in l_ubuf for odd
    void* uold = (term - 1) % 2 == 0 ? (void*)h->rec : (void*)h->l_ubuf;
    void* unew = term % 2 == 0 ? (void*)h->rec : (void*)h->l_ubuf;
    long long so = (term - 1) % 2 == 0 ? rstride : ustride;
    long long sn = term % 2 == 0 ? rstride : ustride;
    void* args[] = {&nint, &bt, &h->t, &h->rec, &uold, &so, &unew, &sn, &h->l_slots, &h->l_active, &term};
    TRY(launch(h, f32 ? F_L_TERM32 : F_L_TERM, nl * (unsigned)h->l_tiles, kLargeThreads, args));
    nterms = term;
    if (term >= 2) {
      unsigned long long act = 0;
      HIPCHK(hipMemcpyAsync(&act, h->l_active + term, sizeof act, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      if (act == 0) break;
    }
  }
  {
    void* args[] = {&nint, &bt, &h->fb, &h->rows0, &h->rec, &h->l_ubuf, &h->l_slots, &nterms, &counters};
    TRY(launch(h, f32 ? F_L_SOUT32 : F_L_SOUT, nl, kLargeThreads, args));
  }
  unsigned long long cnt[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(cnt, counters, sizeof cnt, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->cond_terms += cnt[1];
  if (cnt[2] != 0)
    return fail(CS_E_SINGULAR, "large-block condensation: the series of %llu interval(s) did not converge in %d terms "
                "(the mesh is too coarse for the series)", cnt[2], kLargeMaxTerms);
  return CS_OK;
}

// Large-block cyclic reduction (float64): substitution products on matrix
// cores, Householder pairs on work matrices in device memory, the boundary
// systems by Gauss-Jordan in device memory, back substitution per record.
int babd_large(cs_handle* h, const int* mask) {
  ProfScope ps(h, CS_KERNEL_BABD);
  const bool f32 = h->babd_bits == 32;
  double* dy = h->dir + h->st_y;
  double* dp = h->dir + h->st_p;
  Batch bt = h->batch(mask);
  const size_t rowsz = (size_t)h->ny * h->crw;
  const int nlev = (int)h->lev_nout.size();
  // the Householder pairs' work matrices: one per output row of the widest level that pairs by QR
  size_t need = 0;
  for (int l = 0; l < nlev; ++l)
    if (h->lev_qr[l]) need = std::max(need, (size_t)h->lev_nout[l]);
  need *= (size_t)(2 * h->ny) * (size_t)(3 * h->ny + (h->h_zero ? 0 : h->np) + 1);
  if (need > h->l_work_cap) {
    if (h->l_work) hipFree(h->l_work);
    h->l_work = nullptr;
    h->l_work_cap = 0;
    TRY(dmalloc(&h->l_work, need));
    h->l_work_cap = need;
  }
  for (int l = 0; l < nlev; ++l) {
    const int* oinst = h->cr_oinst + h->lev_obase[l];
    const CRSeg* seg = h->cr_seg + h->lev_sbase[l];
    const int* act = mask;
    double* rin = h->rows0 + h->lev_rbase[l] * rowsz;
    double* rout = h->rows0 + h->lev_rbase[l + 1] * rowsz;
    if (h->lev_subst[l]) {
      void* args[] = {&oinst, &seg, &act, &rin, &rout};
      const Fn fs = !f32 ? F_L_CR_SUBST : (l == 0 ? F_L_CR_SUBST_DF : F_L_CR_SUBST_FF);
      TRY(launch(h, fs, (unsigned)h->lev_nout[l] * (unsigned)h->l_stiles, kLargeThreads, args));
    }
    if (h->lev_qr[l]) {
      void* args[] = {&oinst, &seg, &act, &rin, &rout, &h->crrec, &h->l_work};
      const Fn fp = !f32 ? F_L_CR_PAIR : (l == 0 ? F_L_CR_PAIR_DF : F_L_CR_PAIR_FF);
      TRY(launch(h, fp, (unsigned)h->lev_nout[l], kLargeQrThreads, args));
    }
  }
  {
    double* rin = h->rows0 + h->lev_rbase[nlev] * rowsz;
    void* args[] = {&bt, &rin, &h->Dr, &h->rbc, &dy, &dp, &h->status, &h->l_fin};
    TRY(launch(h, f32 && nlev > 0 ? F_L_CR_FINAL_F : F_L_CR_FINAL, (unsigned)h->B, kLargeFinThreads, args));
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int* oinst = h->cr_oinst + h->lev_obase[l];
    const CRSeg* seg = h->cr_seg + h->lev_sbase[l];
    double* rin = h->rows0 + h->lev_rbase[l] * rowsz;
    void* args[] = {&l, &bt, &oinst, &seg, &h->crrec, &rin, &dy, &dp};
    const Fn fb = !f32 ? F_L_CR_BACK : (l == 0 ? F_L_CR_BACK_FD : F_L_CR_BACK_FF);
    TRY(launch(h, fb, (unsigned)h->lev_nout[l], kLargeThreads, args));
  }
  return CS_OK;
}

int recover(cs_handle* h, const int* mask) {
  ProfScope ps(h, CS_KERNEL_RECOVER);
  int nint = h->nint;
  double* dy = h->dir + h->st_y;
  double* dp = h->dir + h->st_p;
  double* dyd = h->dir + h->st_yd;
  double* dz = h->dir + h->st_z;
  Batch bt = h->batch_listed(mask);
  void* args[] = {&nint, &bt, &h->rec, &dy, &dp, &dyd, &dz};
  return launch(h, h->rec_bits == 32 ? F_RECOVER32 : F_RECOVER, (unsigned)h->launch_intervals(), 256, args);
}

Seg4 state_segs(cs_handle* h) {
  Seg4 s;
  s.off[0] = (long long)h->st_y;  s.len[0] = (long long)h->N * h->ny;
  s.off[1] = (long long)h->st_yd; s.len[1] = (long long)h->nint * h->m * h->ny;
  s.off[2] = (long long)h->st_z;  s.len[2] = (long long)h->nint * h->m * h->nz;
  s.off[3] = (long long)h->st_p;  s.len[3] = (long long)h->B * h->np;
  return s;
}

// out = x + steps[b] d per instance b (steps on the device).
int axpy_state(cs_handle* h, const double* x, const double* d, double* out) {
  Seg4 s = state_segs(h);
  if (h->listed()) {   // a launch list is set: just the listed intervals' entries
    Batch bt = h->batch_listed(nullptr);
    void* args[] = {&s, &bt, &x, &d, &h->steps, &out};
    return launch(h, F_AXPY_LISTED, (unsigned)h->launch_intervals(), 64, args);
  }
  long long total = s.len[0] + s.len[1] + s.len[2] + s.len[3];
  Batch bt = h->batch(nullptr);
  void* args[] = {&s, &bt, &x, &d, &h->steps, &out};
  return launch(h, F_AXPY, (unsigned)((total + 255) / 256), 256, args);
}

// rho = F + J delta for the direction in h->dir (see colloc_linres): per-
// instance max |rho| into slot[B]; write: rho replaces f_a, f_b, r_bc.
int linear_residual(cs_handle* h, const int* mask, unsigned long long* slot, bool write) {
  TRY(fill_u64(h, slot, h->B));
  int nint = h->nint;
  Batch bt = h->batch_listed(mask);
  const double* dy = h->dir + h->st_y;
  const double* dyd = h->dir + h->st_yd;
  const double* dz = h->dir + h->st_z;
  const double* dp = h->dir + h->st_p;
  int wr = write ? 1 : 0;
  void* args[] = {&nint, &bt, &h->t, &h->jac, &dy, &dyd, &dz, &dp, &h->fa, &h->fb, &h->rbc, &h->Dr, &slot, &wr};
  const long long rows = (long long)h->launch_intervals() * (h->m * (h->ny + h->nz) + h->ny) +
                         (long long)h->B * (h->ny + h->np);
  return launch(h, h->jac_bits == 32 ? F_LINRES32 : F_LINRES, (unsigned)((rows + 255) / 256), 256, args);
}

int state_add(cs_handle* h, const double* a, const double* b, double* out) {
  long long n = (long long)h->st_p + (long long)h->B * h->np;
  void* args[] = {&n, &a, &b, &out};
  return launch(h, F_STATE_ADD, (unsigned)((n + 255) / 256), 256, args);
}

// Iterative refinement of the Newton direction: rho = F + J delta in
// float64 (with the stored Jacobians), the correction e from the same
// direction kernels with rho in place of F (J e = -rho), delta <- delta + e.
// The trial state holds delta meanwhile (the line search writes it later).
int refine_direction(cs_handle* h, const int* mask, long long ncond) {
  TRY(linear_residual(h, mask, h->scal, true));
  TRY(state_add(h, h->dir, nullptr, h->trial));
  TRY(condense(h, mask, ncond));
  TRY(babd(h, mask));
  TRY(recover(h, mask));
  return state_add(h, h->trial, h->dir, h->dir);
}

int copy_out(cs_handle* h, double* host, const double* dev, size_t n) {
  if (!host || n == 0) return CS_OK;
  HIPCHK(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return CS_OK;
}

int copy_in(cs_handle* h, double* dev, const double* host, size_t n) {
  if (!host || n == 0) return CS_OK;
  HIPCHK(hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  return CS_OK;
}

int sync(cs_handle* h) {
  HIPCHK(hipStreamSynchronize(h->stream));
  return CS_OK;
}

int init_from_saved(cs_handle* h);
int set_batch_mesh(cs_handle* h, int B, const int* nodes, const double* t);

}  // namespace

extern "C" {

const char* cs_last_error(void) { return g_err.c_str(); }

const char* cs_version(void) { return "colloc-mi355x 0.1"; }

void cs_free(void* ptr) { std::free(ptr); }

int cs_jit_compile(const char* source, const char* name, const char* include_dir, void** code, size_t* size,
                   char** log) {
  if (!source || !code || !size) return fail(CS_E_ARG, "null argument");
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source, name ? name : "colloc_problem.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    return fail(CS_E_COMPILE, "hiprtcCreateProgram failed");
  std::string inc = std::string("-I") + (include_dir ? include_dir : ".");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", inc.c_str(),
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t lsz = 0;
  hiprtcGetProgramLogSize(prog, &lsz);
  std::string lg(lsz + 1, '\0');
  if (lsz) hiprtcGetProgramLog(prog, &lg[0]);
  if (log) {
    *log = (char*)std::malloc(lg.size() + 1);
    std::memcpy(*log, lg.c_str(), lg.size() + 1);
  }
  if (rc != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(CS_E_COMPILE, "hiprtc compile failed: %.900s", lg.c_str());
  }
  size_t csz = 0;
  hiprtcGetCodeSize(prog, &csz);
  void* buf = std::malloc(csz);
  hiprtcGetCode(prog, (char*)buf);
  hiprtcDestroyProgram(&prog);
  *code = buf;
  *size = csz;
  return CS_OK;
}

int cs_create(const cs_dims* d, cs_handle** out) {
  if (!d || !out) return fail(CS_E_ARG, "null argument");
  const int bmax = d->batch > 0 ? d->batch : 1;
  if (d->ny < 1 || d->nz < 0 || d->np < 0 || d->m < 2 || d->n_max < 2 || d->nc < 0 || d->batch < 0 ||
      d->n_max < 2 * bmax)
    return fail(CS_E_SHAPE, "invalid dims ny=%d nz=%d np=%d m=%d n_max=%d nc=%d batch=%d", d->ny, d->nz, d->np,
                d->m, d->n_max, d->nc, d->batch);
  cs_handle* h = new cs_handle();
  h->ny = d->ny; h->nz = d->nz; h->np = d->np; h->m = d->m; h->n_max = d->n_max; h->nc = d->nc; h->bmax = bmax;
  const int ny = h->ny, nz = h->nz, np = h->np, m = h->m;
  h->nj = ny + nz + np;
  h->ns = (m - 1) * ny;
  h->crw = 2 * ny + np + 1;
  h->recw = 3 * ny + np + 1;
  h->jstage = (ny + nz) * h->nj;
  h->drsz = (ny + np) * (2 * ny + np);
  h->nrb = ny + np;
  const size_t nmax = (size_t)h->n_max, nint = nmax - 1, bm = (size_t)bmax;
  h->st_y = 0;
  h->st_yd = nmax * ny;
  h->st_z = h->st_yd + nint * m * ny;
  h->st_p = h->st_z + nint * m * nz;
  h->st_len = h->st_p + bm * np;
  int r = CS_OK;
  auto A = [&](double** p, size_t n) { if (r == CS_OK) r = dmalloc(p, n); };
  auto I = [&](int** p, size_t n) { if (r == CS_OK) r = imalloc(p, n); };
  A(&h->t, nmax);
  A(&h->state, h->st_len);
  A(&h->trial, h->st_len);
  A(&h->dir, h->st_len);
  A(&h->c, bm * std::max(h->nc, 1));
  A(&h->ytil, nint * m * ny);
  A(&h->fa, nint * m * (ny + nz));
  A(&h->fb, nint * ny);
  A(&h->rbc, bm * h->nrb);
  A(&h->jac, nint * m * (size_t)h->jstage);
  A(&h->Dr, bm * h->drsz);
  // the rows of all levels: an instance of n rows has at most n + 32 output
  // rows over its levels (ceil(n / 2^k) <= n / 2^k + 1, at most 32 levels)
  h->rows_cap = 2 * nint + 32 * bm;
  A(&h->rows0, h->rows_cap * ny * (size_t)h->crw);
  A(&h->crrec, nint * ny * (size_t)h->recw);
  A(&h->segres, nmax);
  A(&h->zbuf, nmax * std::max(nz, 1));
  A(&h->y0buf, nmax * (ny + std::max(nz, 1)) + bm * np);
  A(&h->steps, bm);
  I(&h->iinst, nint);
  I(&h->ninst, nmax);
  I(&h->noff, bm + 1);
  I(&h->act, bm);
  I(&h->srch, bm);
  I(&h->fbk, nint);
  I(&h->alist_buf, nint);
  I(&h->slist_buf, nint);
  I(&h->dlist, nint);
  A(&h->nt_nrm, bm);
  A(&h->nt_taken, bm);
  A(&h->nt_next, bm);
  I(&h->nt_ints, 4 * bm + 2);
  if (r == CS_OK) {
    hipError_t e = hipMalloc((void**)&h->nt_cnt, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&h->stamps, 4 * kStampSlots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&h->nt_par, sizeof(NtParams));
    int dev = 0, khz = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (e == hipSuccess && khz <= 0) e = hipErrorInvalidValue;
    if (e != hipSuccess) r = fail(CS_E_HIP, "allocation failed: %s", hipGetErrorString(e));
    h->stamp_khz = (double)khz;
    h->g_dir.base = 0;
    h->g_trial.base = kStampSlots;
    h->g_acc.base = 2 * kStampSlots;
    h->g_dir0.base = 3 * kStampSlots;
  }
  if (r == CS_OK) {
    hipError_t e = hipMalloc((void**)&h->scal, (2 * bm + 4) * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&h->status, sizeof(int));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemset(h->status, 0, sizeof(int));
    if (e == hipSuccess) e = hipMemset(h->c, 0, sizeof(double) * bm * std::max(h->nc, 1));
    if (e != hipSuccess) r = fail(CS_E_HIP, "allocation failed: %s", hipGetErrorString(e));
  }
  if (r == CS_OK) {
    void* sp = nullptr;
    r = scratch_address(&sp);
    h->scratch_slots = (unsigned long long*)sp;
  }
  if (r != CS_OK) {
    std::string msg = g_err;
    cs_destroy(h);
    g_err = msg;
    return r;
  }
  *out = h;
  return CS_OK;
}

void cs_destroy(cs_handle* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (Graph* g : {&h->g_dir0, &h->g_dir, &h->g_trial, &h->g_acc})
    if (g->exec) hipGraphExecDestroy(g->exec);
  if (h->stamps) hipFree(h->stamps);
  if (h->nt_par) hipFree(h->nt_par);
  double* bufs[] = {h->t, h->state, h->trial, h->dir, h->c, h->ytil, h->fa, h->fb, h->rbc, h->jac, h->Dr,
                    h->rows0, h->rec, h->crrec, h->segres, h->zbuf, h->y0buf, h->steps, h->nt_nrm, h->nt_taken, h->nt_next,
                    h->nt_hist, h->nt_ring, h->l_ubuf, h->l_work, h->l_fin};
  if (h->l_slots) hipFree(h->l_slots);
  if (h->l_active) hipFree(h->l_active);
  for (double* b : bufs)
    if (b) hipFree(b);
  int* ibufs[] = {h->iinst, h->ninst, h->noff, h->act, h->srch, h->fbk, h->alist_buf, h->slist_buf, h->cr_oinst,
                  h->dlist, h->nt_ints};
  if (h->nt_cnt) hipFree(h->nt_cnt);
  for (int* b : ibufs)
    if (b) hipFree(b);
  if (h->cr_seg) hipFree(h->cr_seg);
  if (h->scal) hipFree(h->scal);
  if (h->status) hipFree(h->status);
  for (int k = 0; k < CS_KERNEL_COUNT; ++k)
    for (auto& pr : h->pending[k]) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
  for (hipEvent_t e : h->pool) hipEventDestroy(e);
  if (h->module) hipModuleUnload(h->module);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

// The captured Newton graphs name the module's kernels and the buffers of
// the moment: they go with the module.
static void drop_graphs(cs_handle* h) {
  if (h->stream) hipStreamSynchronize(h->stream);
  for (Graph* g : {&h->g_dir0, &h->g_dir, &h->g_trial, &h->g_acc}) {
    if (g->exec) hipGraphExecDestroy(g->exec);
    g->exec = nullptr;
    g->prof.clear();
    g->used = 0;
  }
  h->graph_key.clear();
}

int cs_load_module(cs_handle* h, const void* code, size_t size) {
  if (!h || !code || size == 0) return fail(CS_E_ARG, "null argument");
  drop_graphs(h);
  if (h->module) {
    hipModuleUnload(h->module);
    h->module = nullptr;
  }
  HIPCHK(hipModuleLoadData(&h->module, code));
  // which of the two kernel sets the module has
  hipFunction_t probe = nullptr;
  const bool mod_large = hipModuleGetFunction(&probe, h->module, kFnNames[F_L_STAGE]) == hipSuccess;
  // the failed lookup of a small module is the runtime's last error now: clear
  // it, or the next caller that checks hipGetLastError (torch does) reports it
  if (!mod_large) (void)hipGetLastError();
  for (int f = 0; f < F_COUNT; ++f) {
    h->fn[f] = nullptr;
    if ((mod_large && small_kernel(f)) || (!mod_large && large_kernel(f))) continue;
    if (optional_kernel(f)) {
      if (hipModuleGetFunction(&h->fn[f], h->module, kFnNames[f]) != hipSuccess) {
        h->fn[f] = nullptr;
        (void)hipGetLastError();   // absent: not an error (see the probe above)
      }
      continue;
    }
    hipError_t e = hipModuleGetFunction(&h->fn[f], h->module, kFnNames[f]);
    if (e != hipSuccess) return fail(CS_E_HIP, "kernel %s missing: %s", kFnNames[f], hipGetErrorString(e));
  }
  // module metadata: sizes and launch shapes baked into the code object
  hipDeviceptr_t gp;
  size_t gsz;
  HIPCHK(hipModuleGetGlobal(&gp, &gsz, h->module, "colloc_meta"));
  int meta[8] = {};
  HIPCHK(hipMemcpy(meta, (void*)gp, sizeof meta, hipMemcpyDeviceToHost));
  if (meta[0] != h->ny || meta[1] != h->nz || meta[2] != h->np || meta[3] != h->m || meta[4] != h->nc)
    return fail(CS_E_SHAPE, "module sizes (%d,%d,%d,%d,%d) differ from handle (%d,%d,%d,%d,%d)", meta[0], meta[1],
                meta[2], meta[3], meta[4], h->ny, h->nz, h->np, h->m, h->nc);
  h->cw = meta[5];
  h->fw = meta[6];
  h->jac_gemm = 0;
  if (h->fn[F_JAC32_GEMM]) {   // its stages per workgroup
    HIPCHK(hipModuleGetGlobal(&gp, &gsz, h->module, "colloc_jac_gemm"));
    HIPCHK(hipMemcpy(&h->jac_gemm, (void*)gp, sizeof(int), hipMemcpyDeviceToHost));
  }
  HIPCHK(hipModuleGetGlobal(&gp, &gsz, h->module, "colloc_caps"));
  int caps[8] = {};
  if (gsz < sizeof caps) return fail(CS_E_SHAPE, "module capabilities hold %zu bytes, expected %zu", gsz, sizeof caps);
  HIPCHK(hipMemcpy(caps, (void*)gp, sizeof caps, hipMemcpyDeviceToHost));
  h->mf_ok = caps[0];
  h->pair_threads = caps[1];
  h->h_zero = caps[2];
  h->jac_parts = caps[3] > 0 ? caps[3] : 1;
  h->tab_g = caps[4];
  h->large = caps[5];
  h->l_tiles = caps[6];
  h->l_stiles = caps[7];
  if (h->large) {
    // the large-block series' buffers, once per handle
    const size_t nint = (size_t)(h->n_max - 1), un = (size_t)h->ns * (h->ny + meta[7] + 1);
    if (!h->l_ubuf) TRY(dmalloc(&h->l_ubuf, nint * un));
    if (!h->l_slots) HIPCHK(hipMalloc((void**)&h->l_slots, nint * kLargeSlot * sizeof(unsigned long long)));
    if (!h->l_active) HIPCHK(hipMalloc((void**)&h->l_active, (kLargeMaxTerms + 1) * sizeof(unsigned long long)));
    const size_t nfin = (size_t)(2 * h->ny + h->np);
    if (!h->l_fin) TRY(dmalloc(&h->l_fin, (size_t)h->bmax * nfin * (nfin + 1)));
  }
  // the static stage-Jacobian entries are the new module's constants: the
  // buffer holds the previous module's until they are written again
  h->jac_static_bits = 0;
  // the condensation record keeps the columns [dy_i | active dp | 1]
  const int npa = meta[7];
  if (npa < 0 || npa > h->np) return fail(CS_E_SHAPE, "module has %d active parameters of %d", npa, h->np);
  const size_t recn = (size_t)(h->ns + h->m * h->ny + h->m * h->nz) * (h->ny + npa + 1);
  if (recn != h->recn) {
    if (h->rec) hipFree(h->rec);
    h->rec = nullptr;
    h->recn = 0;
    TRY(dmalloc(&h->rec, (size_t)(h->n_max - 1) * recn));
    h->recn = recn;
  }
  h->npa = npa;
  // the new module's constant row blocks (its active parameters may differ)
  return fill_rows_constant(h);
}

int cs_set_mesh(cs_handle* h, const double* t, int N) {
  if (!h || !t) return fail(CS_E_ARG, "null argument");
  return set_batch_mesh(h, 1, &N, t);
}

int cs_set_batch_mesh(cs_handle* h, int B, const int* nodes, const double* t) {
  if (!h || !t || !nodes) return fail(CS_E_ARG, "null argument");
  return set_batch_mesh(h, B, nodes, t);
}

int cs_set_constants(cs_handle* h, const double* c, int nc) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (nc != h->nc) return fail(CS_E_SHAPE, "nc=%d, handle has %d", nc, h->nc);
  if (nc == 0) return CS_OK;
  if (!c) return fail(CS_E_ARG, "null argument");
  for (int b = 0; b < h->bmax; ++b) TRY(copy_in(h, h->c + (size_t)b * nc, c, (size_t)nc));
  return sync(h);
}

int cs_set_batch_constants(cs_handle* h, const double* c, int B, int nc) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (nc != h->nc) return fail(CS_E_SHAPE, "nc=%d, handle has %d", nc, h->nc);
  if (B < 1 || B > h->bmax) return fail(CS_E_SHAPE, "B=%d outside [1, %d]", B, h->bmax);
  if (nc == 0) return CS_OK;
  if (!c) return fail(CS_E_ARG, "null argument");
  TRY(copy_in(h, h->c, c, (size_t)B * nc));
  return sync(h);
}

int cs_set_state(cs_handle* h, const double* y, const double* y_dot, const double* z_tilde, const double* p) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  const size_t nint = h->nint;
  TRY(copy_in(h, h->y(), y, (size_t)h->N * h->ny));
  TRY(copy_in(h, h->yd(), y_dot, nint * h->m * h->ny));
  TRY(copy_in(h, h->zt(), z_tilde, nint * h->m * h->nz));
  TRY(copy_in(h, h->p(), p, (size_t)h->B * h->np));
  return sync(h);
}

int cs_get_state(cs_handle* h, double* y, double* y_dot, double* z_tilde, double* p) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  const size_t nint = h->nint;
  TRY(copy_out(h, y, h->y(), (size_t)h->N * h->ny));
  TRY(copy_out(h, y_dot, h->yd(), nint * h->m * h->ny));
  TRY(copy_out(h, z_tilde, h->zt(), nint * h->m * h->nz));
  TRY(copy_out(h, p, h->p(), (size_t)h->B * h->np));
  return sync(h);
}

int cs_form_initial_input(cs_handle* h, const double* y0, const double* z0, const double* p0) {
  if (!h || !y0 || (!z0 && h->nz) || (!p0 && h->np)) return fail(CS_E_ARG, "null argument");
  TRY(need_mesh(h));
  const size_t N = h->N;
  TRY(copy_in(h, h->y0buf, y0, N * h->ny));
  TRY(copy_in(h, h->y0buf + N * h->ny, z0, N * h->nz));
  TRY(copy_in(h, h->y0buf + N * (h->ny + h->nz), p0, (size_t)h->B * h->np));
  h->init_N = h->N;
  return init_from_saved(h);
}

int cs_form_initial_input_device(cs_handle* h, const double* y0, const double* z0, const double* p0) {
  if (!h || !y0 || (!z0 && h->nz) || (!p0 && h->np)) return fail(CS_E_ARG, "null argument");
  TRY(need_mesh(h));
  const size_t N = h->N;
  HIPCHK(hipMemcpyAsync(h->y0buf, y0, N * h->ny * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  if (h->nz)
    HIPCHK(hipMemcpyAsync(h->y0buf + N * h->ny, z0, N * h->nz * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  if (h->np)
    HIPCHK(hipMemcpyAsync(h->y0buf + N * (h->ny + h->nz), p0, (size_t)h->B * h->np * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
  h->init_N = h->N;
  return init_from_saved(h);
}

int cs_restart(cs_handle* h) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (h->init_N != h->N || h->N < 2) return fail(CS_E_ARG, "no initial input for the current mesh");
  return init_from_saved(h);
}

}  // extern "C"

namespace {
int set_batch_mesh(cs_handle* h, int B, const int* nodes, const double* t) {
  if (B < 1 || B > h->bmax) return fail(CS_E_SHAPE, "B=%d outside [1, %d]", B, h->bmax);
  std::vector<int> noff(B + 1, 0);
  long long total = 0;
  for (int b = 0; b < B; ++b) {
    if (nodes[b] < 2) return fail(CS_E_SHAPE, "instance %d has %d nodes (at least 2)", b, nodes[b]);
    total += nodes[b];
    if (total > h->n_max) return fail(CS_E_SHAPE, "N=%lld nodes in total, handle holds %d", total, h->n_max);
    noff[b + 1] = (int)total;
  }
  const int N = (int)total, nint = N - B;
  std::vector<int> iinst(nint), ninst(N);
  for (int b = 0; b < B; ++b) {
    for (int n = noff[b]; n < noff[b + 1]; ++n) ninst[n] = b;
    for (int i = noff[b] - b; i < noff[b + 1] - b - 1; ++i) iinst[i] = b;
  }
  h->N = N;
  h->B = B;
  h->nint = nint;
  h->hnoff = noff;
  h->init_N = 0;
  ++h->mesh_version;
  TRY(copy_in(h, h->t, t, (size_t)N));
  TRY(upload_ints(h, h->noff, noff));
  TRY(upload_ints(h, h->iinst, iinst));
  TRY(upload_ints(h, h->ninst, ninst));
  TRY(sync(h));
  TRY(build_cr_schedule(h));
  return fill_rows_constant(h);
}

int init_from_saved(cs_handle* h) {
  const size_t N = h->N;
  double* zb = h->y0buf + N * h->ny;
  HIPCHK(hipMemcpyAsync(h->y(), h->y0buf, N * h->ny * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  if (h->np)
    HIPCHK(hipMemcpyAsync(h->p(), h->y0buf + N * (h->ny + h->nz), (size_t)h->B * h->np * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
  int nint = h->nint;
  double* y = h->y();
  double* p = h->p();
  double* yd = h->yd();
  double* zt = h->zt();
  Batch bt = h->batch(nullptr);
  void* args[] = {&nint, &bt, &y, &zb, &p, &h->c, &yd, &zt};
  TRY(launch(h, F_INIT, h->tab_g ? (unsigned)((nint * h->m + h->tab_g - 1) / h->tab_g)
                                  : (unsigned)((nint * h->m + 255) / 256), 256, args));
  return sync(h);
}

// The damped Newton loop of every instance of the batch (solver:109-174):
// each instance stops when its own ||f_q||_inf < tol, and its line search
// takes the first step 2^-k (k < 20) that strictly decreases its own norm.
// The kernels of an iteration run over the instances still iterating (act),
// the line-search residuals over those still searching (srch); a step is
// accepted by adding it to the state of its instance only.  A line search
// that accepts no step leaves the instance's state as it was, and nothing
// else changes between iterations (alpha, mesh and constants are fixed), so
// every later iteration of that instance would repeat this one exactly: the
// instance stops here with the outcome those iterations would give (loop
// counter max_iter - 1, the same norm, the norm history continued).
// Launch lists of the intervals of the active instances (set) and of the
// line-searching ones (search), each in its own buffer and re-uploaded only
// when its set changed; none (every interval) when all instances are in the
// set.  The handle's current list is the one selected last.
struct LaunchList {
  cs_handle* h;
  struct Slot {
    int* buf;
    std::vector<int> last;
    int n = 0;   // list length; 0: no list
  };
  Slot act_, srch_;
  explicit LaunchList(cs_handle* h_) : h(h_), act_{h_->alist_buf, {}, 0}, srch_{h_->slist_buf, {}, 0} {}
  ~LaunchList() {   // other entry points launch over every interval
    h->alist = nullptr;
    h->nalist = 0;
  }
  int fill(Slot& sl, const std::vector<int>& in) {
    if (in != sl.last) {
      sl.last = in;
      int nin = 0;
      for (int a : in) nin += a;
      sl.n = 0;
      if (nin != h->B && nin != 0) {
        std::vector<int> list;
        for (int b = 0; b < h->B; ++b)
          if (in[b])
            for (int i = h->hnoff[b] - b; i < h->hnoff[b + 1] - b - 1; ++i) list.push_back(i);
        TRY(upload_ints(h, sl.buf, list));
        sl.n = (int)list.size();
      }
    }
    h->alist = sl.n ? sl.buf : nullptr;
    h->nalist = sl.n;
    return CS_OK;
  }
  int set(const std::vector<int>& active) { return fill(act_, active); }
  int search(const std::vector<int>& searching) {
    if (searching == act_.last) return fill(act_, searching);   // the same set: its list
    return fill(srch_, searching);
  }
};

// patience > 0: an instance whose norm, from iteration `patience` on, falls
// by less than a factor min_ratio over the last `window` iterations stops
// there (reported with its loop counter at that point and not converged);
// off for the reference's loop (patience = 0).
int newton_graph(cs_handle* h, double alpha, double tol, int max_iter, int* iter_time, double* norm, double* hist,
                 int hist_len, int patience, double min_ratio, int window);

int newton(cs_handle* h, double alpha, double tol, int max_iter, int* iter_time, double* norm, double* hist,
           int hist_len, int patience = 0, double min_ratio = 0.0, int window = 1) {
  if (h->use_graph) return newton_graph(h, alpha, tol, max_iter, iter_time, norm, hist, hist_len, patience, min_ratio,
                                        window);
  const int B = h->B;
  std::vector<double> nrm(B, 0.0), nn(B, 0.0), nxt(B, 0.0), step(B, 0.0), taken(B, 0.0);
  // the last window + 1 norms of each instance (ring), for the progress test
  const int ring = patience > 0 ? window + 1 : 1;
  std::vector<double> past((size_t)B * ring, 0.0);
  std::vector<int> gave_up(B, -1);
  std::vector<int> active(B, 1), search(B, 0), done(B, -1);
  LaunchList lst(h);
  for (int it = 0; it < max_iter; ++it) {
    // The residual of the iterate: evaluated at the start; later iterates
    // are the accepted trial states, whose residual (norm, and the stage
    // states, f_a, f_b and r_bc the next direction reads) the line search
    // evaluated with the same kernel and the same inputs (the accepted
    // update and the trial step are one axpy with the same step).
    std::vector<double> cur(B, 0.0);
    if (it == 0) {
      TRY(upload_ints(h, h->act, active));
      TRY(lst.set(active));
      TRY(residual(h, h->state, alpha, true, h->act, cur.data()));
    } else {
      cur = nxt;
    }
    int nact = 0;
    for (int b = 0; b < B; ++b) {
      if (!active[b]) continue;
      nrm[b] = cur[b];
      if (hist && it < hist_len) hist[(size_t)b * hist_len + it] = cur[b];
      if (cur[b] < tol) {
        active[b] = 0;
        done[b] = it;
      } else if (patience > 0 && it >= patience && it >= window &&
                 !(cur[b] * min_ratio <= past[(size_t)b * ring + (it - window) % ring])) {
        active[b] = 0;      // too slow: not converging the way a good start does
        gave_up[b] = it;
      } else {
        ++nact;
      }
      past[(size_t)b * ring + it % ring] = cur[b];
    }
    if (nact == 0) break;
    TRY(upload_ints(h, h->act, active));
    TRY(lst.set(active));
    TRY(jacobian(h, alpha, h->act));
    long long ncond = 0;
    for (int b = 0; b < B; ++b)
      if (active[b]) ncond += h->hnoff[b + 1] - h->hnoff[b] - 1;
    TRY(condense(h, h->act, ncond));
    TRY(babd(h, h->act));
    TRY(recover(h, h->act));
    for (int r = 0; r < h->refine; ++r) TRY(refine_direction(h, h->act, ncond));
    for (int b = 0; b < B; ++b) {
      search[b] = active[b];
      step[b] = active[b] ? 1.0 : 0.0;
      taken[b] = 0.0;
    }
    bool any_taken = false;
    for (int ls = 0; ls < 20; ++ls) {
      TRY(upload_ints(h, h->srch, search));
      TRY(lst.search(search));
      TRY(upload_doubles(h, h->steps, step));
      TRY(axpy_state(h, h->state, h->dir, h->trial));
      TRY(residual(h, h->trial, alpha, true, h->srch, nn.data()));
      int left = 0;
      for (int b = 0; b < B; ++b) {
        if (!search[b]) continue;
        if (nn[b] < nrm[b]) {
          taken[b] = step[b];
          nxt[b] = nn[b];
          any_taken = true;
          search[b] = 0;
          step[b] = 0.0;
        } else {
          step[b] *= 0.5;
          ++left;
        }
      }
      if (left == 0) break;
    }
    TRY(lst.set(active));
    if (any_taken) {
      TRY(upload_doubles(h, h->steps, taken));
      TRY(axpy_state(h, h->state, h->dir, h->state));
    }
    for (int b = 0; b < B; ++b) {
      if (!active[b] || taken[b] != 0.0) continue;
      active[b] = 0;   // stalled: done[b] stays -1
      if (hist)
        for (int k = it + 1; k < std::min(max_iter, hist_len); ++k) hist[(size_t)b * hist_len + k] = nrm[b];
    }
  }
  for (int b = 0; b < B; ++b) {
    // the reference's loop variable stops at max_iter - 1
    if (iter_time) iter_time[b] = done[b] >= 0 ? done[b] : (gave_up[b] >= 0 ? gave_up[b] : max_iter - 1);
    if (norm) norm[b] = nrm[b];
  }
  return sync(h);
}
// ------------------------------------------- device-resident Newton loop
// The same loop as newton() (identical arithmetic and decisions) with every
// per-instance decision made by kernels (colloc_nt_*) and the launch list
// of the iterating instances' intervals built on the device; the work of
// an iteration is captured once per mesh and settings as three HIP graphs:
//   direction: residual of the iterate (first iteration; the later ones
//     start from the accepted trial step's residual, as newton() does), the
//     convergence / progress check, the launch list, Jacobian, condensation,
//     BABD, recovery, and the first trial step of the line search;
//   trial: one more trial step (replayed while instances are searching);
//   accept: the accepted steps added to the state, the iteration counter.
// The host reads three counters per replay of the direction graph (and per
// extra trial) instead of synchronising for every residual and the
// condensation's counters.
NtState nt_state(cs_handle* h, double tol, int max_iter, int patience, double min_ratio, int window, int hl) {
  const size_t bm = (size_t)h->bmax;
  NtState st;
  st.nrm = h->nt_nrm;
  st.taken = h->nt_taken;
  st.next = h->nt_next;
  st.trials = h->nt_ints;
  st.done = h->nt_ints + bm;
  st.gave = h->nt_ints + 2 * bm;
  st.lofs = h->nt_ints + 3 * bm;
  st.it = h->nt_ints + 4 * bm;
  st.nlist = h->nt_ints + 4 * bm + 1;
  st.hist = h->nt_hist;
  st.ring = h->nt_ring;
  st.cnt = h->nt_cnt;
  st.par = h->nt_par;
  NtParams& P = h->nt_par_host;
  P.hl = hl;
  P.window = window;
  P.patience = patience;
  P.max_iter = max_iter;
  P.max_trials = 20;   // solver:64
  P.pad = 0;
  P.tol = tol;
  P.min_ratio = min_ratio;
  return st;
}

int nt_list(cs_handle* h, const NtState& st) {
  Batch bt = h->batch(h->act);
  NtState s2 = st;
  void* a1[] = {&bt, &s2};
  TRY(launch(h, F_NT_SCAN, 1, 1024, a1));
  int nint = h->nint;
  int* list = h->dlist;
  void* a2[] = {&nint, &bt, &s2, &list};
  return launch(h, F_NT_FILL, (unsigned)((nint + 255) / 256), 256, a2);
}

// One trial step of the line search: trial state, its residual norms, the
// decision (the still-searching count in cnt[1]).
int nt_trial(cs_handle* h, const NtState& st, double alpha) {
  TRY(axpy_state(h, h->state, h->dir, h->trial));
  TRY(residual_to(h, h->trial, alpha, true, h->srch, h->scal));
  TRY(fill_u64(h, h->nt_cnt + 1, 1));
  NtState s2 = st;
  int B = h->B;
  unsigned long long* slot = h->scal;
  void* a[] = {&B, &s2, &slot, &h->srch, &h->steps};
  return launch(h, F_NT_LS, (unsigned)((B + 255) / 256), 256, a);
}

int nt_capture(cs_handle* h, Graph& g, const std::function<int()>& body) {
  if (g.exec) {
    HIPCHK(hipGraphExecDestroy(g.exec));
    g.exec = nullptr;
  }
  g.prof.clear();
  g.used = 0;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
  h->capturing = true;
  h->cap_prof = h->prof ? &g : nullptr;
  const int r = body();
  h->cap_prof = nullptr;
  h->capturing = false;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &graph);
  if (r != CS_OK) {
    if (e == hipSuccess && graph) hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) return fail(CS_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(CS_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
  return CS_OK;
}

int nt_replay(cs_handle* h, Graph& g, unsigned long long* cnt) {
  HIPCHK(hipGraphLaunch(g.exec, h->stream));
  HIPCHK(hipMemcpyAsync(cnt, h->nt_cnt, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  unsigned long long ts[kStampSlots];
  if (!g.prof.empty())
    HIPCHK(hipMemcpyAsync(ts, h->stamps + g.base, g.used * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (const StampPair& pr : g.prof) {
    h->prof_ms[pr.k] += (double)(ts[pr.s1] - ts[pr.s0]) / h->stamp_khz;
    h->prof_n[pr.k] += 1;
  }
  return CS_OK;
}

int newton_graph(cs_handle* h, double alpha, double tol, int max_iter, int* iter_time, double* norm, double* hist,
                 int hist_len, int patience, double min_ratio, int window) {
  const int B = h->B;
  const int hl = std::max(hist_len, 1);
  const int rs = patience > 0 ? window + 1 : 1;
  // history and progress ring (grown as needed; outside any graph)
  if ((size_t)B * hl > h->nt_hist_cap) {
    if (h->nt_hist) hipFree(h->nt_hist);
    h->nt_hist = nullptr;
    h->nt_hist_cap = 0;
    TRY(dmalloc(&h->nt_hist, (size_t)h->bmax * hl));
    h->nt_hist_cap = (size_t)h->bmax * hl;
    h->graph_key.clear();
  }
  if ((size_t)B * rs > h->nt_ring_cap) {
    if (h->nt_ring) hipFree(h->nt_ring);
    h->nt_ring = nullptr;
    h->nt_ring_cap = 0;
    TRY(dmalloc(&h->nt_ring, (size_t)h->bmax * rs));
    h->nt_ring_cap = (size_t)h->bmax * rs;
    h->graph_key.clear();
  }
  const NtState st = nt_state(h, tol, max_iter, patience, min_ratio, window, hl);
  // start: every instance iterating, nothing converged, iteration 0, the
  // list of all intervals; static Jacobian entries in place
  {
    std::vector<int> ints(4 * (size_t)h->bmax + 2, 0);
    for (int b = 0; b < B; ++b) {
      ints[h->bmax + b] = -1;
      ints[2 * h->bmax + b] = -1;
    }
    HIPCHK(hipMemcpyAsync(h->nt_ints, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
    TRY(upload_ints(h, h->act, std::vector<int>(B, 1)));
    HIPCHK(hipMemcpyAsync(h->nt_par, &h->nt_par_host, sizeof(NtParams), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));   // the host vectors above are temporaries
  }
  h->jac_bits = h->dir_bits;
  TRY(jacobian_static(h));
  TRY(fill_u64(h, h->scal + 2 * h->bmax, 3));   // condensation counters, summed over the solve
  TRY(nt_list(h, st));
  // (re)capture when the mesh or any setting baked into the graphs changed
  // (the solve's tolerance, budget and progress test are read from device
  // memory, so they are not part of it)
  const std::vector<double> key = {(double)h->mesh_version, (double)B, (double)h->N, alpha, (double)h->dir_bits,
                                   (double)h->cond_kind, (double)h->prof, (double)h->subst_spans, (double)h->refine,
                                   (double)h->babd_bits,
                                   (double)(size_t)h->nt_hist, (double)(size_t)h->nt_ring};
  h->alist = h->dlist;
  h->nalist = 0;
  h->nalistp = st.nlist;
  struct Unlist {
    cs_handle* h;
    ~Unlist() { h->alist = nullptr; h->nalist = 0; h->nalistp = nullptr; }
  } unlist{h};
  if (key != h->graph_key) {
    h->graph_key.clear();
    // the first iteration evaluates the residual of the iterate; the later
    // ones start from the accepted trial's (see newton())
    for (const bool first : {true, false})
      TRY(nt_capture(h, first ? h->g_dir0 : h->g_dir, [&]() -> int {
        TRY(fill_u64(h, h->nt_cnt, 4));
        if (first) TRY(residual_to(h, h->state, alpha, true, h->act, h->scal));
        int Bv = B;
        NtState s2 = st;
        unsigned long long* slot = first ? h->scal : nullptr;
        void* a[] = {&Bv, &s2, &slot, &h->act, &h->srch, &h->steps};
        TRY(launch(h, F_NT_CHECK, (unsigned)((B + 255) / 256), 256, a));
        TRY(nt_list(h, st));
        TRY(jacobian(h, alpha, h->act));
        TRY(condense(h, h->act, 0));
        TRY(babd(h, h->act));
        TRY(recover(h, h->act));
        for (int r = 0; r < h->refine; ++r) TRY(refine_direction(h, h->act, 0));
        return nt_trial(h, st, alpha);
      }));
    TRY(nt_capture(h, h->g_trial, [&]() -> int { return nt_trial(h, st, alpha); }));
    TRY(nt_capture(h, h->g_acc, [&]() -> int {
      int Bv = B;
      NtState s2 = st;
      void* a[] = {&Bv, &s2, &h->act, &h->steps};
      TRY(launch(h, F_NT_ACCEPT, (unsigned)((B + 255) / 256), 256, a));
      TRY(axpy_state(h, h->state, h->dir, h->state));
      void* t[] = {&s2};
      return launch(h, F_NT_TICK, 1, 64, t);
    }));
    h->graph_key = key;
  }
  unsigned long long cnt[3] = {0, 0, 0};
  for (int it = 0; it < max_iter; ++it) {
    TRY(nt_replay(h, it == 0 ? h->g_dir0 : h->g_dir, cnt));
    if (cnt[0] == 0) break;
    h->cond_intervals += cnt[2] * (1ull + (unsigned long long)h->refine);
    while (cnt[1] > 0) TRY(nt_replay(h, h->g_trial, cnt));
    HIPCHK(hipGraphLaunch(h->g_acc.exec, h->stream));
  }
  // results, as newton() reports them
  std::vector<int> ints(4 * (size_t)h->bmax + 2);
  std::vector<double> nrm(B);
  std::vector<unsigned long long> ctr(3);
  HIPCHK(hipMemcpyAsync(ints.data(), h->nt_ints, ints.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(nrm.data(), h->nt_nrm, B * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(ctr.data(), h->scal + 2 * h->bmax, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        h->stream));
  std::vector<double> dh;
  if (hist && hist_len > 0) {
    dh.resize((size_t)B * hl);
    HIPCHK(hipMemcpyAsync(dh.data(), h->nt_hist, dh.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  h->cond_fallback += ctr[2];
  h->cond_terms += ctr[1];
  const int last_all = std::min(max_iter, hist_len) - 1;
  for (int b = 0; b < B; ++b) {
    const int done = ints[h->bmax + b], gave = ints[2 * h->bmax + b];
    if (iter_time) iter_time[b] = done >= 0 ? done : (gave >= 0 ? gave : max_iter - 1);
    if (norm) norm[b] = nrm[b];
    if (!dh.empty()) {
      // the entries newton() writes: up to the stopping iteration, or the
      // whole budget (run out, or a stall continued at the same norm)
      const int last = done >= 0 ? std::min(done, hist_len - 1) : (gave >= 0 ? std::min(gave, hist_len - 1) : last_all);
      for (int k = 0; k <= last; ++k) hist[(size_t)b * hist_len + k] = dh[(size_t)b * hl + k];
    }
  }
  return CS_OK;
}
}  // namespace

extern "C" {

int cs_compute_f_q(cs_handle* h, double alpha, double* norm_inf, double* y_tilde, double* f_a, double* f_b,
                   double* r_bc) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  std::vector<double> nrm(h->B);
  TRY(residual(h, h->state, alpha, true, nullptr, nrm.data()));
  if (norm_inf) std::memcpy(norm_inf, nrm.data(), nrm.size() * sizeof(double));
  const size_t nint = h->nint;
  TRY(copy_out(h, y_tilde, h->ytil, nint * h->m * h->ny));
  TRY(copy_out(h, f_a, h->fa, nint * h->m * (h->ny + h->nz)));
  TRY(copy_out(h, f_b, h->fb, nint * h->ny));
  TRY(copy_out(h, r_bc, h->rbc, (size_t)h->B * h->nrb));
  return sync(h);
}

int cs_compute_f_q_noalloc(cs_handle* h, double alpha, double* norm_inf) {
  if (!h || !norm_inf) return fail(CS_E_ARG, "null argument");
  TRY(need_mesh(h));
  if (h->B > kScratchSlots)
    return fail(CS_E_SHAPE, "B=%d instances: cs_compute_f_q_noalloc holds at most %d", h->B, kScratchSlots);
  // slots cleared by a fill kernel, the norms copied straight into the
  // caller's array: no allocation on the device or the host
  TRY(residual_to(h, h->state, alpha, true, nullptr, h->scratch_slots));
  HIPCHK(hipMemcpyAsync(norm_inf, h->scratch_slots, (size_t)h->B * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  return sync(h);
}

int cs_construct_jacobian(cs_handle* h, double alpha, double* Df, double* Dg, double* Dr) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  TRY(jacobian(h, alpha, nullptr));
  const size_t nst = (size_t)h->nint * h->m;
  const size_t nf = (size_t)h->ny * h->nj, ng = (size_t)h->nz * h->nj;
  if (Df || Dg) {
    std::vector<double> buf(nst * h->jstage);
    if (h->jac_bits == 32) {   // stored in float32: widen
      std::vector<float> fbuf(buf.size());
      HIPCHK(hipMemcpyAsync(fbuf.data(), h->jac, fbuf.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
      TRY(sync(h));
      for (size_t e = 0; e < buf.size(); ++e) buf[e] = fbuf[e];
    } else {
      TRY(copy_out(h, buf.data(), h->jac, buf.size()));
    }
    TRY(sync(h));
    for (size_t s = 0; s < nst; ++s) {
      if (Df) std::memcpy(Df + s * nf, buf.data() + s * h->jstage, nf * sizeof(double));
      if (Dg && ng) std::memcpy(Dg + s * ng, buf.data() + s * h->jstage + nf, ng * sizeof(double));
    }
  }
  TRY(copy_out(h, Dr, h->Dr, (size_t)h->B * h->drsz));
  return sync(h);
}

int cs_reduce_jacobian(cs_handle* h, double* A, double* H, double* b) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  TRY(condense(h, nullptr, h->nint));
  if (A || H || b) {
    const size_t nint = h->nint;
    std::vector<double> buf(nint * h->ny * h->crw);
    TRY(copy_out(h, buf.data(), h->rows0, buf.size()));
    TRY(sync(h));
    for (size_t i = 0; i < nint; ++i)
      for (int r = 0; r < h->ny; ++r) {
        const double* row = buf.data() + (i * h->ny + r) * h->crw;
        if (A) std::memcpy(A + (i * h->ny + r) * h->ny, row, h->ny * sizeof(double));
        if (H && h->np) {
          if (h->h_zero) std::memset(H + (i * h->ny + r) * h->np, 0, h->np * sizeof(double));
          else std::memcpy(H + (i * h->ny + r) * h->np, row + 2 * h->ny, h->np * sizeof(double));
        }
        if (b) b[i * h->ny + r] = row[h->crw - 1];
      }
  }
  return sync(h);
}

int cs_solve_babd(cs_handle* h, double* delta_y, double* delta_p) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  TRY(babd(h, nullptr));
  TRY(copy_out(h, delta_y, h->dir + h->st_y, (size_t)h->N * h->ny));
  TRY(copy_out(h, delta_p, h->dir + h->st_p, (size_t)h->B * h->np));
  return sync(h);
}

int cs_recover_delta_k(cs_handle* h, double* delta_y_dot, double* delta_z_tilde) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  TRY(recover(h, nullptr));
  const size_t nint = h->nint;
  TRY(copy_out(h, delta_y_dot, h->dir + h->st_yd, nint * h->m * h->ny));
  TRY(copy_out(h, delta_z_tilde, h->dir + h->st_z, nint * h->m * h->nz));
  return sync(h);
}

int cs_newton(cs_handle* h, double alpha, double tol, int max_iter, int* iter_time, double* norm, double* norms,
              int hist_len) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  if (max_iter < 1) return fail(CS_E_ARG, "max_iter must be >= 1");
  if (hist_len < 0) return fail(CS_E_ARG, "hist_len must be >= 0");
  return newton(h, alpha, tol, max_iter, iter_time, norm, norms, hist_len);
}

int cs_newton_progress(cs_handle* h, double alpha, double tol, int max_iter, int patience, double min_ratio,
                       int* iter_time, double* norm, double* norms, int hist_len) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  if (max_iter < 1) return fail(CS_E_ARG, "max_iter must be >= 1");
  if (hist_len < 0) return fail(CS_E_ARG, "hist_len must be >= 0");
  if (patience < 1 || !(min_ratio >= 1.0)) return fail(CS_E_ARG, "patience %d (>= 1), min_ratio %g (>= 1)", patience,
                                                       min_ratio);
  return newton(h, alpha, tol, max_iter, iter_time, norm, norms, hist_len, patience, min_ratio);
}

int cs_newton_stall(cs_handle* h, double alpha, double tol, int max_iter, int start, int window, double min_ratio,
                    int* iter_time, double* norm, double* norms, int hist_len) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  if (max_iter < 1) return fail(CS_E_ARG, "max_iter must be >= 1");
  if (hist_len < 0) return fail(CS_E_ARG, "hist_len must be >= 0");
  if (start < 1 || window < 1 || window > max_iter || !(min_ratio >= 1.0))
    return fail(CS_E_ARG, "start %d (>= 1), window %d (1..max_iter), min_ratio %g (>= 1)", start, window, min_ratio);
  return newton(h, alpha, tol, max_iter, iter_time, norm, norms, hist_len, start, min_ratio, window);
}

int cs_segment_residual_ex(cs_handle* h, double alpha, double tol, int residual_type, int scale_by_time,
                           int relative, double* residual_out, double* max_residual) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (residual_type < 0 || residual_type > 2) return fail(CS_E_ARG, "residual_type %d (0, 1 or 2)", residual_type);
  TRY(need_mesh(h));
  unsigned long long* slot = h->scal + h->bmax;
  TRY(fill_u64(h, slot, h->B));
  int nint = h->nint;
  double* y = h->y();
  double* yd = h->yd();
  double* zt = h->zt();
  double* p = h->p();
  int st = scale_by_time ? 1 : 0, rel = relative ? 1 : 0;
  Batch bt = h->batch(nullptr);
  void* args[] = {&nint, &bt, &h->t, &y, &yd, &zt, &p, &h->c, &alpha, &tol, &residual_type, &st, &rel,
                  &h->segres, &slot};
  TRY(launch(h, F_SEGRES, h->tab_g ? (unsigned)(nint + (h->B + h->tab_g - 1) / h->tab_g)
                                    : (unsigned)((nint + h->B + 255) / 256), 256, args));
  std::vector<double> mr(h->B);
  TRY(read_slots(h, slot, mr.data()));
  if (max_residual) std::memcpy(max_residual, mr.data(), mr.size() * sizeof(double));
  TRY(copy_out(h, residual_out, h->segres, (size_t)h->N));
  return sync(h);
}

int cs_segment_residual(cs_handle* h, double alpha, double tol, double* residual_out, double* max_residual) {
  return cs_segment_residual_ex(h, alpha, tol, 1, 1, 0, residual_out, max_residual);
}

int cs_recover_solution(cs_handle* h, double* z) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  int n = h->N;
  double* zt = h->zt();
  Batch bt = h->batch(nullptr);
  void* args[] = {&n, &bt, &zt, &h->zbuf};
  TRY(launch(h, F_RECZ, (unsigned)((n + 255) / 256), 256, args));
  TRY(copy_out(h, z, h->zbuf, (size_t)n * h->nz));
  return sync(h);
}

int cs_node_solution_device(cs_handle* h, const double** y, const double** z, const double** p) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  int n = h->N;
  double* zt = h->zt();
  Batch bt = h->batch(nullptr);
  void* args[] = {&n, &bt, &zt, &h->zbuf};
  TRY(launch(h, F_RECZ, (unsigned)((n + 255) / 256), 256, args));
  TRY(sync(h));
  if (y) *y = h->y();
  if (z) *z = h->zbuf;
  if (p) *p = h->p();
  return CS_OK;
}

int cs_set_direction_precision(cs_handle* h, int bits) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (bits != 32 && bits != 64) return fail(CS_E_ARG, "precision %d bits (32 or 64)", bits);
  h->dir_bits = bits;
  return CS_OK;
}

int cs_set_newton_graph(cs_handle* h, int on) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (on && h->large) return fail(CS_E_UNSUPPORTED, "the large-block kernels run the host-driven Newton loop");
  h->use_graph = on != 0;
  return CS_OK;
}

int cs_set_babd_precision(cs_handle* h, int bits) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (bits != 32 && bits != 64) return fail(CS_E_ARG, "BABD precision %d (32 or 64)", bits);
  h->babd_bits = bits;
  return CS_OK;
}

int cs_set_refinement(cs_handle* h, int steps) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (steps < 0 || steps > 8) return fail(CS_E_ARG, "refinement steps %d (0..8)", steps);
  h->refine = steps;
  return CS_OK;
}

int cs_get_direction(cs_handle* h, double* delta_y, double* delta_y_dot, double* delta_z_tilde, double* delta_p) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(need_mesh(h));
  const size_t nint = h->nint;
  TRY(copy_out(h, delta_y, h->dir + h->st_y, (size_t)h->N * h->ny));
  TRY(copy_out(h, delta_y_dot, h->dir + h->st_yd, nint * h->m * h->ny));
  TRY(copy_out(h, delta_z_tilde, h->dir + h->st_z, nint * h->m * h->nz));
  TRY(copy_out(h, delta_p, h->dir + h->st_p, (size_t)h->B * h->np));
  return sync(h);
}

int cs_refine_direction(cs_handle* h, int steps) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (steps < 0 || steps > 8) return fail(CS_E_ARG, "refinement steps %d (0..8)", steps);
  TRY(need_mesh(h));
  for (int r = 0; r < steps; ++r) TRY(refine_direction(h, nullptr, h->nint));
  return sync(h);
}

int cs_linear_residual(cs_handle* h, double* norm) {
  if (!h || !norm) return fail(CS_E_ARG, "null argument");
  TRY(need_mesh(h));
  TRY(linear_residual(h, nullptr, h->scal, false));
  return read_slots(h, h->scal, norm);
}

int cs_set_condense_kernel(cs_handle* h, int kind) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (kind != 0 && kind != 1) return fail(CS_E_ARG, "condensation kernel %d (0 or 1)", kind);
  h->cond_kind = kind;
  return CS_OK;
}

int cs_set_launch(cs_handle* h, int which, int value) {
  if (!h) return fail(CS_E_ARG, "null handle");
  switch (which) {
    case CS_LAUNCH_CONDENSE_GRID:
      if (value < 0) return fail(CS_E_ARG, "condensation grid %d (>= 0)", value);
      h->cond_grid = value;
      return CS_OK;
    case CS_LAUNCH_RESIDUAL_BLOCK:
      if (value != 64 && value != 128 && value != 256)
        return fail(CS_E_ARG, "residual workgroup %d (64, 128 or 256)", value);
      h->res_block = value;
      return CS_OK;
    case CS_LAUNCH_JACOBIAN_GEMM:
      if (value != 0 && value != 1) return fail(CS_E_ARG, "matrix-core Jacobian switch %d (0 or 1)", value);
      h->jac_gemm_on = value;
      return CS_OK;
    default:
      return fail(CS_E_ARG, "launch setting %d", which);
  }
}

int cs_get_launch(cs_handle* h, int which, int* value) {
  if (!h || !value) return fail(CS_E_ARG, "null argument");
  switch (which) {
    case CS_LAUNCH_CONDENSE_GRID: *value = h->cond_grid; return CS_OK;
    case CS_LAUNCH_RESIDUAL_BLOCK: *value = h->res_block; return CS_OK;
    case CS_LAUNCH_JACOBIAN_GEMM: *value = (h->jac_gemm > 0 && h->jac_gemm_on) ? 1 : 0; return CS_OK;
    case CS_LAUNCH_COMPUTE_UNITS: {
      int dev = 0;
      HIPCHK(hipGetDevice(&dev));
      HIPCHK(hipDeviceGetAttribute(value, hipDeviceAttributeMultiprocessorCount, dev));
      return CS_OK;
    }
    default:
      return fail(CS_E_ARG, "launch setting %d", which);
  }
}

int cs_set_babd_substitution(cs_handle* h, int min_spans) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (min_spans < 0) return fail(CS_E_ARG, "min_spans=%d (>= 0)", min_spans);
  h->subst_spans = min_spans;
  if (h->B >= 1 && h->N >= 2) return build_cr_schedule(h);
  return CS_OK;
}

int cs_batch_info(cs_handle* h, int* B, int* N) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (B) *B = h->B;
  if (N) *N = h->N;
  return CS_OK;
}

int cs_profile(cs_handle* h, int enable) {
  if (!h) return fail(CS_E_ARG, "null handle");
  TRY(sync(h));
  drain_prof(h);
  for (int k = 0; k < CS_KERNEL_COUNT; ++k) { h->prof_ms[k] = 0.0; h->prof_n[k] = 0; }
  h->prof = enable != 0;
  return CS_OK;
}

int cs_profile_get(cs_handle* h, int kernel, double* total_ms, int* launches) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (kernel < 0 || kernel >= CS_KERNEL_COUNT) return fail(CS_E_ARG, "bad kernel id %d", kernel);
  TRY(sync(h));
  drain_prof(h);
  if (total_ms) *total_ms = h->prof_ms[kernel];
  if (launches) *launches = h->prof_n[kernel];
  return CS_OK;
}

int cs_condense_stats(cs_handle* h, long long* fallback_intervals, long long* series_terms,
                      long long* intervals) {
  if (!h) return fail(CS_E_ARG, "null handle");
  if (intervals) *intervals = (long long)h->cond_intervals;
  if (fallback_intervals) *fallback_intervals = (long long)h->cond_fallback;
  if (series_terms) *series_terms = (long long)h->cond_terms;
  return CS_OK;
}

int cs_status(cs_handle* h, int* status) {
  if (!h || !status) return fail(CS_E_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(status, h->status, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemsetAsync(h->status, 0, sizeof(int), h->stream));
  return sync(h);
}

}  // extern "C"
