"""Throughput benchmark: converged BVP solves per second on one MI355X.

Default workload (--workload batch): the batched family, 1024 instances of
the n_x=20, n_u=10 tanh OCP sharing A, B, C (seed 1), each with its own
x0 ~ U[-1,1]^20, T ~ U[0.5, 2] and uniform mesh of Nint ~ U{256..2048}
intervals, m=5 Lobatto IIIA stages.  One step = every instance solved once,
in chunks that fit one handle, each chunk one batch (every launch covers all
its instances).  The default form (--form learned --precision 32): the LQ
solve, the trained network's correction, Newton capped at 12 iterations with
a progress test, then the continuation of the instances it leaves, with the
Newton direction condensed in float32 (residuals, BABD solve and
convergence test in float64).  --form plain: the all-ones guess with the
two-step homotopy in the tanh scale (0 -> 0.5), float64.  A solve counts
only if it converged (residual < 1e-8, max segment residual <= 1).  Beside
it: the same learned pass in float64, one plain batched pass, and the plain
single-instance solve looped over a sample of the same instances.

--workload single: one instance, Nint=4096 (the single-instance
configuration), one complete solve per step.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(condensation, timed live with HIP events on the solver's stream) and the CPU
baseline (the NumPy oracle on a bounded sample of the same workload).
"""
import argparse
import importlib.util
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "collocation-solver-cuda_amd")
FP64_PEAK_TFLOPS = 78.6     # MI355X, FP64 vector = FP64 matrix peak
FP32_PEAK_TFLOPS = 157.3    # MI355X, FP32 peak (matrix cores; vector with packed FP32)
HBM_PEAK_GBS = 8000.0


def load_package():
    if "colloc_amd" in sys.modules:
        return sys.modules["colloc_amd"]
    spec = importlib.util.spec_from_file_location("colloc_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["colloc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def stage_elim_flops(ny, nz, npa, m):
    """Executed FP64 flops of the per-stage eliminations of one interval:
    Gauss-Jordan on [g_z | g_y g_p f_a,g] (2 nz^2 (nz + nv)) and
    T_j = [h_y h_p f_a,h] - h_z G_j (2 ny nv nz) for each of the m stages,
    nv = ny + npa + 1 (npa: parameters the stage functions depend on; the
    other parameter columns are structurally zero and not formed)."""
    nv = ny + npa + 1
    return m * (2 * nz * nz * (nz + nv) + 2 * ny * nv * nz)


def series_term_flops(ny, npa, m):
    """One term of the series solve of the stage system of one interval:
    Y_k = M_k U_k for the m - 1 stage blocks (2 (m-1) ny^2 nv) and the
    mixing U = R + h (A^ x I) Y (2 (m-1)^2 ny nv)."""
    nv = ny + npa + 1
    return 2 * (m - 1) * ny * ny * nv + 2 * (m - 1) ** 2 * ny * nv


def gauss_jordan_flops(ny, npa, m):
    """Gauss-Jordan on the ns x (ns + nv) stage system: ns^2 (ns - 1 + 2 nv)."""
    nv = ny + npa + 1
    ns = (m - 1) * ny
    return ns * ns * (ns - 1 + 2 * nv)


def dense_condense_flops(ny, nz, np_, m):
    """Flops of the reference's dense condensation per interval:
    LU of W (K x K), W^-1 [J V f_a] and D (W^-1 [J V f_a])."""
    K = m * (ny + nz)
    return (2.0 / 3.0) * K ** 3 + 2 * K * K * (ny + np_ + 1) + 2 * m * ny * (ny + np_ + 1)


def measured_traffic(workload, kernel):
    """HBM bytes per interval condensed by `kernel` at this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a
    copy kernel); None when no measurement exists for it."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_m5.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_interval %s:" % kernel):
            return float(line.split(":")[1])
    return None


# condensation kernel -> (bound, peak TFLOP/s).  The FP32 vector peak without
# packed FP32 math (disabled in this build) is half the published 157.3.
CONDENSE_KERNELS = {
    "colloc_condense_mf32": ("mfma", FP32_PEAK_TFLOPS),
    "colloc_condense_ns32": ("valu", FP32_PEAK_TFLOPS / 2),
    "colloc_condense_ns": ("valu", FP64_PEAK_TFLOPS),
    # the large-block series (ny > 64): products on matrix cores in 64 x 64 tiles
    "colloc_l_series_term32": ("mfma", FP32_PEAK_TFLOPS),
    "colloc_l_series_term": ("mfma", FP64_PEAK_TFLOPS),
}


def condense_kernel_name(precision, kind):
    if precision == 64:
        return "colloc_condense_ns"
    return "colloc_condense_mf32" if kind == 0 else "colloc_condense_ns32"


def cpu_baseline(inst, m, nint_sample, nint_target):
    """Oracle (NumPy/LAPACK) plain solve of a family member at a reduced mesh;
    work per solve is linear in Nint at equal Newton counts, so the rate is
    scaled by nint_sample / nint_target."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import families
    from oracle import solver as osolver
    prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)
    hom = [families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=0.0, q=inst.q, c_u=inst.c_u)]
    N = nint_sample + 1
    t = np.linspace(0.0, inst.T, N)
    ny, nz, np_ = prob.size_y, prob.size_z, prob.size_p
    t0 = time.perf_counter()
    res = osolver.collocation_solver(prob, m, t, np.ones((N, ny)), np.ones((N, nz)), np.ones(np_),
                                     tol=inst.tolerance, homotopy=hom)
    dt = time.perf_counter() - t0
    ok = res["success"] == 1
    scale = nint_target / float(nint_sample)
    return {"value": (1.0 / (dt * scale)) if ok else 0.0, "unit": "solves/s",
            "cores": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": "NumPy oracle, full plain solve (homotopy 0->0.5) of one instance of the same family at "
                      "Nint=%d, T=%.3g (%.1f s), rate scaled by %d/%.0f to Nint=%.0f; LAPACK threads = all cores"
                      % (nint_sample, inst.T, dt, nint_sample, nint_target, nint_target),
            "seconds": dt, "converged": ok}


def cpu_protocol(runs=5):
    """The oracle timed as the measurement plan asks: the configuration-0
    solve (LQ, n_x=4, Nint=128, m=4) as the median of `runs` wall-clock runs
    with BLAS on all the process's cores and on one thread, and one Newton
    iteration (residual, direction, full-step trial) of the single-instance
    configuration (n_x=20, m=5) on all cores, timed on a 256-interval mesh
    and scaled to its 4096 intervals (every step of an iteration is linear
    in the interval count)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from threadpoolctl import threadpool_limits
    from colloc_amd import problems
    from oracle import families
    from oracle import solver as osolver
    from oracle import tables as otables
    cores = len(os.sched_getaffinity(0))
    inst, m = problems.config(0)
    prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=None, q=inst.q, c_u=inst.c_u)
    N = inst.nodes
    t = np.linspace(0.0, inst.T, N)

    def cfg0():
        t0 = time.perf_counter()
        r = osolver.collocation_solver(prob, m, t, np.ones((N, prob.size_y)), np.ones((N, prob.size_z)),
                                       np.ones(prob.size_p), tol=inst.tolerance)
        assert r["success"] == 1
        return time.perf_counter() - t0
    out = {}
    for label, nthreads in (("all_cores", cores), ("1_thread", 1)):
        with threadpool_limits(nthreads):
            out["cfg0_solve_s_median_%s" % label] = float(np.median([cfg0() for _ in range(runs)]))
    sample_nint = 256
    inst1, m1 = problems.config(1, nint=sample_nint)
    prob1 = families.TanhFamily(inst1.A, inst1.B, inst1.x0, C=inst1.C, s=inst1.s, q=inst1.q, c_u=inst1.c_u)
    N1 = inst1.nodes
    t1 = np.linspace(0.0, inst1.T, N1)
    rk = otables.lobatto(m1)
    y, yd, zt = osolver.form_initial_input(prob1, m1, N1, np.ones((N1, prob1.size_y)), np.ones((N1, prob1.size_z)),
                                           np.ones(prob1.size_p), rk)
    p = np.ones(prob1.size_p)
    t0 = time.perf_counter()
    f_q, yt, fa, fb, rbc = osolver.compute_f_q(prob1, m1, N1, rk, t1, y, yd, zt, p, 0.1)
    dy, dyd, dz, dp = osolver.newton_direction(prob1, m1, N1, rk, t1, y, yt, zt, p, 0.1, fa, fb, rbc)
    osolver.compute_f_q(prob1, m1, N1, rk, t1, y + dy, yd + dyd, zt + dz, p + dp, 0.1)
    out["cfg1_newton_iteration_s_all_cores"] = (time.perf_counter() - t0) * 4096 / sample_nint
    out["cfg1_sample_nint"] = sample_nint
    out["cores"] = cores
    return out


def condense_roofline(S, before, after, cond_ms, cond_n, ny, nz, np_, npa, m, traffic, kernel):
    """Roofline entry of the condensation from the algorithmic-flop counters
    (stage eliminations, series terms, Gauss-Jordan fallbacks) against the
    peak of the unit the kernel's products run on."""
    bound, peak = CONDENSE_KERNELS[kernel]
    fb = after[0] - before[0]
    terms = after[1] - before[1]
    ncond = after[2] - before[2]
    fl_total = ncond * stage_elim_flops(ny, nz, npa, m) + terms * series_term_flops(ny, npa, m) \
        + fb * gauss_jordan_flops(ny, npa, m)
    fl = fl_total / max(cond_n, 1)
    avg_s = cond_ms / max(cond_n, 1) / 1e3
    achieved = fl / avg_s / 1e12 if avg_s > 0 else 0.0
    dense = dense_condense_flops(ny, nz, np_, m) * ncond / (cond_ms / 1e3) / 1e12 if cond_ms > 0 else 0.0
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic * ncond / max(cond_n, 1) if traffic else None,
            "traffic_bytes_per_interval": traffic,
            "kernel": "%s (+ colloc_condense for intervals the series leaves)" % kernel,
            "flops_per_launch": fl, "avg_launch_ms": avg_s * 1e3, "launches": cond_n,
            "intervals_per_launch": ncond / max(cond_n, 1),
            "series_terms_per_interval": terms / max(ncond, 1), "gauss_jordan_intervals": fb,
            "dense_equivalent_tflops": dense}


def check_gpus(args):
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (a batch would shard by instance, "
              "one chunk per GPU, with no communication inside a solve). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 2 batch passes / 5 solves)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("batch", "single", "precision"), default="batch",
                    help="batch: the metric's workload; single: one n_x=20 instance; precision: the n_x=64 "
                         "instance with the float64 and the FP32 Newton direction")
    ap.add_argument("--nint", type=int, default=None, help="single: override Nint (default 4096)")
    ap.add_argument("--instances", type=int, default=1024, help="batch: instances of the family")
    ap.add_argument("--chunk-nodes", type=int, default=320000,
                    help="batch: most nodes held at once (divided among the --streams handles)")
    ap.add_argument("--streams", type=int, default=3,
                    help="batch: handles solving chunks concurrently, one host thread and HIP stream each")
    ap.add_argument("--loop-sample", type=int, default=64,
                    help="batch: instances solved one by one for the looped plain rate")
    ap.add_argument("--form", choices=("learned", "plain"), default="learned",
                    help="batch: network warm start (default) or the plain guess with continuation")
    ap.add_argument("--model", default=None, help="batch, learned: trained model (default: the family's)")
    ap.add_argument("--plain-pass", type=int, default=1, help="batch, learned: also time one plain batched pass")
    ap.add_argument("--precision", type=int, choices=(32, 64), default=32,
                    help="batch: arithmetic of the Newton direction's condensation (32: the FP32 direction)")
    ap.add_argument("--condense-kernel", type=int, choices=(0, 1), default=0,
                    help="batch, precision 32: series condensation on matrix cores (0) or the vector ALU (1)")
    ap.add_argument("--fp64-pass", type=int, default=1,
                    help="batch, learned, precision 32: also time one learned pass with the float64 direction")
    ap.add_argument("--babd-spans", type=int, default=16,
                    help="batch: substitution pairing while a pair spans <= 1/N of an instance (0: Householder only; "
                         "the library's default is 64)")
    ap.add_argument("--warm-patience", type=int, default=8,
                    help="batch, learned: from this warm iteration on, stop instances whose norm falls by less than "
                         "--warm-ratio per iteration (0: off)")
    ap.add_argument("--warm-ratio", type=float, default=2.0)
    ap.add_argument("--warm-cap", type=int, default=12,
                    help="batch, learned: Newton iterations from the network's guess before the fallback")
    ap.add_argument("--warm-cont", default="8,5,2",
                    help="batch, learned: the instances the capped warm stage leaves continue Newton from there "
                         "for the rest of the iteration budget, with the stall test start,window,ratio (an "
                         "instance stops when its norm fell by less than ratio over the last window "
                         "iterations); 0: no continuation")
    ap.add_argument("--fallback", choices=("none", "plain"), default="plain",
                    help="batch, learned: what the instances the warm-started Newton leaves get: nothing, or the "
                         "plain path again (all-ones guess + continuation)")
    ap.add_argument("--fallback-scales", default="0",
                    help="batch, learned: tanh scales of the plain fallback's continuation stages before the "
                         "final one")
    ap.add_argument("--fallback-stall", default="10,5,2",
                    help="batch, learned: the plain fallback's stall test start,window,ratio; 0: off")
    ap.add_argument("--babd-precision", type=int, choices=(32, 64), default=None,
                    help="stored type of the cyclic reduction's rows and records of the FP32-direction solves "
                         "(32: float32 levels, float64 boundary solve and back substitution); default 64 for "
                         "the batch workload, 32 for the precision workload")
    ap.add_argument("--refine", type=int, default=0,
                    help="batch: iterative refinement steps per Newton direction (float64 linear residual with the "
                         "stored Jacobians, correction from the same direction kernels)")
    ap.add_argument("--graph", type=int, choices=(0, 1), default=0,
                    help="Newton loops device resident, each iteration a HIP-graph replay (same results)")
    ap.add_argument("--cpu-sample", type=int, default=640, help="Nint of the CPU baseline sample (0: skip)")
    ap.add_argument("--autotune", type=int, choices=(0, 1), default=1,
                    help="batch: choose the condensation launch, the residual workgroup and the BABD pairing by "
                         "measurement (autotune.py) after the warmup, outside the timed region")
    args = ap.parse_args()
    check_gpus(args)
    load_package()
    from colloc_amd import build as b
    if not os.path.exists(b.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    if args.babd_precision is None:
        args.babd_precision = 32 if args.workload == "precision" else 64
    if args.workload == "batch":
        line = batch_workload(args)
    elif args.workload == "precision":
        line = precision_workload(args)
    else:
        line = single_workload(args)
    print(json.dumps(line), flush=True)


def S_spans(S, tune_rec, default):
    """The BABD pairing the batch runs with (the autotuner's choice, if it ran)."""
    if tune_rec and "babd_spans" in tune_rec["choices"]:
        return tune_rec["choices"]["babd_spans"]
    return default


def batch_workload(args):
    if args.form == "learned":
        from colloc_amd import learned   # imports torch before libcolloc.so is loaded (shared HIP runtime)
    from colloc_amd import jit, problems, solver
    steps = args.steps if args.steps is not None else 2
    insts = problems.batch_family(args.instances)
    m = 5
    bvp = solver.BvpDae(insts[0].ocp_text())
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    tol, max_iter = bvp.tolerance, bvp.maximum_newton_iterations
    nodes = [inst.nodes for inst in insts]
    # --streams handles solve chunks concurrently, each driven by a host
    # thread of its own on its own HIP stream (the library releases the
    # interpreter inside its calls); the chunks shrink so the handles
    # together hold what one handle of --chunk-nodes would
    nstreams = max(1, args.streams)
    chunk_nodes = args.chunk_nodes // nstreams

    def balanced_chunks(nds, cap, bcap):
        """Consecutive chunks within cap nodes and bcap instances: as many as
        solver.plan_chunks makes, rounded up to a multiple of the handles,
        split at even shares of the nodes so the handles finish together
        (plan_chunks fills each chunk to the cap and leaves a short last one).
        Each instance's result does not depend on its chunk."""
        base = solver.plan_chunks(nds, cap, bcap)
        k = -(-len(base) // nstreams) * nstreams
        tot = float(sum(nds))
        out = [[] for _ in range(k)]
        run = 0
        for i, n in enumerate(nds):
            out[min(k - 1, int((run + 0.5 * n) * k / tot))].append(i)
            run += n
        out = [c for c in out if c]
        if any(sum(nds[i] for i in c) > cap or len(c) > bcap for c in out):
            return base
        return out
    chunks = balanced_chunks(nodes, chunk_nodes, len(insts))
    n_max = max(sum(nodes[i] for i in ch) for ch in chunks)
    b_max = max(len(ch) for ch in chunks)
    handles = [solver.CollocationSolver(bvp, m, n_max=n_max, batch=b_max) for _ in range(nstreams)]
    S = handles[0]

    def each(fn):
        for S_ in handles:
            fn(S_)
    each(lambda S_: S_.set_direction_precision(args.precision))
    each(lambda S_: S_.set_condense_kernel(args.condense_kernel))
    each(lambda S_: S_.set_babd_substitution(args.babd_spans))
    each(lambda S_: S_.set_newton_graph(bool(args.graph)))
    each(lambda S_: S_.set_refinement(args.refine))
    each(lambda S_: S_.set_babd_precision(args.babd_precision))

    def run_chunks(chs, work):
        """work(handle, chunk) for every chunk: chunk k on handle k mod
        --streams, each handle's chunks in order on a thread of its own.
        Returns [(chunk, result)] in the order of chs."""
        groups = [list(range(k, len(chs), nstreams)) for k in range(nstreams)]
        out = [None] * len(chs)
        if nstreams == 1:
            for k, ch in enumerate(chs):
                out[k] = (ch, work(S, ch))
            return out
        errors = []

        def body(h, ks):
            try:
                for k in ks:
                    out[k] = (chs[k], work(handles[h], chs[k]))
            except BaseException as e:   # re-raised on the main thread
                errors.append(e)
        threads = [threading.Thread(target=body, args=(h, groups[h])) for h in range(nstreams)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        return out
    ts_all = [np.linspace(0.0, inst.T, inst.nodes) for inst in insts]
    cons_all = np.array([inst.constant_values() for inst in insts])
    hom_all = [np.array([inst.constant_values(s=0.0) for inst in insts])]

    fb_scales = [float(v) for v in args.fallback_scales.split(",") if v.strip()]
    fb_hom = [np.array([inst.constant_values(s=v) for inst in insts]) for v in fb_scales]
    def stall_rule(txt):
        if txt == "0":
            return None
        v = [float(x) for x in txt.split(",")]
        return int(v[0]), int(v[1]), v[2]
    fb_stall = stall_rule(args.fallback_stall)
    warm_cont = stall_rule(args.warm_cont)

    def plain_batch(idx, hom=None, stall=None):
        """All-ones guess + continuation for the instances idx, chunked: by
        default the plain path (homotopy S: 0 -> final, no stall test)."""
        hom = hom_all if hom is None else hom
        conv = np.zeros(len(idx), dtype=bool)
        its = np.zeros((len(hom) + 1, len(idx)), dtype=int)

        def work(S_, ch):
            sel = [idx[k] for k in ch]
            Nt = sum(nodes[i] for i in sel)
            S_.set_batch_mesh([ts_all[i] for i in sel])
            S_.form_initial_input(np.ones((Nt, ny)), np.ones((Nt, nz)), np.ones((len(sel), np_)))
            return solver.newton_stages(S_, cons_all[sel], [h[sel] for h in hom], 0.1, tol, max_iter, stall=stall)
        for ch, r in run_chunks(balanced_chunks([nodes[i] for i in idx], S.n_max, S.batch), work):
            conv[ch] = r["converged"]
            its[:, ch] = r["iterations"]
        return conv, its

    ws = None
    model_info = None
    if args.form == "learned":
        model = args.model or learned.FAMILY_MODEL
        if not os.path.exists(model):
            print("no trained model at %s: training one now (see train_warmstart.py)" % model, file=sys.stderr)
            learned.fit_family(path=model, log=lambda s: print(s, file=sys.stderr, flush=True))
        ws = learned.WarmStart.load(model)
        if os.path.exists(model + ".json"):
            model_info = json.load(open(model + ".json"))
        feats_all = learned.features([inst.x0 for inst in insts], [inst.T for inst in insts])

    def learned_pass():
        """Network guess + Newton at the final constants, batched; the
        instances whose Newton solve does not converge go to plain_batch."""
        conv = np.zeros(len(insts), dtype=bool)
        its = np.zeros(len(insts), dtype=int)
        newton_ok = np.zeros(len(insts), dtype=bool)
        t_warm = time.perf_counter()

        def work(S_, ch):
            return learned.warm_solve(ws, S_, [ts_all[i] for i in ch], feats_all[ch], cons_all[ch], hom_all[0][ch],
                                      tol=tol, max_iter=max_iter, max_iter_warm=args.warm_cap,
                                      patience=args.warm_patience, min_ratio=args.warm_ratio, cont=warm_cont)
        for ch, r in run_chunks(chunks, work):
            conv[ch] = r["converged"]
            its[ch] = r["iterations"].sum(axis=0) + r["lq_iterations"]
            newton_ok[ch] = r["norm"] < tol
            phase_s["warm_converged"] += int(r["warm_converged"].sum())
            phase_s["learned_converged"] += int(r["converged"].sum())
        back = [i for i in range(len(insts)) if not newton_ok[i]]
        t_back = time.perf_counter()
        phase_s["warm"] += t_back - t_warm
        phase_s["fallback_instances"] += len(back)
        if back and args.fallback == "plain":
            c2, _ = plain_batch(back, fb_hom, fb_stall)
            conv[back] = c2
            phase_s["fallback_converged"] += int(c2.sum())
        phase_s["fallback"] += time.perf_counter() - t_back
        return conv, its, len(back)

    def plain_pass():
        conv, its = plain_batch(list(range(len(insts))))
        return conv, its, 0

    one_pass = learned_pass if args.form == "learned" else plain_pass
    phase_s = {"warm": 0.0, "fallback": 0.0, "fallback_converged": 0, "warm_converged": 0, "learned_converged": 0,
               "fallback_instances": 0}
    for _ in range(args.warmup):
        one_pass()
    tune_rec = None
    if args.autotune:
        # on each handle's last chunk (mesh and state as the warmup left them)
        from colloc_amd import autotune
        for S_ in handles:
            tune_rec = autotune.tune(S_, log=lambda s: print("autotune: " + s, file=sys.stderr, flush=True))
    spans_used = S_spans(S, tune_rec, args.babd_spans)
    phase_s = {"warm": 0.0, "fallback": 0.0, "fallback_converged": 0, "warm_converged": 0, "learned_converged": 0,
               "fallback_instances": 0}
    each(lambda S_: S_.profile(True))
    st0 = tuple(int(x) for x in np.sum([S_.condense_stats() for S_ in handles], axis=0))
    t0 = time.perf_counter()
    results = [one_pass() for _ in range(steps)]
    elapsed = time.perf_counter() - t0
    st1 = tuple(int(x) for x in np.sum([S_.condense_stats() for S_ in handles], axis=0))
    prof = {}
    for k in ("residual", "jacobian", "condense", "babd", "recover"):
        v = [S_.profile_get(k) for S_ in handles]
        prof[k] = (sum(x[0] for x in v), sum(x[1] for x in v))
    each(lambda S_: S_.profile(False))
    phases = {k: v / steps for k, v in phase_s.items()}
    converged = int(sum(r[0].sum() for r in results))
    value = converged / elapsed
    conv, its, n_back = results[-1]
    npa = len(jit.unit_info(bvp.ocp_text, m)["active_params"])
    kname = condense_kernel_name(args.precision, args.condense_kernel)
    roof = condense_roofline(S, st0, st1, prof["condense"][0], prof["condense"][1], ny, nz, np_, npa, m,
                             measured_traffic("batch_%s" % args.form, kname), kname)
    roof["arithmetic"] = "fp%d" % args.precision
    if nstreams > 1:
        roof["concurrent_streams"] = nstreams
        roof["note"] = ("launch durations are HIP-event times on each handle's stream while %d handles run "
                        "concurrently: a launch shares the GPU with the other streams' kernels" % nstreams)
    extra = {}
    if args.form == "learned" and args.precision == 32 and args.fp64_pass:
        # the same learned + batched solve with the float64 direction
        each(lambda S_: S_.set_direction_precision(64))
        each(lambda S_: S_.set_babd_precision(64))
        each(lambda S_: S_.set_refinement(0))
        t0 = time.perf_counter()
        c64, it64, b64 = learned_pass()
        lt = time.perf_counter() - t0
        extra.update({"learned_fp64_solves_per_s": c64.sum() / lt, "learned_fp64_converged": int(c64.sum()),
                      "learned_fp64_unconverged_after_newton": b64})
    if args.form == "learned" and args.plain_pass:
        # the plain batched form (float64 direction) on the same instances, for the split of the gain
        each(lambda S_: S_.set_direction_precision(64))
        each(lambda S_: S_.set_babd_precision(64))
        each(lambda S_: S_.set_refinement(0))
        t0 = time.perf_counter()
        pconv, pits, _ = plain_pass()
        pt = time.perf_counter() - t0
        extra.update({"plain_batched_solves_per_s": pconv.sum() / pt, "plain_batched_converged": int(pconv.sum()),
                      "plain_newton_iterations_mean": pits.sum(axis=0).mean()})
    each(lambda S_: S_.close())

    # the plain single-instance solve looped over a sample of the same instances
    S1 = solver.CollocationSolver(bvp, m, n_max=max(nodes), batch=1)
    sample = list(range(min(args.loop_sample, len(insts))))
    t0 = time.perf_counter()
    loop_ok = 0
    for i in sample:
        r = solver.solve_batch(S1, [ts_all[i]], [np.ones((nodes[i], ny))], [np.ones((nodes[i], nz))], [np.ones(np_)],
                               cons_all[i:i + 1], [h[i:i + 1] for h in hom_all], tol=tol, max_iter=max_iter)
        loop_ok += int(r["converged"][0])
    loop_s = time.perf_counter() - t0
    S1.close()
    loop_rate = loop_ok / loop_s
    mean_nint = float(np.mean([n - 1 for n in nodes]))
    it_mean = np.asarray(its).reshape(-1, len(insts)).sum(axis=0).mean()
    print("batch (%s): %d chunks, converged %d/%d per pass, Newton iterations mean %.2f, unconverged after the "
          "warm-started Newton %d" % (args.form, len(chunks), conv.sum(), len(insts), it_mean, n_back),
          file=sys.stderr)
    print("kernel totals over %d passes (ms, launches): %s" % (steps, prof), file=sys.stderr)
    if args.form == "learned":
        print("per pass: warm-started solves %.2f s (%.1f converged in the capped stage), %s fallback %.2f s "
              "(%.1f converged)" % (phases["warm"], phases["warm_converged"], args.fallback, phases["fallback"],
                                    phases["fallback_converged"]), file=sys.stderr)
    print("looped plain solve: %d instances in %.2f s (%d converged)" % (len(sample), loop_s, loop_ok),
          file=sys.stderr)
    form_txt = ("LQ solve (1 Newton iteration) + network correction (MLP 21->4x1024->3250, bf16, trained on "
                "solver solutions) + at most %d Newton iterations at the final constants (stopped from iteration "
                "%d on when the norm falls by less than %gx per iteration)%s; %s"
                % (args.warm_cap, args.warm_patience, args.warm_ratio,
                   "" if warm_cont is None else
                   (", the instances it leaves continuing damped Newton for the rest of the %d-iteration budget "
                    "(stopped from iteration %d on when the norm fell by less than %gx over the last %d "
                    "iterations)" % (max_iter, warm_cont[0], warm_cont[2], warm_cont[1])),
                   "no fallback" if args.fallback == "none" else
                   ("fallback: plain guess + continuation S = %s -> final%s"
                    % (args.fallback_scales,
                       "" if fb_stall is None else (", each stage stopping an instance from iteration %d on when "
                                                     "its norm fell by less than %gx over the last %d iterations"
                                                     % (fb_stall[0], fb_stall[2], fb_stall[1])))))
                if args.form == "learned"
                else "plain all-ones guess + homotopy S: 0->0.5")
    form_txt += ("; Newton direction condensed in float32 (residuals, BABD, convergence test in float64)"
                 if args.precision == 32 else "; float64 throughout")
    form_txt += ("; BABD pairs by substitution while a pair spans <= 1/%d of its instance, by Householder QR above"
                 % spans_used if spans_used > 0 else "; BABD pairs by Householder QR on every level")
    if args.babd_precision == 32:
        form_txt += "; cyclic reduction with float32 rows and records"
    if args.refine:
        form_txt += "; %d iterative refinement step(s) per Newton direction" % args.refine
    if args.graph:
        form_txt += "; Newton loops device resident, each iteration a HIP-graph replay"
    line = {
        "metric": "converged BVP solves/s @ residual <=1e-8, %s vs plain FP64, 1 GPU"
                  % (("learned+batched" if args.form == "learned" else "batched")
                     + ("+FP32-direction" if args.precision == 32 else "")),
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 64 else "f64 (residuals, BABD, convergence) + f32 (direction condensation)",
        "data": "synthetic (seeded OCP family, generated in-process)",
        "config": {"workload": "batched family: %d instances of the tanh OCP n_x=20 n_u=10 (ny=40 nz=10 np=40), "
                               "shared A,B,C (seed 1), x0~U[-1,1]^20, T~U[0.5,2], Nint~U{256..2048}, m=%d; %s; "
                               "%d chunk(s) of <= %d nodes%s" % (len(insts), m, form_txt, len(chunks), chunk_nodes,
                                                                 "" if nstreams == 1 else
                                                                 ", solved %d at a time (one handle, host thread "
                                                                 "and HIP stream each)" % nstreams),
                   "form": args.form, "direction_precision": args.precision, "instances": len(insts), "mean_nint": mean_nint, "m": m, "tolerance": tol,
                   "converged_per_pass": int(conv.sum()), "chunks": len(chunks), "streams": nstreams,
                   "newton_iterations_mean": float(it_mean),
                   "warm_stage_converged_per_pass": phases["warm_converged"] if args.form == "learned" else None,
                   "unconverged_after_newton_per_pass": n_back if args.form == "learned" else None,
                   "fallback": args.fallback if args.form == "learned" else None,
                   "seconds_per_pass": {"warm_start_solves": phases["warm"], "fallback": phases["fallback"]}
                   if args.form == "learned" else None,
                   "fallback_converged_per_pass": phases["fallback_converged"] if args.form == "learned" else None,
                   "fallback_rate": phases["fallback_instances"] / len(insts) if args.form == "learned" else None,
                   "no_fallback_solves_per_s": (phases["learned_converged"] / phases["warm"])
                   if args.form == "learned" and phases["warm"] > 0 else None,
                   "no_fallback_converged_per_pass": phases["learned_converged"] if args.form == "learned" else None,
                   "looped_plain_solves_per_s": loop_rate, "looped_sample": len(sample),
                   "looped_sample_converged": loop_ok,
                   "vs_looped_plain": value / loop_rate if loop_rate > 0 else None,
                   "model": model_info,
                   "autotune": ({"choices": tune_rec["choices"], "timings_ms": tune_rec["timings_ms"],
                                 "cached": tune_rec["cached"]} if tune_rec else None),
                   "pcie_included": "per chunk: mesh (and for the plain form the all-ones guess) copied in, "
                                    "network input copied in (inside the timed step)", **extra},
        "roofline": roof,
    }
    if args.cpu_sample > 0:
        line["cpu_baseline"] = cpu_baseline(insts[0], m, args.cpu_sample, mean_nint)
        line["cpu_baseline"]["protocol"] = cpu_protocol()
    return line


def single_workload(args):
    from colloc_amd import jit, problems, solver
    steps = args.steps if args.steps is not None else 5
    inst, m = problems.config(1, nint=args.nint)
    bvp = solver.BvpDae(inst.ocp_text())
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    S.set_newton_graph(bool(args.graph))
    S.set_mesh(bvp.T0)
    S.form_initial_input(bvp.Y0, bvp.Z0, bvp.P0)
    schedule = inst.homotopy_schedule()
    final_c = bvp.constants
    tol = bvp.tolerance

    def one_solve():
        S.restart()
        its = []
        for cv in schedule:
            S.set_constants(cv)
            it, nrm, _ = S.newton(0.1, tol, bvp.maximum_newton_iterations)
            its.append(it)
        S.set_constants(final_c)
        it, nrm, _ = S.newton(0.1, tol, bvp.maximum_newton_iterations)
        its.append(it)
        _, mx = S.segment_residual(0.1, tol)
        ok = nrm < tol and mx <= 1.0 and it < bvp.maximum_newton_iterations - 1
        return ok, its, nrm, mx

    for _ in range(args.warmup):
        one_solve()
    S.profile(True)
    st0 = S.condense_stats()
    t0 = time.perf_counter()
    results = [one_solve() for _ in range(steps)]
    elapsed = time.perf_counter() - t0
    st1 = S.condense_stats()
    prof = {k: S.profile_get(k) for k in ("residual", "jacobian", "condense", "babd", "recover")}
    S.profile(False)
    converged = sum(1 for r in results if r[0])
    value = converged / elapsed
    nint = bvp.N - 1
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    npa = len(jit.unit_info(bvp.ocp_text, m)["active_params"])
    roof = condense_roofline(S, st0, st1, prof["condense"][0], prof["condense"][1], ny, nz, np_, npa, m,
                             measured_traffic("tanh20_nint%d" % nint, "colloc_condense_ns"), "colloc_condense_ns")
    print("per-solve Newton iterations %s, final residual %.2e, max segment residual %.3g"
          % (results[-1][1], results[-1][2], results[-1][3]), file=sys.stderr)
    print("kernel totals over %d solves (ms, launches): %s" % (steps, prof), file=sys.stderr)
    line = {
        "metric": "converged BVP solves/s @ residual <=1e-8 (plain FP64 path, 1 GPU)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded OCP family, generated in-process)",
        "config": {"workload": "plain FP64 single-instance solve, tanh OCP n_x=20 n_u=10 (ny=40 nz=10 np=40), "
                               "Nint=%d, m=%d, all-ones guess + homotopy S: 0->0.5%s"
                               % (nint, m, "; Newton loop device resident (HIP-graph replay per iteration)"
                                  if args.graph else ""),
                   "nint": nint, "m": m, "ny": ny, "nz": nz, "np": np_, "np_in_stage_functions": npa,
                   "tolerance": tol,
                   "converged_steps": converged, "newton_iterations": results[-1][1]},
        "roofline": roof,
    }
    if args.cpu_sample > 0:
        line["cpu_baseline"] = cpu_baseline(inst, m, args.cpu_sample, nint)
    S.close()
    return line


def precision_workload(args):
    """The precision form on the dense instance: tanh OCP n_x=64 n_u=32 (ny=128,
    nz=32, np=128), Nint=4096, m=4, plain all-ones guess + homotopy S: 0 ->
    0.5, solved with the float64 Newton direction and with the FP32 one
    (float32 Jacobians and condensation; residuals, BABD, convergence test in
    float64; --refine steps of iterative refinement), timed alternately."""
    from colloc_amd import jit, problems, solver
    steps = args.steps if args.steps is not None else 3
    inst, m = problems.config(2, nint=args.nint)
    bvp = solver.BvpDae(inst.ocp_text())
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    S.set_babd_substitution(args.babd_spans)
    S.set_mesh(bvp.T0)
    S.form_initial_input(bvp.Y0, bvp.Z0, bvp.P0)
    schedule = inst.homotopy_schedule()
    final_c = bvp.constants
    tol = bvp.tolerance

    def one_solve(bits):
        S.set_direction_precision(bits)
        S.set_refinement(args.refine if bits == 32 else 0)
        S.set_babd_precision(args.babd_precision if bits == 32 else 64)
        S.restart()
        its = []
        for cv in schedule:
            S.set_constants(cv)
            it, nrm, _ = S.newton(0.1, tol, bvp.maximum_newton_iterations)
            its.append(it)
        S.set_constants(final_c)
        it, nrm, _ = S.newton(0.1, tol, bvp.maximum_newton_iterations)
        its.append(it)
        _, mx = S.segment_residual(0.1, tol)
        ok = nrm < tol and mx <= 1.0 and it < bvp.maximum_newton_iterations - 1
        return ok, its, nrm, mx

    forms = {}
    sols = {}
    tuned = {}
    for bits in (64, 32):
        for _ in range(args.warmup):
            one_solve(bits)
        if args.autotune:
            # the pairing rule on the solved state, outside the timed region
            from colloc_amd import autotune
            S.set_direction_precision(bits)
            rec = autotune.tune(S, log=lambda s: print("autotune fp%d: %s" % (bits, s), file=sys.stderr, flush=True))
            tuned[bits] = {"choices": rec["choices"], "timings_ms": rec["timings_ms"], "cached": rec["cached"]}
        S.profile(True)
        st0 = S.condense_stats()
        t0 = time.perf_counter()
        results = [one_solve(bits) for _ in range(steps)]
        elapsed = time.perf_counter() - t0
        st1 = S.condense_stats()
        prof = {k: S.profile_get(k) for k in ("residual", "jacobian", "condense", "babd", "recover")}
        S.profile(False)
        y, _, _, p = S.get_state()
        sols[bits] = (y, p)
        forms[bits] = dict(results=results, elapsed=elapsed, st=(st0, st1), prof=prof)
        print("direction fp%d: %.3f s per solve, Newton iterations %s, kernel totals %s"
              % (bits, elapsed / steps, results[-1][1], prof), file=sys.stderr)
    nint = bvp.N - 1
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    npa = len(jit.unit_info(bvp.ocp_text, m)["active_params"])
    f32 = forms[32]
    conv32 = sum(1 for r in f32["results"] if r[0])
    value = conv32 / f32["elapsed"]
    roof = condense_roofline(S, f32["st"][0], f32["st"][1], f32["prof"]["condense"][0], f32["prof"]["condense"][1],
                             ny, nz, np_, npa, m, None, "colloc_l_series_term32")
    roof["kernel"] = "large-block condensation (colloc_l_stage_elim32 + colloc_l_series_term32 on matrix cores)"
    roof["arithmetic"] = "fp32"
    scale = max(np.max(np.abs(sols[64][0])), 1e-300)
    line = {
        "metric": "converged BVP solves/s @ residual <=1e-8, FP32-direction precision form, 1 GPU",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": f32["elapsed"] / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (residuals, BABD, convergence) + f32 (Jacobians, condensation)",
        "data": "synthetic (seeded OCP family, generated in-process)",
        "config": {"workload": "precision form: tanh OCP n_x=64 n_u=32 (ny=128 nz=32 np=128), Nint=%d, m=%d, "
                               "all-ones guess + homotopy S: 0->0.5; FP32 Newton direction (float32 Jacobians and "
                               "condensation, %s cyclic reduction, %d refinement steps) against the float64 one; "
                               "callbacks in polynomial table form, large-block kernels"
                               % (nint, m, "float32" if args.babd_precision == 32 else "float64", args.refine),
                   "babd_precision": args.babd_precision, "refine": args.refine,
                   "nint": nint, "m": m, "ny": ny, "nz": nz, "np": np_, "tolerance": tol,
                   "fp32_converged_steps": conv32, "fp32_newton_iterations": f32["results"][-1][1],
                   "fp64_solves_per_s": sum(1 for r in forms[64]["results"] if r[0]) / forms[64]["elapsed"],
                   "fp64_newton_iterations": forms[64]["results"][-1][1],
                   "fp32_vs_fp64_solution_rel_diff": float(np.max(np.abs(sols[32][0] - sols[64][0])) / scale),
                   "kernel_ms_fp32": f32["prof"], "kernel_ms_fp64": forms[64]["prof"],
                   "autotune": {"fp%d" % b: v for b, v in tuned.items()} or None},
        "roofline": roof,
    }
    S.close()
    return line


if __name__ == "__main__":
    main()
