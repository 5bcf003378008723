"""Throughput benchmark: converged BVP solves per second on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): the
n_x=20, n_u=10 tanh OCP (seed 1, T=2), Nint=4096 uniform intervals, m=5
Lobatto IIIA stages, FP64, from the plain all-ones initial guess with the
two-step homotopy in the tanh scale (0 -> 0.5).  One step = one complete
solve: restart from the device-resident initial guess, Newton at each
homotopy stage to ||f_q||_inf < 1e-8, segment-residual acceptance.  A step
counts only if it converged (residual <= 1e-8 and max segment residual <= 1).

Printed: one JSON line with the metric, the roofline of the dominant kernel
(condensation, timed live with HIP events on the solver's stream) and the CPU
baseline (the NumPy oracle on a bounded sample of the same workload).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "collocation-solver-cuda_amd")
FP64_PEAK_TFLOPS = 78.6     # MI355X, FP64 vector = FP64 matrix peak
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


def measured_traffic(nint):
    """HBM bytes per colloc_condense launch at this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy kernel);
    None when no measurement exists for this mesh size."""
    path = os.path.join(ROOT, "profiles", "traffic_tanh20_nint%d_m5.md" % nint)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch colloc_condense:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(inst, m, nint_sample):
    """Oracle (NumPy/LAPACK) plain solve of the same family at a reduced mesh;
    work per solve is linear in Nint at equal Newton counts, so the rate is
    scaled by nint_sample / Nint."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import families
    from oracle import solver as osolver
    from colloc_amd import problems
    small, _ = problems.config(1, nint=nint_sample)
    prob = families.TanhFamily(small.A, small.B, small.x0, C=small.C, s=small.s, q=small.q, c_u=small.c_u)
    hom = [families.TanhFamily(small.A, small.B, small.x0, C=small.C, s=0.0, q=small.q, c_u=small.c_u)]
    N = small.nodes
    t = np.linspace(0.0, small.T, N)
    ny, nz, np_ = prob.size_y, prob.size_z, prob.size_p
    t0 = time.perf_counter()
    res = osolver.collocation_solver(prob, m, t, np.ones((N, ny)), np.ones((N, nz)), np.ones(np_),
                                     tol=small.tolerance, homotopy=hom)
    dt = time.perf_counter() - t0
    ok = res["success"] == 1
    scale = (inst.nodes - 1) / float(nint_sample)
    return {"value": (1.0 / (dt * scale)) if ok else 0.0, "unit": "solves/s",
            "cores": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": "NumPy oracle, full plain solve (homotopy 0->0.5) of the same family at Nint=%d "
                      "(%.1f s), rate scaled by %d/%d to Nint=%d; LAPACK threads = all cores"
                      % (nint_sample, dt, nint_sample, inst.nodes - 1, inst.nodes - 1),
            "seconds": dt, "converged": ok}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nint", type=int, default=None, help="override Nint (default: the configuration's 4096)")
    ap.add_argument("--cpu-sample", type=int, default=128, help="Nint of the CPU baseline sample (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the solve shards by "
              "interval ranges or by instance; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    load_package()
    from colloc_amd import build as b
    if not os.path.exists(b.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    from colloc_amd import problems, solver

    inst, m = problems.config(1, nint=args.nint)
    bvp = solver.BvpDae(inst.ocp_text())
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
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
    fb0, terms0 = S.condense_stats()
    t0 = time.perf_counter()
    results = [one_solve() for _ in range(args.steps)]
    elapsed = time.perf_counter() - t0
    fb1, terms1 = S.condense_stats()
    cond_ms, cond_n = S.profile_get("condense")
    prof = {k: S.profile_get(k) for k in ("residual", "jacobian", "condense", "babd", "recover")}
    S.profile(False)
    converged = sum(1 for r in results if r[0])
    value = converged / elapsed
    nint = bvp.N - 1
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    from colloc_amd import jit
    npa = len(jit.unit_info(bvp.ocp_text, m)["active_params"])
    # executed flops of all timed condensation launches: stage eliminations of
    # every interval, the series terms actually run, Gauss-Jordan for the
    # intervals the series left
    fl_total = (cond_n * nint * stage_elim_flops(ny, nz, npa, m) + (terms1 - terms0) * series_term_flops(ny, npa, m)
                + (fb1 - fb0) * gauss_jordan_flops(ny, npa, m))
    fl = fl_total / max(cond_n, 1)
    terms_per_interval = (terms1 - terms0) / max(cond_n * nint, 1)
    avg_s = cond_ms / max(cond_n, 1) / 1e3
    achieved = fl / avg_s / 1e12 if avg_s > 0 else 0.0
    print("per-solve Newton iterations %s, final residual %.2e, max segment residual %.3g"
          % (results[-1][1], results[-1][2], results[-1][3]), file=sys.stderr)
    print("kernel totals over %d solves (ms, launches): %s" % (args.steps, prof), file=sys.stderr)
    line = {
        "metric": "converged BVP solves/s @ residual <=1e-8 (plain FP64 path, 1 GPU)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded OCP family, generated in-process)",
        "config": {"workload": "plain FP64 single-instance solve, tanh OCP n_x=20 n_u=10 (ny=40 nz=10 np=40), "
                               "Nint=%d, m=%d, all-ones guess + homotopy S: 0->0.5" % (nint, m),
                   "nint": nint, "m": m, "ny": ny, "nz": nz, "np": np_, "np_in_stage_functions": npa,
                   "tolerance": tol,
                   "converged_steps": converged, "newton_iterations": results[-1][1]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(nint),
                     "kernel": "colloc_condense_ns (+ colloc_condense for intervals the series leaves)",
                     "flops_per_launch": fl, "avg_launch_ms": avg_s * 1e3, "launches": cond_n,
                     "series_terms_per_interval": terms_per_interval, "gauss_jordan_intervals": fb1 - fb0,
                     "dense_equivalent_tflops": dense_condense_flops(ny, nz, np_, m) * nint / avg_s / 1e12
                     if avg_s > 0 else 0.0},
    }
    if args.cpu_sample > 0:
        line["cpu_baseline"] = cpu_baseline(inst, m, args.cpu_sample)
    S.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
