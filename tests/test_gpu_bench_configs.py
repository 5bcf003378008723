"""The benchmarked workloads at the settings bench.py runs them with, checked
against the oracle.

* The BABD solve at Nint = 2048 with substitution pairing while a pair spans
  at most 1/16 of the intervals (the bench's setting): 7 substitution levels,
  transition products over up to 128 intervals, then Householder levels.
  The Newton direction must equal the oracle's M = 8 partitioned QR solve of
  the same BABD system (1e-10 relative, float64 condensation; 1e-4 with the
  float32 condensation of the FP32 direction).
* A sample of the 1024-instance family (seed 1), the two longest meshes
  (Nint close to 2048) among them, solved batched with the FP32 direction and
  the bench's BABD pairing: every converged instance must satisfy the
  oracle's float64 residual (||f_q||_inf < 1e-8) and acceptance test (segment
  residual <= 1), and match the float64-direction solve of the same batch
  (1e-7 relative: two solutions of the same discrete problem to 1e-8).
* The learned warm start with the bench's stopping rules and the plain
  fallback on 64 held-out members (the trained family model; skipped when
  the model file is absent): fallback rate and Newton iterations are
  reported, every converged solution passes the oracle's residual test, the
  learned form converges more instances than the plain path, and those both
  solve take no more iterations (median) from the network's guess.
"""
import os

import numpy as np
import pytest

from colloc_amd import problems, solver
from oracle import babd as obabd
from oracle import families
from oracle import solver as osolver
from oracle import tables as otables

pytestmark = pytest.mark.gpu

SPANS = 16          # bench.py --babd-spans default
M_STAGES = 5


def say(msg):
    print(msg, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def family_of(inst):
    return families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)


def substitution_levels(nint, spans):
    lev = 0
    while spans << (lev + 1) <= nint:
        lev += 1
    return lev


@pytest.mark.parametrize("bits", [64, 32])
def test_seven_level_substitution_babd(bits):
    inst, m = problems.config(1, nint=2048)
    assert substitution_levels(2048, SPANS) == 7
    bvp = solver.BvpDae(inst.ocp_text())
    prob = family_of(inst)
    N = bvp.N
    S = solver.CollocationSolver(bvp, m, n_max=N)
    S.set_babd_substitution(SPANS)
    S.set_mesh(bvp.T0)
    S.form_initial_input(bvp.Y0, bvp.Z0, bvp.P0)
    # a state on the bench's path: the LQ solution (the first continuation
    # stage), then the direction at the final constants
    say("Nint=2048: LQ solve")
    S.set_constants(inst.homotopy_schedule()[0])
    it, nrm, _ = S.newton(0.1, bvp.tolerance, bvp.maximum_newton_iterations)
    assert nrm < bvp.tolerance
    S.set_constants(bvp.constants)
    S.set_direction_precision(bits)
    y, yd, zt, p = S.get_state()
    say("Nint=2048: direction with %d-bit condensation, %d substitution levels" % (bits, 7))
    S.compute_f_q(0.1, outputs=False)
    S.construct_jacobian(0.1, outputs=False)
    A, C, H, b = S.reduce_jacobian()
    dy, dp = S.solve_babd()
    assert S.status() == 0 and S.condense_fallbacks() == 0
    S.close()
    t = bvp.T0
    rk = otables.lobatto(m)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    dy_o, dp_o = obabd.solve_babd(prob.size_y, prob.size_p, N, A_o, C_o, H_o, b_o, B0, Bn, Vn, rbc_o, M=8)
    errs = (rel_err(A, A_o), rel_err(b, b_o), rel_err(dy, dy_o), rel_err(dp, dp_o))
    say("A, b, dy, dp errors vs the oracle: %.1e %.1e %.1e %.1e" % errs)
    # the same BABD system solved by substitution levels + Householder on the
    # GPU, by the partitioned QR on the CPU
    dy_s, dp_s = obabd.solve_babd(prob.size_y, prob.size_p, N, A, C_o, H, b, B0, Bn, Vn, rbc_o, M=8)
    say("dy, dp vs the oracle's solve of the GPU's rows: %.1e %.1e" % (rel_err(dy, dy_s), rel_err(dp, dp_s)))
    assert rel_err(dy, dy_s) < 1e-10 and rel_err(dp, dp_s) < 1e-10
    if bits == 64:
        assert max(errs) < 1e-10
    else:
        assert errs[0] < 1e-5 and errs[1] < 1e-5 and errs[2] < 1e-4 and errs[3] < 1e-4


def sample_members():
    fam = problems.batch_family(1024)
    order = np.argsort([-i.nodes for i in fam])
    idx = sorted(set(order[:2].tolist()) | {0, 1, 2, 3})
    return fam, idx


def oracle_accepts(inst, sol, m, tol):
    """The oracle's float64 residual and acceptance residual of a GPU solution."""
    prob = family_of(inst)
    N = sol["t"].size
    rk = otables.lobatto(m)
    f_q = osolver.compute_f_q(prob, m, N, rk, sol["t"], sol["y"], sol["y_dot"], sol["z_tilde"], sol["p"], 0.1)[0]
    _, mx = osolver.compute_segment_residual(prob, m, N, sol["t"], sol["y"], sol["y_dot"], sol["z_tilde"], sol["p"],
                                             0.1, tol)
    return np.max(np.abs(f_q)), mx


def full_solutions(S):
    """Per-instance t, y, y_dot, z_tilde, p, z of the handle's batch."""
    y, yd, zt, p = S.get_state()
    z = S.recover_solution()
    p = np.atleast_2d(p)
    no, io = S.node_offsets(), S.interval_offsets() * S.m
    return [{"t": S.t_span[no[b]:no[b + 1]], "y": y[no[b]:no[b + 1]], "z": z[no[b]:no[b + 1]],
             "y_dot": yd[io[b]:io[b + 1]], "z_tilde": zt[io[b]:io[b + 1]], "p": p[b]} for b in range(S.B)]


def test_family_sample_fp32_direction_against_oracle():
    fam, idx = sample_members()
    insts = [fam[i] for i in idx]
    say("members %s, Nint %s" % (idx, [i.nodes - 1 for i in insts]))
    assert max(i.nodes - 1 for i in insts) > 1900
    bvp = solver.BvpDae(insts[0].ocp_text())
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    cons = np.array([i.constant_values() for i in insts])
    hom = [np.array([i.constant_values(s=0.0) for i in insts])]
    S = solver.CollocationSolver(bvp, M_STAGES, n_max=sum(i.nodes for i in insts), batch=len(insts))
    S.set_babd_substitution(SPANS)
    out = {}
    for bits in (32, 64):
        S.set_direction_precision(bits)
        say("plain batched solve, %d-bit direction" % bits)
        r = solver.solve_batch(S, ts, [np.ones((len(t), bvp.size_y)) for t in ts],
                               [np.ones((len(t), bvp.size_z)) for t in ts], [np.ones(bvp.size_p) for _ in ts], cons,
                               hom, tol=bvp.tolerance, max_iter=bvp.maximum_newton_iterations)
        out[bits] = (r, full_solutions(S))
        say("  converged %s, iterations %s" % (r["converged"].astype(int).tolist(), r["iterations"].tolist()))
    S.close()
    r32, s32 = out[32]
    r64, s64 = out[64]
    assert r32["converged"].sum() >= 1
    for k, inst in enumerate(insts):
        if not r32["converged"][k]:
            continue
        res, mx = oracle_accepts(inst, s32[k], M_STAGES, bvp.tolerance)
        say("member %d (Nint %d): oracle ||f_q|| %.2e, segment residual %.3f" % (idx[k], inst.nodes - 1, res, mx))
        assert res < bvp.tolerance and mx <= 1.0
        if r64["converged"][k]:
            for key in ("y", "z", "p"):
                assert rel_err(s32[k][key], s64[k][key]) < 1e-7, (idx[k], key)
    # the float32 direction converges what the float64 one converges
    assert np.all(r32["converged"] >= r64["converged"])


def test_learned_warm_start_with_plain_fallback_held_out():
    from colloc_amd import learned
    if not os.path.exists(learned.FAMILY_MODEL):
        pytest.skip("no trained family model at %s (train_warmstart.py makes it)" % learned.FAMILY_MODEL)
    ws = learned.WarmStart.load(learned.FAMILY_MODEL)
    insts = problems.batch_family(64)           # members 0..63: held out (training uses 100000+)
    bvp = solver.BvpDae(insts[0].ocp_text())
    tol, max_iter = bvp.tolerance, bvp.maximum_newton_iterations
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    cons = np.array([i.constant_values() for i in insts])
    lq = np.array([i.constant_values(s=0.0) for i in insts])
    feats = learned.features([i.x0 for i in insts], [i.T for i in insts])
    nodes = [i.nodes for i in insts]
    chunks = solver.plan_chunks(nodes, 40000, 64)
    S = solver.CollocationSolver(bvp, M_STAGES, n_max=max(sum(nodes[i] for i in c) for c in chunks),
                                 batch=max(len(c) for c in chunks))
    S.set_babd_substitution(SPANS)
    conv = np.zeros(len(insts), dtype=bool)
    its = np.zeros(len(insts), dtype=int)
    fallback = []
    sols = [None] * len(insts)
    for ch in chunks:
        S.set_direction_precision(32)
        say("learned chunk of %d" % len(ch))
        r = learned.warm_solve(ws, S, [ts[i] for i in ch], feats[ch], cons[ch], lq[ch], tol=tol, max_iter=max_iter,
                               max_iter_warm=12, patience=8, min_ratio=2.0, cont=(8, 5, 2.0))
        conv[ch] = r["converged"]
        its[ch] = r["iterations"].sum(axis=0) + r["lq_iterations"]
        for k, s in zip(ch, full_solutions(S)):
            sols[k] = s
        fallback += [ch[k] for k in range(len(ch)) if not r["norm"][k] < tol]
    warm_conv = conv.copy()
    if fallback:
        say("plain fallback for %d instances" % len(fallback))
        for ch in solver.plan_chunks([nodes[i] for i in fallback], S.n_max, S.batch):
            sel = [fallback[k] for k in ch]
            r = solver.solve_batch(S, [ts[i] for i in sel], [np.ones((nodes[i], bvp.size_y)) for i in sel],
                                   [np.ones((nodes[i], bvp.size_z)) for i in sel], [np.ones(bvp.size_p) for _ in sel],
                                   cons[sel], [lq[sel]], tol=tol, max_iter=max_iter)
            conv[sel] = r["converged"]
            its[sel] += r["iterations"].sum(axis=0)
            for k, s in zip(sel, full_solutions(S)):
                sols[k] = s
    # the plain path on the same instances
    say("plain path")
    S.set_direction_precision(64)
    pconv = np.zeros(len(insts), dtype=bool)
    pits = np.zeros(len(insts), dtype=int)
    for ch in chunks:
        r = solver.solve_batch(S, [ts[i] for i in ch], [np.ones((nodes[i], bvp.size_y)) for i in ch],
                               [np.ones((nodes[i], bvp.size_z)) for i in ch], [np.ones(bvp.size_p) for _ in ch],
                               cons[ch], [lq[ch]], tol=tol, max_iter=max_iter)
        pconv[ch] = r["converged"]
        pits[ch] = r["iterations"].sum(axis=0)
    S.close()
    both = conv & pconv
    say("held-out 64: learned converged %d (%d without the fallback), fallback rate %.3f; plain converged %d; "
        "Newton iterations where both converged (LQ stage included): learned mean %.2f median %.1f, plain mean "
        "%.2f median %.1f" % (conv.sum(), warm_conv.sum(), len(fallback) / len(insts), pconv.sum(), its[both].mean(),
                              np.median(its[both]), pits[both].mean(), np.median(pits[both])))
    # the gain: more instances converge (the plain guess stalls on many
    # members of this family), and those both forms solve take no more
    # iterations from the network's guess (median: an instance converted by
    # the fallback carries the iterations of every stage it went through)
    assert conv.sum() > pconv.sum()
    assert np.median(its[both]) <= np.median(pits[both])
    for k in np.nonzero(conv)[0][:12]:
        res, mx = oracle_accepts(insts[k], sols[k], M_STAGES, tol)
        assert res < tol and mx <= 1.0, (k, res, mx)
