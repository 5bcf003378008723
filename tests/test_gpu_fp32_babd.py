"""The FP32 cyclic reduction (cs_set_babd_precision(h, 32)) on the GPU.

Every level stores float32 rows and records (the substitution pairing
computes in float32, the Householder pairing in float64); the boundary-block
solve and the back substitution compute in float64.  Its index arithmetic is
the float64 reduction's (tests/test_cr_schedule_audit.py checks the element
ranges of both stored types against the allocations on the CPU).

Tolerances (relative to the largest entry of the float64 oracle's
direction; n_x = 20 family, seeded state):
* float64 stage Jacobians and condensation, FP32 cyclic reduction: the
  direction (dy, dp, dydot, dz) within 2e-4 (float32 rows over 9 to 11
  levels; up to 9e-5 measured with 7 substitution levels at Nint = 2048);
  after one refinement step within 1e-7, after two within 1e-11: the
  refinement's residual uses the float64 Jacobians, so the refined direction
  is the float64 one;
* with the FP32 direction as well: within 2e-4, refined within 1e-6 (the
  float32 Jacobian entries' bound);
* whole solves (FP32 direction, FP32 cyclic reduction, one refinement
  step): converged, at most one Newton iteration more per continuation
  stage than the float64 path, its solution to 1e-8, the oracle's residual
  of the result < 1e-8;
* a batch of 12 members of the bench's family at the bench's pairing
  (1/16), plain guess + continuation, FP32 direction and FP32 reduction:
  every instance that converges also converges with the float64 reduction,
  to the same solution within 1e-8; at most one of the float64 reduction's
  converged instances is lost (the float32 rows' 1e-4 direction error can
  turn the damped-Newton path of a hard plain-guess instance: member 4
  needs 33 iterations in its second stage with float64 rows and does not
  converge with float32 rows); at least 5 of 12 converge; graph mode is
  bit-equal to the host loop.
"""
import numpy as np
import pytest

from colloc_amd import problems, solver
from oracle import babd as obabd
from oracle import families
from oracle import solver as osolver
from oracle import tables as otables

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def family_of(inst):
    return families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)


@pytest.mark.parametrize("nint,spans", [(512, 0), (512, 16), (512, 64), (2048, 16)])
def test_fp32_cyclic_reduction_direction_against_oracle(nint, spans):
    inst, m = problems.config(1, nint=nint)
    bvp = solver.BvpDae(inst.ocp_text())
    prob = family_of(inst)
    N = bvp.N
    t = np.linspace(0.0, inst.T, N)
    rng = np.random.default_rng(21)
    y = rng.normal(scale=0.3, size=(N, prob.size_y))
    yd = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_y))
    zt = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_z))
    p = rng.normal(scale=0.3, size=prob.size_p)
    rk = otables.lobatto(m)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    dy_o, dp_o = obabd.solve_babd(prob.size_y, prob.size_p, N, A_o, C_o, H_o, b_o, B0, Bn, Vn, rbc_o, M=8)
    _, dyd_o, dz_o = osolver.recover_delta_k(prob, m, N, dy_o, dp_o, fa_o, J, V, W)
    d_o = (dy_o, dyd_o, dz_o, dp_o)
    S = solver.CollocationSolver(bvp, m, n_max=N)
    S.set_mesh(t)
    S.set_state(y, yd, zt, p)
    S.set_babd_substitution(spans)
    S.set_babd_precision(32)
    for bits, bound0, bound2 in ((64, 2e-4, 1e-11), (32, 2e-4, 1e-6)):
        S.set_direction_precision(bits)
        S.compute_f_q(0.1, outputs=False)
        S.construct_jacobian(0.1, outputs=False)
        S.reduce_jacobian(outputs=False)
        S.solve_babd()
        S.recover_delta_k()
        errs = [max(rel_err(a, b) for a, b in zip(S.get_direction(), d_o))]
        for _ in range(2):
            S.refine_direction(1)
            S.compute_f_q(0.1, outputs=False)
            errs.append(max(rel_err(a, b) for a, b in zip(S.get_direction(), d_o)))
        assert S.status() == 0
        say("Nint %d spans %d: FP%d direction + FP32 cyclic reduction: error %.1e, refined %.1e, %.1e"
            % (nint, spans, bits, *errs))
        assert errs[0] < bound0
        assert errs[2] < bound2
        if bits == 64:
            assert errs[1] < 1e-7
    S.close()


def test_fp32_cyclic_reduction_solve():
    inst, m = problems.config(1, nint=1024)
    bvp = solver.BvpDae(inst.ocp_text())
    hom = inst.homotopy_schedule()
    say("float64 solve")
    r64 = solver.collocation_solver_parallel(bvp, m, homotopy=hom)
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    S.set_direction_precision(32)
    S.set_babd_precision(32)
    S.set_refinement(1)
    say("FP32 direction + FP32 cyclic reduction + 1 refinement step")
    r32 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    S.close()
    it64 = [h["iterations"] for h in r64["history"]]
    it32 = [h["iterations"] for h in r32["history"]]
    errs = {k: rel_err(r32[k], r64[k]) for k in ("y", "z", "p")}
    say("iterations float64 %s, FP32 %s; solution difference %s"
        % (it64, it32, {k: "%.1e" % v for k, v in errs.items()}))
    assert r32["success"] == 1 and r32["max_residual"] <= 1.0
    assert len(it32) == len(it64) and all(a <= b <= a + 1 for a, b in zip(it64, it32))
    for k, v in errs.items():
        assert v < 1e-8, k
    f_q = osolver.compute_f_q(family_of(inst), m, r32["N"], otables.lobatto(m), r32["t"], r32["y"], r32["y_dot"],
                              r32["z_tilde"], r32["p"], 0.1)[0]
    assert np.max(np.abs(f_q)) < bvp.tolerance


def test_fp32_cyclic_reduction_batch_bench_settings():
    fam = problems.batch_family(1024)
    insts = [fam[i] for i in range(12)]
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(i.nodes for i in insts), batch=len(insts))
    S.set_direction_precision(32)
    S.set_babd_substitution(16)
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    cons = np.array([i.constant_values() for i in insts])
    hom = [np.array([i.constant_values(s=0.0) for i in insts])]
    say("Nint %s" % [len(t) - 1 for t in ts])

    def run():
        r = solver.solve_batch(S, ts, [np.ones((len(t), S.ny)) for t in ts], [np.ones((len(t), S.nz)) for t in ts],
                               [np.ones(S.np) for _ in ts], cons, hom, tol=bvp.tolerance,
                               max_iter=bvp.maximum_newton_iterations)
        return r, S.get_state()
    out = {}
    for bits, graph in ((64, False), (32, False), (32, True)):
        S.set_babd_precision(bits)
        S.set_newton_graph(graph)
        out[(bits, graph)] = run()
        r = out[(bits, graph)][0]
        say("FP%d cyclic reduction, graph %d: iterations %s, converged %s"
            % (bits, graph, r["iterations"].tolist(), r["converged"].tolist()))
    assert S.status() == 0
    S.close()
    r64, s64 = out[(64, False)]
    r32, s32 = out[(32, False)]
    assert not np.any(r32["converged"] & ~r64["converged"])
    assert r32["converged"].sum() >= r64["converged"].sum() - 1
    assert r32["converged"].sum() >= 5   # 5 or 6 of these 12 converge from the plain guess
    off = np.concatenate([[0], np.cumsum([len(t) for t in ts])])
    for b in np.flatnonzero(r32["converged"]):
        sl = slice(off[b], off[b + 1])
        assert rel_err(s32[0][sl], s64[0][sl]) < 1e-8, b
    rg, sg = out[(32, True)]
    for k in ("iterations", "norm", "max_residual", "converged"):
        np.testing.assert_array_equal(r32[k], rg[k], err_msg=k)
    for a, b in zip(s32, sg):
        np.testing.assert_array_equal(a, b)
