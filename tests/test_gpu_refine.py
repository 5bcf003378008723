"""Iterative refinement of the Newton direction (cs_set_refinement,
cs_refine_direction) and the linear residual it is built on
(cs_linear_residual: rho = F + J delta of the linearised collocation system,
in float64, with the stored stage Jacobians and Dr).

Tolerances (relative to max |F| of the state, or to the largest entry of
the compared oracle quantity):
* the float64 direction solves J delta = -F (it matches the oracle's
  direction to 1e-10), so its linear residual vanishes to rounding:
  <= 1e-11 max |F|, on the n_x = 20 family and on a problem whose dynamics
  depend on a parameter (the p columns of Df, Dg and the condensation's
  active-parameter columns); after the state y moves by 10% (F changes, the
  direction does not) it is >= 1e-3 max |F|;
* the FP32 direction leaves a linear residual of order 1e-6 max |F|; one
  refinement step takes it below 1e-9 max |F|, two below 1e-12: the
  direction converges to the exact solution of the stored system, whose
  float32 Jacobian entries (a few float32 roundings) bound its distance to
  the float64 oracle's direction: 1e-6 after refinement, against 1e-4
  without (the FP32-direction bound);
* whole solves with the FP32 direction and one refinement step: at most one
  Newton iteration more per continuation stage than the float64 path (the
  linear-quadratic stage, which one exact float64 step solves, takes a
  second: the refined direction is 1.6e-7 from the exact one, not 0), its
  solution to 1e-8, the oracle's residual of the result < 1e-8; in graph
  mode bit-equal to the host loop.
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


def seeded(S, bvp, m, t, seed):
    rng = np.random.default_rng(seed)
    N = len(t)
    y = rng.normal(scale=0.3, size=(N, bvp.size_y))
    yd = rng.normal(scale=0.3, size=((N - 1) * m, bvp.size_y))
    zt = rng.normal(scale=0.3, size=((N - 1) * m, bvp.size_z))
    p = rng.normal(scale=0.3, size=bvp.size_p)
    S.set_mesh(t)
    S.set_state(y, yd, zt, p)
    return y, yd, zt, p


def direction(S):
    fnorm = S.compute_f_q(0.1, outputs=False)[0]
    S.construct_jacobian(0.1, outputs=False)
    S.reduce_jacobian(outputs=False)
    S.solve_babd()
    S.recover_delta_k()
    return fnorm


def oracle_direction(prob, m, t, y, yd, zt, p):
    N = len(t)
    rk = otables.lobatto(m)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    dy_o, dp_o = obabd.solve_babd(prob.size_y, prob.size_p, N, A_o, C_o, H_o, b_o, B0, Bn, Vn, rbc_o, M=8)
    _, dyd_o, dz_o = osolver.recover_delta_k(prob, m, N, dy_o, dp_o, fa_o, J, V, W)
    return dy_o, dyd_o, dz_o, dp_o


def check_problem(bvp, prob, m, t, seed, label):
    S = solver.CollocationSolver(bvp, m, n_max=len(t))
    y, yd, zt, p = seeded(S, bvp, m, t, seed)
    d_o = oracle_direction(prob, m, t, y, yd, zt, p)
    # float64 direction: rho vanishes to rounding
    S.set_direction_precision(64)
    fnorm = direction(S)
    lr64 = S.linear_residual()
    err64 = max(rel_err(a, b) for a, b in zip(S.get_direction(), d_o))
    S.set_state(y * 1.1, yd, zt, p)
    S.compute_f_q(0.1, outputs=False)
    lr_moved = S.linear_residual()
    S.set_state(y, yd, zt, p)
    # FP32 direction, then refinement steps (F recomputed before each: a
    # refinement step leaves rho in f_a, f_b, r_bc)
    S.set_direction_precision(32)
    direction(S)
    lr32 = S.linear_residual()
    err32 = max(rel_err(a, b) for a, b in zip(S.get_direction(), d_o))
    lrs, errs = [], []
    for _ in range(2):
        S.refine_direction(1)
        S.compute_f_q(0.1, outputs=False)
        lrs.append(S.linear_residual())
        errs.append(max(rel_err(a, b) for a, b in zip(S.get_direction(), d_o)))
    assert S.status() == 0
    S.close()
    say("%s N=%d: max|F| %.3g; float64 direction: rho %.1e, error %.1e; state moved: rho %.1e"
        % (label, len(t), fnorm, lr64, err64, lr_moved))
    say("  FP32 direction: rho %.1e, error %.1e; refined: rho %s, error %s"
        % (lr32, err32, ["%.1e" % v for v in lrs], ["%.1e" % v for v in errs]))
    assert err64 < 1e-10
    assert lr64 <= 1e-11 * fnorm
    assert lr_moved >= 1e-3 * fnorm
    assert lr32 >= 1e-9 * fnorm and err32 < 1e-4
    assert lrs[0] <= 1e-9 * fnorm and lrs[1] <= 1e-12 * fnorm
    assert errs[0] < 1e-6 and errs[1] < 1e-6 and errs[0] < err32


@pytest.mark.parametrize("nint", [512, 2048])
def test_linear_residual_and_refinement_family(nint):
    inst, m = problems.config(1, nint=nint)
    bvp = solver.BvpDae(inst.ocp_text())
    prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)
    check_problem(bvp, prob, m, np.linspace(0.0, inst.T, bvp.N), 21, "tanh n_x=20")


def test_linear_residual_and_refinement_active_parameter():
    from symbolic_problem import SymbolicProblem
    text = problems.PARAM_OSC_OCP
    bvp = solver.BvpDae(text)
    prob = SymbolicProblem(text)
    rng = np.random.default_rng(3)
    N = 65
    t = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.0, 1.0, N - 2)]))
    check_problem(bvp, prob, 4, t, 8, "param_osc")


def test_refined_fp32_solve_matches_float64():
    inst, m = problems.config(1, nint=1024)
    bvp = solver.BvpDae(inst.ocp_text())
    hom = inst.homotopy_schedule()
    say("float64 solve")
    r64 = solver.collocation_solver_parallel(bvp, m, homotopy=hom)
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    S.set_direction_precision(32)
    S.set_refinement(1)
    say("FP32 direction + 1 refinement step")
    r32 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    S.close()
    it64 = [h["iterations"] for h in r64["history"]]
    it32 = [h["iterations"] for h in r32["history"]]
    errs = {k: rel_err(r32[k], r64[k]) for k in ("y", "z", "p")}
    say("iterations float64 %s, refined FP32 %s; solution difference %s"
        % (it64, it32, {k: "%.1e" % v for k, v in errs.items()}))
    assert r32["success"] == 1 and r32["max_residual"] <= 1.0
    assert len(it32) == len(it64) and all(a <= b <= a + 1 for a, b in zip(it64, it32))
    for k, v in errs.items():
        assert v < 1e-8, k
    prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)
    f_q = osolver.compute_f_q(prob, m, r32["N"], otables.lobatto(m), r32["t"], r32["y"], r32["y_dot"],
                              r32["z_tilde"], r32["p"], 0.1)[0]
    assert np.max(np.abs(f_q)) < bvp.tolerance


def test_refinement_in_graph_equals_host_loop():
    fam = problems.batch_family(1024)
    insts = [fam[i] for i in (0, 2, 5)]
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(i.nodes for i in insts), batch=len(insts))
    S.set_direction_precision(32)
    S.set_babd_substitution(16)
    S.set_refinement(1)
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    cons = np.array([i.constant_values() for i in insts])
    hom = [np.array([i.constant_values(s=0.0) for i in insts])]
    out = {}
    for graph in (False, True):
        S.set_newton_graph(graph)
        r = solver.solve_batch(S, ts, [np.ones((len(t), S.ny)) for t in ts], [np.ones((len(t), S.nz)) for t in ts],
                               [np.ones(S.np) for _ in ts], cons, hom, tol=bvp.tolerance,
                               max_iter=bvp.maximum_newton_iterations)
        out[graph] = (r, S.get_state())
        say("graph %d: iterations %s, converged %s" % (graph, r["iterations"].tolist(), r["converged"].tolist()))
    S.close()
    for k in ("iterations", "norm", "max_residual", "converged"):
        np.testing.assert_array_equal(out[False][0][k], out[True][0][k], err_msg=k)
    for a, b in zip(out[False][1], out[True][1]):
        np.testing.assert_array_equal(a, b)
