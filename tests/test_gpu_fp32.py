"""The FP32 direction (cs_set_direction_precision(h, 32)) on the GPU.

The condensation of the Newton direction runs in float32: the stage
eliminations and the series solve of the stage systems.  Residuals, the BABD
solve, the recovery, the line-search merit and the convergence test stay
float64.  The direction is then an inexact-Newton direction.

Tolerances (relative to the largest entry of the compared float64 oracle
quantity):
  - the condensed rows A, H, b: 1e-5 (float32 arithmetic on O(1) values);
  - the direction delta: 1e-4 (the condensed system's error, amplified by the
    BABD solve);
  - whole solves: converged to the float64 tolerance 1e-8 by the float64
    residual, to the float64 path's solution within 1e-7 (both solutions meet
    1e-8 on the same discrete problem), in at most two more Newton
    iterations than the float64 path.
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


@pytest.mark.parametrize("kernel", [0, 1])
def test_fp32_direction_stages_against_oracle(kernel):
    """kernel 0: the series products on matrix cores; 1: the vector-ALU series."""
    inst, m = problems.config(1, nint=512)
    bvp = solver.BvpDae(inst.ocp_text())
    prob = family_of(inst)
    N = bvp.N
    S = solver.CollocationSolver(bvp, m, n_max=N)
    S.set_direction_precision(32)
    S.set_condense_kernel(kernel)
    t = np.linspace(0.0, inst.T, N)
    rng = np.random.default_rng(21)
    y = rng.normal(scale=0.3, size=(N, prob.size_y))
    yd = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_y))
    zt = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_z))
    p = rng.normal(scale=0.3, size=prob.size_p)
    S.set_mesh(t)
    S.set_state(y, yd, zt, p)
    say("fp32 direction stages N=%d condensation kernel %d" % (N, kernel))
    S.compute_f_q(0.1, outputs=False)
    S.construct_jacobian(0.1, outputs=False)
    A, C, H, b = S.reduce_jacobian()
    dy, dp = S.solve_babd()
    dyd, dz = S.recover_delta_k()
    fb_count = S.condense_fallbacks()
    assert S.status() == 0
    S.close()
    rk = otables.lobatto(m)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    errs = {"A": rel_err(A, A_o), "b": rel_err(b, b_o)}
    dy_o, dp_o = obabd.solve_babd(prob.size_y, prob.size_p, N, A_o, C_o, H_o, b_o, B0, Bn, Vn, rbc_o, M=8)
    _, dyd_o, dz_o = osolver.recover_delta_k(prob, m, N, dy_o, dp_o, fa_o, J, V, W)
    errs.update({"dy": rel_err(dy, dy_o), "dp": rel_err(dp, dp_o), "dyd": rel_err(dyd, dyd_o),
                 "dz": rel_err(dz, dz_o)})
    say("fp32 direction errors vs the float64 oracle: %s (float64 fallbacks %d)"
        % ({k: "%.1e" % v for k, v in errs.items()}, fb_count))
    assert fb_count == 0
    assert np.all(H == 0.0)
    assert errs["A"] < 1e-5 and errs["b"] < 1e-5
    for k in ("dy", "dp", "dyd", "dz"):
        assert errs[k] < 1e-4, k


@pytest.mark.parametrize("label", ["tanh4_m5", "param_osc_m4", "tanh4_batch_m5"])
def test_matrix_core_condensation_small_blocks(label):
    """The matrix-core series on stage blocks smaller than one 16 x 16 tile
    (ny = 8, 4; an active parameter column; three stage blocks with m = 4),
    against the float64 oracle's condensed rows (float32 arithmetic: 1e-5
    relative to the largest entry) and against the vector-ALU series kernel
    (the same float32 system, solved to the same series tolerance: 1e-5)."""
    from symbolic_problem import SymbolicProblem
    inst, m = [(i, mm) for lab, i, mm in problems.standard_units() if lab == label][0]
    text = inst.ocp_text() if hasattr(inst, "ocp_text") else inst.text
    bvp = solver.BvpDae(text)
    prob = SymbolicProblem(text)
    N = 65
    S = solver.CollocationSolver(bvp, m, n_max=N)
    S.set_direction_precision(32)
    rng = np.random.default_rng(3)
    t = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.0, 1.0, N - 2)]))
    y = rng.normal(scale=0.3, size=(N, prob.size_y))
    yd = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_y))
    zt = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_z))
    p = rng.normal(scale=0.3, size=prob.size_p)
    S.set_mesh(t)
    S.set_state(y, yd, zt, p)
    out = {}
    for kernel in (0, 1):
        say("%s: condensation kernel %d" % (label, kernel))
        S.set_condense_kernel(kernel)
        S.compute_f_q(0.1, outputs=False)
        S.construct_jacobian(0.1, outputs=False)
        out[kernel] = S.reduce_jacobian()
        assert S.condense_fallbacks() == 0 and S.status() == 0
    S.close()
    # the handle's instance constants are the text's values, which the
    # oracle's problem substitutes as well
    rk = otables.lobatto(m)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    for kernel, (A, C, H, b) in out.items():
        errs = (rel_err(A, A_o), rel_err(H, H_o) if np.any(H_o) else np.max(np.abs(H)), rel_err(b, b_o))
        say("kernel %d: A, H, b errors vs the float64 oracle %.1e %.1e %.1e" % ((kernel,) + errs))
        assert max(errs) < 1e-5, (kernel, errs)
    for k in (0, 2, 3):
        assert rel_err(out[0][k], out[1][k]) < 1e-5


@pytest.mark.parametrize("nint", [200, 1024])
def test_fp32_direction_solve(nint):
    inst, m = problems.config(1, nint=nint)
    bvp = solver.BvpDae(inst.ocp_text())
    hom = inst.homotopy_schedule()
    say("float64 solve nint=%d" % nint)
    r64 = solver.collocation_solver_parallel(bvp, m, homotopy=hom)
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    S.set_direction_precision(32)
    say("fp32-direction solve nint=%d" % nint)
    r32 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    S.close()
    it64 = [h["iterations"] for h in r64["history"]]
    it32 = [h["iterations"] for h in r32["history"]]
    errs = {k: rel_err(r32[k], r64[k]) for k in ("y", "z", "p")}
    say("iterations float64 %s, fp32 direction %s; solution difference %s"
        % (it64, it32, {k: "%.1e" % v for k, v in errs.items()}))
    assert r32["success"] == 1 and r32["max_residual"] <= 1.0
    assert all(b <= a + 2 for a, b in zip(it64, it32)) and len(it32) == len(it64)
    for k, v in errs.items():
        assert v < 1e-7, k
    prob = family_of(inst)
    f_q = osolver.compute_f_q(prob, m, r32["N"], otables.lobatto(m), r32["t"], r32["y"], r32["y_dot"],
                              r32["z_tilde"], r32["p"], 0.1)[0]
    assert np.max(np.abs(f_q)) < bvp.tolerance


def test_precision_switches_refill_static_jacobian_entries():
    """One handle switching 32 -> 64 -> 32: the stage Jacobian buffer holds
    float32 or float64 entries, and its static entries (zeros, compile-time
    constants) are written once per stored type.  After every switch the
    whole Jacobian must match the oracle's (float32 evaluation and storage:
    2e-6 relative to the largest entry, a few float32 roundings of O(1)
    terms; float64: 1e-13)."""
    inst, m = problems.config(1, nint=64)
    bvp = solver.BvpDae(inst.ocp_text())
    prob = family_of(inst)
    N = bvp.N
    S = solver.CollocationSolver(bvp, m, n_max=N)
    t = np.linspace(0.0, inst.T, N)
    rng = np.random.default_rng(5)
    y = rng.normal(scale=0.3, size=(N, prob.size_y))
    yd = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_y))
    zt = rng.normal(scale=0.3, size=((N - 1) * m, prob.size_z))
    p = rng.normal(scale=0.3, size=prob.size_p)
    S.set_mesh(t)
    S.set_state(y, yd, zt, p)
    _, yt, _, _, _ = S.compute_f_q(0.1)
    Df_o, Dg_o = prob.Df(yt, zt, p), prob.Dg(yt, zt, p)
    for bits in (32, 64, 32, 64):
        S.set_direction_precision(bits)
        Df, Dg, _ = S.construct_jacobian(0.1)
        errs = (rel_err(Df, Df_o), rel_err(Dg, Dg_o))
        say("precision %d: Jacobian errors %.1e %.1e" % (bits, errs[0], errs[1]))
        bound = 2e-6 if bits == 32 else 1e-13
        assert errs[0] < bound and errs[1] < bound, (bits, errs)
        # entries zero in every stage (structural zeros) are exactly zero
        zero = np.all(Df_o == 0, axis=0)
        assert zero.any() and np.all(Df[:, zero] == 0)
    assert S.status() == 0
    S.close()


@pytest.mark.parametrize("nx,nu", [(20, 10), (4, 2)])
def test_matrix_core_jacobian_equals_straight_line(nx, nu):
    """The float32 stage Jacobians on matrix cores (the entries as the jac
    program's coefficient matrix times the stage's monomials) against the
    straight-line float32 kernel (itself pinned against the oracle by the
    tests above): a batch of family members on uneven meshes, seeded
    states.  The two agree within 2e-6 of the largest entry, with the same
    zero pattern; Dr (straight-line code in both) is identical; the Newton
    directions they give agree to the FP32 direction's 1e-4."""
    from colloc_amd import native
    insts = problems.batch_family(5, seed=1, nx=nx, nu=nu, nint_range=(40, 120))
    bvp = solver.BvpDae(insts[0].ocp_text())
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    Nt = sum(len(t) for t in ts)
    S = solver.CollocationSolver(bvp, 5, n_max=Nt, batch=len(insts))
    assert S.get_launch(native.CS_LAUNCH_JACOBIAN_GEMM) == 1
    S.set_batch_mesh(ts)
    S.set_batch_constants(np.array([i.constant_values() for i in insts]))
    rng = np.random.default_rng(4)
    nst = (Nt - len(ts)) * 5
    S.set_state(rng.normal(scale=0.4, size=(Nt, S.ny)), rng.normal(scale=0.4, size=(nst, S.ny)),
                rng.normal(scale=0.4, size=(nst, S.nz)), rng.normal(scale=0.4, size=(len(ts), S.np)))
    S.set_direction_precision(32)
    S.compute_f_q(0.1, outputs=False)
    out = {}
    for on in (1, 0):
        S.set_launch(native.CS_LAUNCH_JACOBIAN_GEMM, on)
        say("n_x=%d float32 Jacobian, matrix cores %d" % (nx, on))
        Df, Dg, Dr = S.construct_jacobian(0.1)
        S.reduce_jacobian(outputs=False)
        S.solve_babd()
        S.recover_delta_k()
        out[on] = (Df, Dg, Dr) + tuple(S.get_direction())
    for x, y in zip(out[1][:2], out[0][:2]):
        assert rel_err(x, y) < 2e-6
        np.testing.assert_array_equal(x == 0, y == 0)
    np.testing.assert_array_equal(out[1][2], out[0][2])   # Dr: the same straight-line code
    for x, y in zip(out[1][3:], out[0][3:]):
        assert rel_err(x, y) < 1e-4
    assert S.status() == 0
    S.close()
