"""The n_x = 64 tanh configuration (ny = 128, nz = 32, np = 128, m = 4): the
polynomial table form of its callbacks and the large-block kernels
(colloc_large.h: stage eliminations, the series on matrix cores, the
cyclic reduction with work matrices in device memory), against the oracle.

* stage by stage at Nint = 48 on seeded states, with the Householder
  pairing on every level and with substitution levels: f_a, f_b, r_bc to
  1e-13, Df, Dg, Dr to 1e-13, the condensed A, H, b to 1e-11, the Newton
  direction (dy, dp, dydot, dz) to 1e-10 against the oracle's dense
  condensation and M = 8 partitioned QR;
* the FP32 direction (float32 Jacobians, series and records): the direction
  within 1e-4 of the oracle's, one refinement step within 1e-6;
* the whole plain solve (all-ones guess, homotopy S = 0 -> 0.5) at Nint = 64:
  the oracle's Newton iteration counts, y, z, p within 1e-9, the oracle's
  residual of the GPU solution below the tolerance;
* the full configuration (Nint = 4096): float64 and FP32 directions both
  converge; the oracle's residual of the float64 solution is below the
  tolerance, and the FP32-direction solution agrees with it to 1e-7;
* the FP32 cyclic reduction (float32 rows and records after level 0, the
  substitution products on float32 matrix cores): with float64 Jacobians the
  direction within 2e-4 of the oracle's, one refinement step within 1e-7;
  with the FP32 direction as well within 2e-4, refined within 1e-6;
  Householder-only and substitution pairings; and the full configuration
  solved with the FP32 direction, FP32 cyclic reduction and one refinement
  step to the float64 solution within 1e-7.
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


def family_of(inst, s=None):
    return families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s if s is None else s, q=inst.q,
                               c_u=inst.c_u)


def seeded(bvp, m, N, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return (rng.normal(scale=scale, size=(N, bvp.size_y)), rng.normal(scale=scale, size=((N - 1) * m, bvp.size_y)),
            rng.normal(scale=scale, size=((N - 1) * m, bvp.size_z)), rng.normal(scale=scale, size=bvp.size_p))


@pytest.fixture(scope="module")
def large():
    inst, m = problems.config(2, nint=48)
    bvp = solver.BvpDae(inst.ocp_text())
    say("n_x=64: ny=%d nz=%d np=%d m=%d N=%d" % (bvp.size_y, bvp.size_z, bvp.size_p, m, bvp.N))
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    prob = family_of(inst)
    rk = otables.lobatto(m)
    rng = np.random.default_rng(5)
    N = bvp.N
    t = np.sort(np.concatenate([[0.0, inst.T], rng.uniform(0.0, inst.T, N - 2)]))
    y, yd, zt, p = seeded(bvp, m, N, 1)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    dy_o, dp_o = obabd.solve_babd(prob.size_y, prob.size_p, N, A_o, C_o, H_o, b_o, B0, Bn, Vn, rbc_o, M=8)
    _, dyd_o, dz_o = osolver.recover_delta_k(prob, m, N, dy_o, dp_o, fa_o, J, V, W)
    ref = dict(t=t, y=y, yd=yd, zt=zt, p=p, f_q=f_q, yt=yt_o, fa=fa_o, fb=fb_o, rbc=rbc_o, A=A_o, H=H_o, b=b_o,
               dy=dy_o, dp=dp_o, dyd=dyd_o, dz=dz_o)
    yield inst, m, bvp, S, prob, ref
    S.close()


@pytest.mark.parametrize("spans", [0, 4])
def test_large_stage_parity(large, spans):
    inst, m, bvp, S, prob, ref = large
    S.set_direction_precision(64)
    S.set_babd_substitution(spans)
    S.set_mesh(ref["t"])
    S.set_state(ref["y"], ref["yd"], ref["zt"], ref["p"])
    say("n_x=64 stages, substitution spans %d: compute_f_q" % spans)
    nrm, yt, fa, fb, rbc = S.compute_f_q(0.1)
    assert rel_err(yt, ref["yt"]) < 1e-14
    assert rel_err(fa, ref["fa"]) < 1e-13
    assert rel_err(fb, ref["fb"]) < 1e-13
    assert rel_err(rbc, ref["rbc"]) < 1e-14
    assert abs(nrm - np.max(np.abs(ref["f_q"]))) <= 1e-13 * np.max(np.abs(ref["f_q"]))
    say("construct_jacobian")
    Df, Dg, Dr = S.construct_jacobian(0.1)
    assert rel_err(Df, prob.Df(ref["yt"], ref["zt"], ref["p"])) < 1e-13
    assert rel_err(Dg, prob.Dg(ref["yt"], ref["zt"], ref["p"])) < 1e-13
    assert rel_err(Dr, prob.Dr(ref["y"][0], ref["y"][-1], ref["p"])) < 1e-14
    say("reduce_jacobian")
    A, C, H, b = S.reduce_jacobian()
    assert rel_err(A, ref["A"]) < 1e-11
    assert rel_err(H, ref["H"]) < 1e-11 or np.max(np.abs(H)) == np.max(np.abs(ref["H"])) == 0
    assert rel_err(b, ref["b"]) < 1e-11
    say("solve_babd")
    dy, dp = S.solve_babd()
    assert rel_err(dy, ref["dy"]) < 1e-10
    assert rel_err(dp, ref["dp"]) < 1e-10
    say("recover_delta_k")
    dyd, dz = S.recover_delta_k()
    assert rel_err(dyd, ref["dyd"]) < 1e-10
    assert rel_err(dz, ref["dz"]) < 1e-10
    assert S.status() == 0


def test_large_fp32_direction(large):
    inst, m, bvp, S, prob, ref = large
    S.set_babd_substitution(4)
    S.set_mesh(ref["t"])
    S.set_state(ref["y"], ref["yd"], ref["zt"], ref["p"])
    S.set_direction_precision(32)
    S.set_refinement(0)
    say("n_x=64 FP32 direction")
    S.compute_f_q(0.1, outputs=False)
    S.construct_jacobian(0.1, outputs=False)
    S.reduce_jacobian(outputs=False)
    S.solve_babd()
    S.recover_delta_k()
    dy, dyd, dz, dp = S.get_direction()
    e32 = max(rel_err(dy, ref["dy"]), rel_err(dp, ref["dp"]), rel_err(dyd, ref["dyd"]), rel_err(dz, ref["dz"]))
    say("FP32 direction error %.3g" % e32)
    assert e32 < 1e-4
    S.refine_direction(1)
    dy, dyd, dz, dp = S.get_direction()
    e1 = max(rel_err(dy, ref["dy"]), rel_err(dp, ref["dp"]), rel_err(dyd, ref["dyd"]), rel_err(dz, ref["dz"]))
    say("refined once %.3g" % e1)
    assert e1 < 1e-6 and e1 < e32
    S.set_direction_precision(64)
    assert S.status() == 0


def compare_solves(res, ref):
    assert res["success"] == 1 and ref["success"] == 1
    assert [h["iterations"] for h in res["history"]] == [h["iterations"] for h in ref["history"]]
    for k in ("y", "z", "p"):
        assert rel_err(res[k], ref[k]) < 1e-9, k


def test_large_solve_parity():
    inst, m = problems.config(2, nint=64)
    bvp = solver.BvpDae(inst.ocp_text())
    say("n_x=64 solve N=%d" % bvp.N)
    res = solver.collocation_solver_parallel(bvp, m, homotopy=inst.homotopy_schedule())
    say("GPU: %s" % [h["iterations"] for h in res["history"]])
    hom = [family_of(inst, s=v[inst.instance_constants().index("S")]) for v in inst.homotopy_schedule()]
    ref = osolver.collocation_solver(family_of(inst), m, bvp.T0, bvp.Y0, bvp.Z0, bvp.P0, tol=bvp.tolerance,
                                     homotopy=hom)
    say("oracle: %s" % [h["iterations"] for h in ref["history"]])
    compare_solves(res, ref)
    f_q = osolver.compute_f_q(family_of(inst), m, res["N"], otables.lobatto(m), res["t"], res["y"], res["y_dot"],
                              res["z_tilde"], res["p"], 0.1)[0]
    assert np.max(np.abs(f_q)) < bvp.tolerance


def test_large_full_size_fp64_and_fp32():
    inst, m = problems.config(2)
    bvp = solver.BvpDae(inst.ocp_text())
    hom = inst.homotopy_schedule()
    say("n_x=64 full size N=%d, float64 direction" % bvp.N)
    # a handle of exactly the mesh's nodes (the solve is accepted without remeshing)
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    r64 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    say("float64: %s, max residual %.3g" % ([h["iterations"] for h in r64["history"]], r64["max_residual"]))
    assert r64["success"] == 1 and r64["N"] == 4097
    f_q = osolver.compute_f_q(family_of(inst), m, r64["N"], otables.lobatto(m), r64["t"], r64["y"], r64["y_dot"],
                              r64["z_tilde"], r64["p"], 0.1)[0]
    assert np.max(np.abs(f_q)) < bvp.tolerance
    S.set_direction_precision(32)
    say("FP32 direction")
    r32 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    say("FP32: %s" % [h["iterations"] for h in r32["history"]])
    S.close()
    assert r32["success"] == 1
    for k in ("y", "z", "p"):
        assert rel_err(r32[k], r64[k]) < 1e-7, k


@pytest.mark.parametrize("spans", [0, 2])
def test_large_fp32_cyclic_reduction(large, spans):
    inst, m, bvp, S, prob, ref = large
    S.set_babd_substitution(spans)
    S.set_mesh(ref["t"])
    S.set_state(ref["y"], ref["yd"], ref["zt"], ref["p"])

    def direction_err():
        dy, dyd, dz, dp = S.get_direction()
        return max(rel_err(dy, ref["dy"]), rel_err(dp, ref["dp"]), rel_err(dyd, ref["dyd"]), rel_err(dz, ref["dz"]))
    try:
        S.set_babd_precision(32)
        for bits, bound, bound1 in ((64, 2e-4, 1e-7), (32, 2e-4, 1e-6)):
            S.set_direction_precision(bits)
            S.set_refinement(0)
            say("n_x=64 FP32 cyclic reduction, spans %d, direction fp%d" % (spans, bits))
            S.compute_f_q(0.1, outputs=False)
            S.construct_jacobian(0.1, outputs=False)
            S.reduce_jacobian(outputs=False)
            S.solve_babd()
            S.recover_delta_k()
            e0 = direction_err()
            S.refine_direction(1)
            e1 = direction_err()
            say("direction error %.3g, refined once %.3g" % (e0, e1))
            assert e0 < bound and e1 < bound1 and e1 < e0
        assert S.status() == 0
    finally:
        S.set_babd_precision(64)
        S.set_direction_precision(64)
        S.set_refinement(0)


def test_large_full_size_fp32_cyclic_reduction():
    inst, m = problems.config(2)
    bvp = solver.BvpDae(inst.ocp_text())
    hom = inst.homotopy_schedule()
    S = solver.CollocationSolver(bvp, m, n_max=bvp.N)
    S.set_babd_substitution(2)
    say("n_x=64 full size, float64")
    r64 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    S.set_direction_precision(32)
    S.set_babd_precision(32)
    S.set_refinement(1)
    say("FP32 direction, FP32 cyclic reduction, one refinement step")
    r32 = solver.collocation_solver_parallel(bvp, m, solver=S, homotopy=hom)
    it64 = [h["iterations"] for h in r64["history"]]
    it32 = [h["iterations"] for h in r32["history"]]
    say("iterations float64 %s, FP32 %s" % (it64, it32))
    assert S.status() == 0
    S.close()
    assert r64["success"] == 1 and r32["success"] == 1
    assert len(it32) == len(it64) and all(b <= a + 2 for a, b in zip(it64, it32))
    for k in ("y", "z", "p"):
        assert rel_err(r32[k], r64[k]) < 1e-7, k
