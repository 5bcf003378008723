"""Constrained optimal-control problems (the reference's example set,
Notebook/source, restated in problems.CONSTRAINED_EXAMPLES) solved on the GPU
through the whole driver of collocation_solver_parallel_2d_shared.py:105-213:
alpha continuation (alpha <- 0.9 alpha from 0.1 down to 1e-6) of the
smoothed complementarity of the inequalities, the segment-residual
acceptance test and remeshing, from the all-ones guess (the examples'
InputFiles are not shipped with the reference).  The oracle runs the same
driver on the same text, with its problem functions built from the same
symbolic BVP-DAE (tests/symbolic_problem.py, pinned by finite differences).

* trolley (control bounds) and free-space rocket (control and state-variable
  bounds, the final time a parameter): both converge, and the GPU run goes
  through the same continuation steps as the oracle: the same alpha, the
  same mesh size N and the same Newton iteration count at every step, the
  remeshes at the same steps; the final y, z and p agree to 1e-8 relative
  (the problems' tolerance is 1e-6; both runs follow the same iterates up to
  rounding).
* container crane and underwater vehicle: from the all-ones guess the
  reference algorithm does not converge at alpha = 0.1; the GPU's norm
  history over the first 30 Newton iterations equals the oracle's (1e-6
  relative).  The container problem has a parameter in its dynamics (the
  BABD rows carry an H block).
"""
import numpy as np
import pytest

from colloc_amd import problems, solver
from oracle import solver as osolver
from oracle import tables as otables
from symbolic_problem import SymbolicProblem

pytestmark = pytest.mark.gpu
M = 4


def say(msg):
    print(msg, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", ["trolley", "rocket"])
def test_continuation_and_remesh_match_oracle(name):
    text = problems.CONSTRAINED_EXAMPLES[name]
    bvp = solver.BvpDae(text)
    say("%s: ny=%d nz=%d np=%d N=%d, %d + %d inequalities" % (name, bvp.size_y, bvp.size_z, bvp.size_p, bvp.N,
                                                            bvp.size_inequality, bvp.size_sv_inequality))
    res = solver.collocation_solver_parallel(bvp, M)
    say("GPU: success %d, %d steps, final alpha %.3g, N %d" % (res["success"], len(res["history"]), res["alpha"],
                                                               res["N"]))
    prob = SymbolicProblem(text)
    ref = osolver.collocation_solver(prob, M, bvp.T0, bvp.Y0, bvp.Z0, bvp.P0, tol=bvp.tolerance,
                                     max_iter=bvp.maximum_newton_iterations, max_mesh=bvp.maximum_mesh_refinements,
                                     max_nodes=bvp.maximum_nodes, inequalities=True)
    say("oracle: success %d, %d steps, final alpha %.3g, N %d" % (ref["success"], len(ref["history"]), ref["alpha"],
                                                                  ref["t"].size))
    assert res["success"] == 1 and ref["success"] == 1
    gh = [(h["alpha"], h["N"], h["iterations"]) for h in res["history"]]
    oh = [(h["alpha"], h["N"], h["iterations"]) for h in ref["history"]]
    remesh = [k for k, h in enumerate(ref["history"]) if h["max_residual"] > 1.0]
    say("remeshes at steps %s, N %s" % (remesh, sorted(set(n for _, n, _ in oh))))
    assert len(gh) == len(oh)
    for k, (g, o) in enumerate(zip(gh, oh)):
        assert g[1] == o[1] and g[2] == o[2] and abs(g[0] - o[0]) <= 1e-15 * o[0], (k, g, o)
    assert res["alpha"] <= 1e-6 and abs(res["alpha"] - ref["alpha"]) <= 1e-15 * ref["alpha"]
    errs = {k: rel_err(res[k], ref[k]) for k in ("y", "z", "p")}
    say("final solution differences: %s" % {k: "%.1e" % v for k, v in errs.items()})
    np.testing.assert_array_equal(res["t"], ref["t"])
    for k, v in errs.items():
        assert v < 1e-8, k
    assert res["max_residual"] <= 1.0


@pytest.mark.parametrize("name", ["container", "underwater"])
def test_nonconverging_examples_stall_identically(name):
    text = problems.CONSTRAINED_EXAMPLES[name]
    bvp = solver.BvpDae(text)
    its = 30
    S = solver.CollocationSolver(bvp, M, n_max=bvp.N)
    S.set_mesh(bvp.T0)
    S.form_initial_input(bvp.Y0, bvp.Z0, bvp.P0)
    say("%s: %d Newton iterations at alpha = 0.1 (ny=%d nz=%d np=%d)" % (name, its, bvp.size_y, bvp.size_z,
                                                                           bvp.size_p))
    it, nrm, hist = S.newton(0.1, bvp.tolerance, its)
    S.close()
    prob = SymbolicProblem(text)
    rk = otables.lobatto(M)
    y, yd, zt = osolver.form_initial_input(prob, M, bvp.N, bvp.Y0, bvp.Z0, bvp.P0, rk)
    _, _, _, _, it_o, norms_o = osolver.newton_solve(prob, M, rk, bvp.T0, y, yd, zt, np.array(bvp.P0, float), 0.1,
                                                     bvp.tolerance, its)
    say("GPU norms %s ..." % np.array2string(np.asarray(hist[:6]), precision=4))
    say("oracle    %s ..." % np.array2string(np.asarray(norms_o[:6]), precision=4))
    assert it == it_o == its - 1 and nrm >= bvp.tolerance
    np.testing.assert_allclose(hist, norms_o, rtol=1e-6)
