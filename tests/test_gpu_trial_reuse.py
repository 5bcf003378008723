"""The Newton loops take each iterate's residual from its accepted line-search
trial (the trial state and the accepted update are one axpy with the same
step, and the trial's residual launch writes every output), so after a
solve whose last iteration took a step, the residual buffers the loop leaves
(stage states, f_a, f_b, r_bc) are the residual of the final state.  Checked
through what reads them: the stage Jacobians (from the stage states) and the
condensed blocks A, H, b (b from f_a and f_b) built straight after the loop
must equal, bit for bit, those built after recomputing the residual at the
same state; for the host-driven loop and the device-resident one, on a small
batch of n_x = 4 family members stopped after three iterations.
"""
import numpy as np
import pytest

from colloc_amd import problems, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [False, True])
def test_loop_leaves_the_final_state_residual(graph):
    insts = problems.batch_family(3, seed=2, nx=4, nu=2, nint_range=(40, 80))
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(len(t) for t in ts), batch=len(insts))
    S.set_newton_graph(graph)
    S.set_batch_mesh(ts)
    S.set_batch_constants(np.array([i.constant_values() for i in insts]))
    S.form_initial_input(np.ones((sum(len(t) for t in ts), S.ny)), np.ones((sum(len(t) for t in ts), S.nz)),
                         np.ones((len(ts), S.np)))
    alpha = 0.1
    it, nrm, hist = S.newton(alpha, 1e-300, 3)
    for h in hist:   # every iteration took a step: the norms fall
        h = np.asarray(h)[:3]
        assert np.all(np.diff(h) < 0.0), h
    Df0, Dg0, Dr0 = S.construct_jacobian(alpha)
    A0, _, H0, b0 = S.reduce_jacobian()
    fresh = S.compute_f_q(alpha)
    Df1, Dg1, Dr1 = S.construct_jacobian(alpha)
    A1, _, H1, b1 = S.reduce_jacobian()
    for name, x, y in (("Df", Df0, Df1), ("Dg", Dg0, Dg1), ("Dr", Dr0, Dr1), ("A", A0, A1), ("H", H0, H1),
                       ("b", b0, b1)):
        np.testing.assert_array_equal(x, y, err_msg=name)
    # and the norm the loop reports is the residual of the state before the last step
    assert np.all(np.asarray(fresh[0]) < np.asarray(nrm))
    assert S.status() == 0
    S.close()
