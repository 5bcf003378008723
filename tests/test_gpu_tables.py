"""The polynomial table form of the callbacks (polyform.py, colloc_tables.h)
on the GPU, against the oracle and against the straight-line form of the
same problem.

* stage outputs of the n_x = 4 tanh instance in table form against the
  oracle: f_a, f_b, r_bc and y-tilde to 1e-13, Df, Dg, Dr to 1e-13, and the
  Newton direction through the unchanged condensation and BABD to 1e-10;
  a batch of 11 instances (two boundary workgroups of 8 instances);
* the constrained trolley (sqrt nodes of the smoothed complementarity) and
  the parameter problem: every stage output, the initial input and the
  segment residual (all variants) of the table form against the straight-
  line form (1e-13 relative), and the Jacobians in float32 (1e-6);
* the trolley's whole driver (alpha continuation to 1e-6 with remeshing)
  in table form: the straight-line form's steps (alpha, N, Newton
  iterations) and solution (1e-9).
"""
import numpy as np
import pytest

from colloc_amd import problems, solver
from oracle import families
from oracle import solver as osolver
from oracle import tables as otables

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def seeded_state(bvp, m, N, seed, scale=0.5, positive=False):
    rng = np.random.default_rng(seed)
    draw = (lambda size: rng.uniform(0.2, 1.0, size)) if positive else (lambda size: rng.normal(scale=scale, size=size))
    return (draw((N, bvp.size_y)), draw(((N - 1) * m, bvp.size_y)), draw(((N - 1) * m, bvp.size_z)),
            draw(bvp.size_p))


def test_table_form_stage_parity_against_oracle():
    inst, m = problems.config(1, nx=4, nu=2, nint=32)
    bvp = solver.BvpDae(inst.ocp_text())
    N = bvp.N
    S = solver.CollocationSolver(bvp, m, n_max=N, form="tables")
    prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)
    rk = otables.lobatto(m)
    rng = np.random.default_rng(3)
    t = np.sort(np.concatenate([[0.0, inst.T], rng.uniform(0.0, inst.T, N - 2)]))
    y, yd, zt, p = seeded_state(bvp, m, N, 0)
    S.set_mesh(t)
    S.set_state(y, yd, zt, p)
    say("table form, n_x=4: compute_f_q")
    nrm, yt, fa, fb, rbc = S.compute_f_q(0.1)
    f_q, yt_o, fa_o, fb_o, rbc_o = osolver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    assert rel_err(yt, yt_o) < 1e-14
    assert rel_err(fa, fa_o) < 1e-13
    assert rel_err(fb, fb_o) < 1e-13
    assert rel_err(rbc, rbc_o) < 1e-14
    assert abs(nrm - np.max(np.abs(f_q))) <= 1e-13 * np.max(np.abs(f_q))
    say("table form, n_x=4: construct_jacobian")
    Df, Dg, Dr = S.construct_jacobian(0.1)
    assert rel_err(Df, prob.Df(yt_o, zt, p)) < 1e-13
    assert rel_err(Dg, prob.Dg(yt_o, zt, p)) < 1e-13
    assert rel_err(Dr, prob.Dr(y[0], y[-1], p)) < 1e-14
    say("table form, n_x=4: direction")
    S.reduce_jacobian(outputs=False)
    dy, dp = S.solve_babd()
    J, V, D, W, B0, Bn, Vn = osolver.construct_jacobian(prob, m, N, rk, t, y, yt_o, zt, p, 0.1)
    A_o, C_o, H_o, b_o = osolver.reduce_jacobian(prob, m, N, W, D, J, V, fa_o, fb_o)
    from oracle import babd as obabd
    dy_o, dp_o = obabd.solve_babd(prob.size_y, prob.size_p, N, A_o, C_o, H_o, b_o, B0, Bn, Vn, rbc_o, M=8)
    assert rel_err(dy, dy_o) < 1e-10
    assert rel_err(dp, dp_o) < 1e-10
    assert S.status() == 0
    S.close()


def test_table_form_batch_equals_straight_line():
    """11 instances of the batched n_x=4 family (per-instance x0 constants):
    the table form's residual norms, r_bc, Dr and segment residuals equal
    the straight-line form's."""
    insts = problems.batch_family(11, nx=4, nu=2, nint_range=(16, 40))
    m = 5
    bvp = solver.BvpDae(insts[0].ocp_text())
    ts = [np.linspace(0.0, it.T, it.nodes) for it in insts]
    cons = np.array([it.constant_values() for it in insts])
    Ntot = sum(len(t) for t in ts)
    out = {}
    for form in ("code", "tables"):
        say("batch of 11, form %s" % form)
        S = solver.CollocationSolver(bvp, m, n_max=Ntot, batch=len(insts), form=form)
        S.set_batch_mesh(ts)
        S.set_batch_constants(cons)
        y, yd, zt, p = seeded_state(bvp, m, Ntot - len(insts) + len(insts), 5)
        S.set_state(y[:Ntot], yd[:(Ntot - len(insts)) * m], zt[:(Ntot - len(insts)) * m],
                    np.tile(p, len(insts)))
        nrm, yt, fa, fb, rbc = S.compute_f_q(0.1)
        Df, Dg, Dr = S.construct_jacobian(0.1)
        res, mx = S.segment_residual(0.1, 1e-6)
        out[form] = (nrm, yt, fa, fb, rbc, Df, Dg, Dr, res, mx)
        S.close()
    for k, (a, b) in enumerate(zip(out["tables"], out["code"])):
        assert rel_err(a, b) < 1e-13, k


@pytest.mark.parametrize("name", ["trolley", "param_osc"])
def test_table_form_equals_straight_line(name):
    text = problems.CONSTRAINED_EXAMPLES[name] if name in problems.CONSTRAINED_EXAMPLES else problems.PARAM_OSC_OCP
    m = 4
    bvp = solver.BvpDae(text)
    N = 41
    t = np.linspace(bvp.t_initial, bvp.t_final, N)
    y, yd, zt, p = seeded_state(bvp, m, N, 11, positive=True)
    y0 = np.random.default_rng(12).uniform(0.2, 1.0, (N, bvp.size_y))
    z0 = np.random.default_rng(13).uniform(0.2, 1.0, (N, bvp.size_z))
    out = {}
    for form in ("code", "tables"):
        say("%s, form %s" % (name, form))
        S = solver.CollocationSolver(bvp, m, n_max=N, form=form)
        S.set_mesh(t)
        S.set_state(y, yd, zt, p)
        vals = list(S.compute_f_q(0.05))
        vals += list(S.construct_jacobian(0.05))
        for rt in (0, 1, 2):
            for sc in (True, False):
                vals += list(S.segment_residual(0.05, 1e-6, residual_type=rt, scale_by_time=sc))
        vals += list(S.segment_residual(0.05, 1e-6, residual2=True))
        S.set_direction_precision(32)
        vals.append(S.construct_jacobian(0.05)[0])
        S.form_initial_input(y0, z0, p)
        vals += list(S.get_state())
        out[form] = vals
        assert S.status() == 0
        S.close()
    n32 = len(out["code"]) - 5
    for k, (a, b) in enumerate(zip(out["tables"], out["code"])):
        tol = 1e-6 if k == n32 else 1e-13
        assert rel_err(a, b) < tol, (k, rel_err(a, b))


def test_table_form_trolley_driver():
    text = problems.CONSTRAINED_EXAMPLES["trolley"]
    runs = {}
    for form in ("code", "tables"):
        bvp = solver.BvpDae(text)
        S = solver.CollocationSolver(bvp, 4, n_max=max(bvp.maximum_nodes, bvp.N) + 4 * bvp.N, form=form)
        say("trolley driver, form %s" % form)
        runs[form] = solver.collocation_solver_parallel(bvp, 4, solver=S)
        S.close()
    a, b = runs["tables"], runs["code"]
    assert a["success"] == 1 and b["success"] == 1
    assert len(a["history"]) == len(b["history"])
    for ha, hb in zip(a["history"], b["history"]):
        assert ha["iterations"] == hb["iterations"]
        assert abs(ha["alpha"] - hb["alpha"]) <= 1e-15 * hb["alpha"]
        assert ha["N"] == hb["N"]
    for k in ("y", "z", "p"):
        assert rel_err(a[k], b[k]) < 1e-9, k
