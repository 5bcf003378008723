"""Batched solves on the GPU: B instances of one problem in one handle.

Every instance of a batch runs the same per-interval arithmetic as a batch
of one, so each instance's stage outputs, Newton direction, iteration
counts and solution must equal those of the same instance solved alone
(bit for bit: reductions are maxima and nothing is summed across
instances).  Instances of the n_x=4 tanh family with per-instance x0, T and
mesh (uneven sizes, N-1 in 96..160, one with a single interval) are also
compared with the CPU oracle.
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


@pytest.fixture(scope="module")
def fam():
    insts = problems.batch_family(5, seed=1, nx=4, nu=2, nint_range=(96, 160))
    bvp = solver.BvpDae(insts[0].ocp_text())
    m = 5
    total = sum(i.nodes for i in insts)
    say("batch of %d: nodes %s, nc=%d" % (len(insts), [i.nodes for i in insts], bvp.size_c))
    SB = solver.CollocationSolver(bvp, m, n_max=total, batch=len(insts))
    S1 = solver.CollocationSolver(bvp, m, n_max=max(i.nodes for i in insts), batch=1)
    yield insts, m, bvp, SB, S1
    SB.close()
    S1.close()


def meshes(insts, seed, single=3):
    """Random meshes of the instances' sizes; instance `single` gets one
    interval (a batch member without any cyclic-reduction level)."""
    rng = np.random.default_rng(seed)
    out = []
    for k, inst in enumerate(insts):
        n = 2 if k == single else inst.nodes
        out.append(np.sort(np.concatenate([[0.0, inst.T], rng.uniform(0.0, inst.T, n - 2)])))
    return out


def test_batch_stages_equal_single(fam):
    insts, m, bvp, SB, S1 = fam
    ts = meshes(insts, 7)
    rng = np.random.default_rng(3)
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    states = []
    for t in ts:
        n = len(t)
        states.append((rng.normal(scale=0.5, size=(n, ny)), rng.normal(scale=0.5, size=((n - 1) * m, ny)),
                       rng.normal(scale=0.5, size=((n - 1) * m, nz)), rng.normal(scale=0.5, size=np_)))
    consts = np.array([inst.constant_values() for inst in insts])
    SB.set_batch_mesh(ts)
    SB.set_batch_constants(consts)
    SB.set_state(np.concatenate([s[0] for s in states]), np.concatenate([s[1] for s in states]),
                 np.concatenate([s[2] for s in states]), np.stack([s[3] for s in states]))
    say("batched stages")
    nrm, yt, fa, fb, rbc = SB.compute_f_q(0.1)
    Df, Dg, Dr = SB.construct_jacobian(0.1)
    A, _, H, b = SB.reduce_jacobian()
    dy, dp = SB.solve_babd()
    dyd, dz = SB.recover_delta_k()
    assert SB.status() == 0
    noff = SB.node_offsets()
    ioff = SB.interval_offsets()
    for k, (t, st) in enumerate(zip(ts, states)):
        say("instance %d alone (N=%d)" % (k, len(t)))
        S1.set_mesh(t)
        S1.set_constants(consts[k])
        S1.set_state(*st)
        n1, yt1, fa1, fb1, rbc1 = S1.compute_f_q(0.1)
        Df1, Dg1, Dr1 = S1.construct_jacobian(0.1)
        A1, _, H1, b1 = S1.reduce_jacobian()
        dy1, dp1 = S1.solve_babd()
        dyd1, dz1 = S1.recover_delta_k()
        i0, i1 = ioff[k], ioff[k + 1]
        s0, s1 = i0 * m, i1 * m
        assert nrm[k] == n1
        assert np.array_equal(yt[s0:s1], yt1) and np.array_equal(fa[i0:i1], fa1) and np.array_equal(fb[i0:i1], fb1)
        assert np.array_equal(rbc[k], rbc1)
        assert np.array_equal(Df[s0:s1], Df1) and np.array_equal(Dg[s0:s1], Dg1) and np.array_equal(Dr[k], Dr1)
        assert np.array_equal(A[i0:i1], A1) and np.array_equal(H[i0:i1], H1) and np.array_equal(b[i0:i1], b1)
        assert np.array_equal(dy[noff[k]:noff[k + 1]], dy1) and np.array_equal(dp[k], dp1)
        assert np.array_equal(dyd[s0:s1], dyd1) and np.array_equal(dz[s0:s1], dz1)


def solve_inputs(insts, bvp):
    ts = [np.linspace(0.0, inst.T, inst.nodes) for inst in insts]
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    Y0s = [np.ones((len(t), ny)) for t in ts]
    Z0s = [np.ones((len(t), nz)) for t in ts]
    P0s = [np.ones(np_) for _ in ts]
    consts = np.array([inst.constant_values() for inst in insts])
    hom = [np.array([inst.constant_values(s=0.0) for inst in insts])]
    return ts, Y0s, Z0s, P0s, consts, hom


def test_batch_solve_matches_single_and_oracle(fam):
    """Plain all-ones guess + homotopy, as the reference algorithm runs it:
    on these meshes some instances converge and some stall (the oracle's
    line search stalls on instance 0 as well); either way every instance of
    the batch must do exactly what it does alone, and what the oracle does."""
    insts, m, bvp, SB, S1 = fam
    ts, Y0s, Z0s, P0s, consts, hom = solve_inputs(insts, bvp)
    tol, max_iter = bvp.tolerance, bvp.maximum_newton_iterations
    say("batched solve")
    r = solver.solve_batch(SB, ts, Y0s, Z0s, P0s, consts, hom, tol=tol, max_iter=max_iter)
    sols = solver.batch_solutions(SB)
    say("iterations %s norms %s max residual %s" % (r["iterations"].tolist(), r["norm"].tolist(),
                                                   r["max_residual"].tolist()))
    assert (r["norm"] < tol).sum() >= 3
    for k, inst in enumerate(insts):
        say("instance %d alone" % k)
        r1 = solver.solve_batch(S1, [ts[k]], [Y0s[k]], [Z0s[k]], [P0s[k]], consts[k:k + 1], [hom[0][k:k + 1]],
                                tol=tol, max_iter=max_iter)
        s1 = solver.batch_solutions(S1)[0]
        assert np.array_equal(r1["iterations"][:, 0], r["iterations"][:, k])
        assert r1["norm"][0] == r["norm"][k] and r1["max_residual"][0] == r["max_residual"][k]
        for key in ("y", "z", "p"):
            assert np.array_equal(s1[key], sols[k][key]), key
    rk = otables.lobatto(m)
    for k in (1, 2):
        inst = insts[k]
        say("oracle instance %d" % k)
        prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)
        hp = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=0.0, q=inst.q, c_u=inst.c_u)
        N = len(ts[k])
        st = osolver.form_initial_input(prob, m, N, Y0s[k], Z0s[k], P0s[k], rk)
        y, yd, zt = st[:3]
        p = P0s[k].copy()
        its = []
        for pr in (hp, prob):
            y, yd, zt, p, it, _ = osolver.newton_solve(pr, m, rk, ts[k], y, yd, zt, p, 0.1, tol, max_iter)
            its.append(it)
        assert its == r["iterations"][:, k].tolist()
        assert rel_err(sols[k]["y"], y) < 1e-9 and rel_err(sols[k]["p"], p) < 1e-9
        assert rel_err(sols[k]["z"], osolver.recover_solution(m, N, zt)) < 1e-9


def test_chunked_batch_driver(fam):
    insts, m, bvp, SB, S1 = fam
    ts, Y0s, Z0s, P0s, consts, hom = solve_inputs(insts, bvp)
    args = (ts, Y0s, Z0s, P0s, consts, hom)
    say("whole batch, then in chunks of at most 2 instances")
    st_all, sol_all = solver.collocation_solver_batch(bvp, m, *args, solver=SB)
    st_ch, sol_ch = solver.collocation_solver_batch(bvp, m, *args, solver=SB, max_batch=2)
    assert np.array_equal(st_all["iterations"], st_ch["iterations"])
    assert np.array_equal(st_all["converged"], st_ch["converged"])
    for a, b in zip(sol_all, sol_ch):
        for key in ("y", "z", "p"):
            assert np.array_equal(a[key], b[key])
