"""Pins of the oracle's Newton driver: line search, alpha continuation and
remesh (collocation_solver_parallel_2d_shared.py:105-213), on hand-built
cases whose outcome is fixed independently of the oracle's code.

* The step rule is checked on hand-made merit sequences: the strict
  decrease (an equal norm is refused), at most 20 trials (2^-19 is the last
  step tried, 2^-20 never is), the first decreasing step is taken (not the
  best one).
* The whole Newton loop runs on a scalar problem, y' = z, 0 = z,
  atan(y(0)) = 0, whose collocation equations are linear except the
  boundary condition.  From y = c the Newton direction is the scalar
  Newton step -atan(c) (1 + c^2) at every node, so the iterates, the number
  of halvings of each step and the norm history follow from a scalar
  closed form; the oracle gets there through its full pipeline
  (condensation, BABD solve, recovery).
* With the sign of dr/dy flipped the direction is an ascent direction:
  every trial fails, the state never changes, and the boundary function is
  evaluated exactly 1 + 20 times per iteration.
* The alpha continuation of a one-inequality OCP visits alpha = 0.1 * 0.9^k
  and stops at the first alpha <= 1e-6 (111 accepted meshes), and its
  final solution is the constrained optimum (a closed form: the bound is
  active on [0, tau], the unconstrained arc sinh-shaped after it).
* A coarse mesh of a stiff LQ problem is remeshed until the segment
  residual is <= 1, with alpha unchanged across every remesh.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

from colloc_amd import emitter, ocp as ocpmod
from oracle import families
from oracle import solver as osolver
from symbolic_problem import SymbolicProblem


# ----------------------------------------------------------- step rule
def run_rule(merit, norm0):
    steps = []

    def trial(a):
        steps.append(a)
        return merit(a)
    return osolver.line_search(trial, norm0), steps


def test_line_search_refuses_an_equal_norm():
    # the full step lands on a point of the same norm: not a decrease
    a, steps = run_rule(lambda a: 1.0 if a == 1.0 else 0.25, 1.0)
    assert a == 0.5 and steps == [1.0, 0.5]


def test_line_search_takes_the_first_decrease():
    # 1/2 would be better, but 1 already decreases
    a, steps = run_rule(lambda a: {1.0: 0.9, 0.5: 0.1}.get(a, 2.0), 1.0)
    assert a == 1.0 and steps == [1.0]


def test_line_search_tries_twenty_steps():
    # only steps <= 2^-19 decrease: the 20th trial succeeds
    a, steps = run_rule(lambda a: 0.5 if a <= 2.0 ** -19 else 2.0, 1.0)
    assert a == 2.0 ** -19
    assert steps == [2.0 ** -k for k in range(20)]
    # only steps <= 2^-20 decrease: no step is taken, after exactly 20 trials
    a, steps = run_rule(lambda a: 0.5 if a <= 2.0 ** -20 else 2.0, 1.0)
    assert a is None
    assert steps == [2.0 ** -k for k in range(20)]


# ------------------------------------------------ scalar closed-form loop
class AtanBVP:
    """y' = z, 0 = z, r = atan(y(0)); sign = -1 flips dr/dy."""
    size_y, size_z, size_p = 1, 1, 0

    def __init__(self, sign=1.0):
        self.sign = sign
        self.r_calls = 0

    def f(self, Y, Z, P):
        return Z[:, :1].copy()

    def g(self, Y, Z, P, alpha=0.0):
        return Z[:, :1].copy()

    def Df(self, Y, Z, P):
        out = np.zeros((Y.shape[0], 1, 2))
        out[:, 0, 1] = 1.0
        return out

    def Dg(self, Y, Z, P, alpha=0.0):
        return self.Df(Y, Z, P)

    def r(self, y0, y1, p):
        self.r_calls += 1
        return np.array([np.arctan(y0[0])])

    def Dr(self, y0, y1, p):
        return np.array([[self.sign / (1.0 + y0[0] ** 2), 0.0]])


def scalar_newton(c, tol, max_iter):
    """The reference's loop on the scalar equation atan(y) = 0: returns the
    norm history and the halvings of each accepted step."""
    y = c
    norms, halvings = [], []
    for _ in range(max_iter):
        n = abs(np.arctan(y))
        norms.append(n)
        if n < tol:
            break
        d = -np.arctan(y) * (1.0 + y * y)
        for k in range(20):
            if abs(np.arctan(y + 2.0 ** -k * d)) < n:
                y = y + 2.0 ** -k * d
                halvings.append(k)
                break
    return norms, halvings


def start(prob, c, nint=16, m=3):
    t = np.linspace(0.0, 1.0, nint + 1)
    rk = osolver.lobatto(m)
    y0 = np.full((nint + 1, 1), c)
    z0 = np.zeros((nint + 1, 1))
    y, yd, zt = osolver.form_initial_input(prob, m, nint + 1, y0, z0, np.zeros(0), rk)
    return t, rk, y, yd, zt


def test_halvings_follow_the_scalar_newton_map(monkeypatch):
    c, tol, m = 10.0, 1e-8, 3
    ref_norms, ref_halv = scalar_newton(c, tol, 50)
    # from y = 10 the full step and two halvings overshoot: the fourth trial (2^-3) decreases
    assert ref_halv[:2] == [3, 3] and len(ref_norms) > 3
    taken = []
    real = osolver.line_search

    def recording(trial_norm, norm0, max_trials=osolver.MAX_LINESEARCH):
        a = real(trial_norm, norm0, max_trials)
        taken.append(a)
        return a
    monkeypatch.setattr(osolver, "line_search", recording)
    prob = AtanBVP()
    t, rk, y, yd, zt = start(prob, c, m=m)
    y, yd, zt, p, it, norms = osolver.newton_solve(prob, m, rk, t, y, yd, zt, np.zeros(0), 0.1, tol, 50)
    assert it == len(ref_norms) - 1
    assert taken == [2.0 ** -k for k in ref_halv]
    np.testing.assert_allclose(norms, ref_norms, rtol=1e-9, atol=1e-13)
    assert np.max(np.abs(y)) < 1e-8 and np.max(np.abs(zt)) < 1e-12


def test_failed_line_search_leaves_the_state_unchanged():
    c, m, max_iter = 2.0, 3, 4
    prob = AtanBVP(sign=-1.0)
    t, rk, y, yd, zt = start(prob, c, m=m)
    y_in, yd_in, zt_in = y.copy(), yd.copy(), zt.copy()
    prob.r_calls = 0
    y2, yd2, zt2, p2, it, norms = osolver.newton_solve(prob, m, rk, t, y, yd, zt, np.zeros(0), 0.1, 1e-8, max_iter)
    # every iteration: one residual, then 20 failed trials
    assert prob.r_calls == max_iter * (1 + 20)
    assert it == max_iter - 1
    assert norms == [abs(np.arctan(c))] * max_iter
    assert np.array_equal(y2, y_in) and np.array_equal(yd2, yd_in) and np.array_equal(zt2, zt_in)
    # the driver reports the stall as a failed solve
    res = osolver.collocation_solver(AtanBVP(sign=-1.0), m, t, y_in, np.zeros((t.size, 1)), np.zeros(0), tol=1e-8,
                                     max_iter=max_iter)
    assert res["success"] == 0 and len(res["history"]) == 1


# ------------------------------------------------- alpha continuation
BOUNDED_OCP = """
StateVariables = [x];
ControlVariables = [u];
CostFunctional = x*x + u*u;
DifferentialEquations = [u];
InequalityConstraints = [-1.2 - u];
InitialConstraints = [x - 1];
TerminalConstraints = [x];
InitialTime = 0;
FinalTime = 1;
Nodes = 65;
Tolerance = 1e-8;
"""


def constrained_optimum(t):
    """x, u of min int x^2 + u^2, x' = u, u >= -1.2, x(0) = 1, x(1) = 0.
    Unconstrained, x'' = x gives u(0) = -coth(1) < -1.2, so the bound is
    active on [0, tau]: x = 1 - 1.2 t there, and x = k sinh(1 - t) after,
    with x and u = x' continuous at tau (u = max(-lambda/2, -1.2) and lambda
    is continuous): k = 1.2 / cosh(1 - tau), 1 - 1.2 tau = 1.2 tanh(1 - tau)."""
    tau = brentq(lambda s: 1.0 - 1.2 * s - 1.2 * np.tanh(1.0 - s), 0.0, 0.5)
    k = 1.2 / np.cosh(1.0 - tau)
    x = np.where(t <= tau, 1.0 - 1.2 * t, k * np.sinh(1.0 - t))
    u = np.where(t <= tau, -1.2, -k * np.cosh(1.0 - t))
    return x, u, tau


@pytest.fixture(scope="module")
def bounded_solve():
    prob = SymbolicProblem(BOUNDED_OCP)
    assert (prob.sym.size_inequality, prob.size_y, prob.size_z) == (1, 2, 3)
    N = 65
    t = np.linspace(0.0, 1.0, N)
    res = osolver.collocation_solver(prob, 4, t, np.ones((N, prob.size_y)), np.ones((N, prob.size_z)),
                                     np.ones(prob.size_p), tol=1e-8, inequalities=True)
    return res


def test_alpha_continuation_schedule(bounded_solve):
    res = bounded_solve
    assert res["success"] == 1
    hist = res["history"]
    accepted = [h for h in hist if h["max_residual"] <= 1.0]
    # alpha <- 0.9 alpha after every accepted mesh, from 0.1 until alpha <= 1e-6
    want = [0.1]
    while want[-1] > 1e-6:
        want.append(want[-1] * 0.9)
    assert len(want) == 111
    assert [h["alpha"] for h in accepted] == want
    assert res["alpha"] == want[-1] <= 1e-6 < want[-2]
    # a remeshed entry keeps its alpha for the next solve
    for a, b in zip(hist, hist[1:]):
        if a["max_residual"] > 1.0:
            assert b["alpha"] == a["alpha"]


def test_alpha_continuation_reaches_the_constrained_optimum(bounded_solve):
    res = bounded_solve
    x_ref, u_ref, tau = constrained_optimum(res["t"])
    assert 0.0 < tau < 0.2
    print("tau %.6f, N %d, max |x - x*| %.2e, max |u - u*| %.2e" % (
        tau, res["t"].size, np.max(np.abs(res["y"][:, 0] - x_ref)), np.max(np.abs(res["z"][:, 0] - u_ref))))
    assert np.max(np.abs(res["y"][:, 0] - x_ref)) < 1e-4
    assert np.max(np.abs(res["z"][:, 0] - u_ref)) < 1e-3
    assert np.min(res["z"][:, 0]) > -1.2 - 1e-4


def test_without_inequalities_one_alpha():
    inst_A, inst_B = np.array([[0.0]]), np.array([[1.0]])
    prob = families.TanhFamily(inst_A, inst_B, np.array([1.0]), C=None, q=1.0, c_u=1.0)
    N = 33
    t = np.linspace(0.0, 1.0, N)
    res = osolver.collocation_solver(prob, 4, t, np.ones((N, 2)), np.ones((N, 1)), np.ones(2), tol=1e-8)
    assert res["success"] == 1 and [h["alpha"] for h in res["history"]] == [0.1]


# -------------------------------------------------------------- remesh
def test_remesh_until_accepted(monkeypatch):
    prob = families.TanhFamily(np.array([[0.0]]), np.array([[1.0]]), np.array([1.0]), C=None, q=1.0, c_u=1e-4)
    calls = []
    real = osolver.remesh

    def recording(ny, nz, N, t, y, z, residual):
        out = real(ny, nz, N, t, y, z, residual)
        calls.append((N, float(np.max(residual)), out[0]))
        return out
    monkeypatch.setattr(osolver, "remesh", recording)
    N = 9
    t = np.linspace(0.0, 1.0, N)
    res = osolver.collocation_solver(prob, 4, t, np.ones((N, 2)), np.ones((N, 1)), np.ones(2), tol=1e-8)
    hist = res["history"]
    assert res["success"] == 1
    assert len(calls) >= 1 and len(hist) == len(calls) + 1
    for k, (n_in, mx, n_out) in enumerate(calls):
        assert hist[k]["N"] == n_in and mx == hist[k]["max_residual"] > 1.0
        assert hist[k + 1]["N"] == n_out   # remesh inserts and deletes nodes
    assert hist[-1]["max_residual"] <= 1.0
    assert all(h["alpha"] == 0.1 for h in hist)
    # the accepted solution is the LQ optimum: x'' = x / c_u, x(0) = 1, x(1) = 0
    w, t = 100.0, res["t"]
    x = (np.exp(-w * t) - np.exp(-w * (2.0 - t))) / (1.0 - np.exp(-2.0 * w))
    assert np.max(np.abs(res["y"][:, 0] - x)) < 1e-4
