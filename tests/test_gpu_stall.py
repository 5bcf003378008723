"""The early-stopping Newton variants on the GPU: cs_newton_progress and
cs_newton_stall.

Each instance of a batch runs its own arithmetic (masks only skip work), so
an instance that one of these tests stops has, up to its stop, exactly the
norm history of the reference loop (cs_newton) from the same state.  The
stop itself is checked by replaying the rule on that history:

  - cs_newton_stall(start, window, ratio): the first iteration k >= start,
    k >= window, with ||f_q|| >= tol and not ||f_q||_k * ratio <=
    ||f_q||_{k - window};
  - cs_newton_progress(patience, ratio) is cs_newton_stall(patience, 1, ratio).

The batch: n_x = 20 family members at Nint 48..80, all-ones guess straight
at the final constants (no continuation), so that some instances converge
and others sit on a plateau.
"""
import numpy as np
import pytest

from colloc_amd import native, problems, solver

pytestmark = pytest.mark.gpu

MAX_ITER = 60


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def fam():
    insts = problems.batch_family(12, seed=3, nint_range=(48, 80))
    bvp = solver.BvpDae(insts[0].ocp_text())
    total = sum(i.nodes for i in insts)
    S = solver.CollocationSolver(bvp, 5, n_max=total, batch=len(insts))
    say("stall batch: nodes %s" % [i.nodes for i in insts])
    yield insts, bvp, S
    S.close()


def start_state(insts, bvp, S):
    Nt = sum(i.nodes for i in insts)
    S.set_batch_mesh([np.linspace(0.0, i.T, i.nodes) for i in insts])
    S.form_initial_input(np.ones((Nt, bvp.size_y)), np.ones((Nt, bvp.size_z)), np.ones((len(insts), bvp.size_p)))
    S.set_batch_constants(np.array([i.constant_values() for i in insts]))


def newton_stall(S, tol, start, window, ratio):
    """cs_newton_stall through the C ABI (any window, 1 included)."""
    it = np.zeros(S.B, dtype=np.int32)
    nrm = np.zeros(S.B)
    hist = np.zeros((S.B, MAX_ITER))
    native.check(native.lib().cs_newton_stall(S.h, 0.1, tol, MAX_ITER, start, window, ratio, native.iptr(it),
                                              native.ptr(nrm), native.ptr(hist), MAX_ITER))
    return it, nrm, hist


def executed(hist, it, nrm, tol):
    """Iterations the reference loop evaluated: all up to its stop, but for
    an instance whose line search accepted no step (its history is then
    continued at that norm) only up to that iteration."""
    if nrm < tol or it < MAX_ITER - 1:
        return it
    j = len(hist) - 1
    while j > 0 and hist[j - 1] == hist[-1]:
        j -= 1
    return j


def replay(hist, last, tol, start, window, ratio):
    for k in range(max(start, window), last + 1):
        if hist[k] < tol:
            return None
        if not hist[k] * ratio <= hist[k - window]:
            return k
    return None


def test_stall_and_progress_tests_replay(fam):
    insts, bvp, S = fam
    tol = bvp.tolerance
    start_state(insts, bvp, S)
    it0, n0, h0 = S.newton(0.1, tol, MAX_ITER)
    say("reference loop: iterations %s, converged %d" % (list(it0), int((n0 < tol).sum())))
    assert (n0 < tol).any() and (n0 >= tol).any()
    stopped_total = 0
    for start, window, ratio in ((6, 4, 2.0), (10, 8, 4.0)):
        start_state(insts, bvp, S)
        it, nrm, h = S.newton(0.1, tol, MAX_ITER, start, ratio, window=window)
        for b in range(len(insts)):
            last = executed(h0[b], it0[b], n0[b], tol)
            k = replay(h0[b], last, tol, start, window, ratio)
            if k is None:
                assert it[b] == it0[b] and nrm[b] == n0[b], (b, start, window)
            else:
                stopped_total += 1
                assert it[b] == k and nrm[b] == h0[b][k] and nrm[b] >= tol, (b, k, it[b])
                np.testing.assert_array_equal(h[b][:k + 1], h0[b][:k + 1])
        say("stall test (%d, %d, %g): iterations %s" % (start, window, ratio, list(it)))
    assert stopped_total > 0
    # window 1 is the per-iteration progress test
    start_state(insts, bvp, S)
    ip, npg, hp = S.newton(0.1, tol, MAX_ITER, 5, 1.5)
    start_state(insts, bvp, S)
    iw, nw, hw = newton_stall(S, tol, 5, 1, 1.5)
    np.testing.assert_array_equal(ip, iw)
    np.testing.assert_array_equal(npg, nw)
    for b in range(len(insts)):
        last = executed(h0[b], it0[b], n0[b], tol)
        k = replay(h0[b], last, tol, 5, 1, 1.5)
        assert ip[b] == (it0[b] if k is None else k), (b, k, ip[b])


def test_stall_arguments(fam):
    insts, bvp, S = fam
    start_state(insts, bvp, S)
    for start, window, ratio in ((0, 4, 2.0), (5, 0, 2.0), (5, MAX_ITER + 1, 2.0), (5, 4, 0.5)):
        with pytest.raises(native.ColloError, match="window"):
            newton_stall(S, bvp.tolerance, start, window, ratio)
