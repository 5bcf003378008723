"""cs_compute_f_q_noalloc through the C ABI against cs_compute_f_q.

The allocation-free variant evaluates the same residual kernel with the
per-instance max-norm slots in the library's fixed device array instead of
the handle's workspace.  The norms must be bit-identical to
cs_compute_f_q's (the same kernel; a max over the same values), on two calls
in a row, the second with a smaller batch whose norms are smaller than the
first's, so that a slot the second call failed to clear would show the first
call's value.  A batch larger than the array holds is rejected before any
work is queued.
"""
import numpy as np
import pytest

from colloc_amd import native, problems, solver

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def norms(S, noalloc):
    out = np.full(S.B, -1.0)
    if noalloc:
        native.check(native.lib().cs_compute_f_q_noalloc(S.h, 0.1, native.ptr(out)))
    else:
        native.check(native.lib().cs_compute_f_q(S.h, 0.1, native.ptr(out), None, None, None, None))
    return out


def test_noalloc_residual_norms_equal_compute_f_q():
    insts = problems.batch_family(6, nx=4, nu=2, nint_range=(16, 64))
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(i.nodes for i in insts), batch=len(insts))
    rng = np.random.default_rng(11)
    try:
        for B, scale in ((6, 10.0), (3, 0.01)):
            sel = insts[:B]
            S.set_batch_mesh([np.linspace(0.0, i.T, i.nodes) for i in sel])
            S.set_batch_constants(np.array([i.constant_values() for i in sel]))
            Nt = sum(i.nodes for i in sel)
            nst = (Nt - B) * S.m
            S.set_state(rng.normal(scale=scale, size=(Nt, S.ny)), rng.normal(scale=scale, size=(nst, S.ny)),
                        rng.normal(scale=scale, size=(nst, S.nz)), rng.normal(scale=scale, size=(B, S.np)))
            say("batch of %d, state scale %g" % (B, scale))
            got = norms(S, True)
            ref = norms(S, False)
            again = norms(S, True)
            say("  norms %s" % np.array2string(got, precision=4))
            assert np.array_equal(got, ref) and np.array_equal(again, ref)
            assert np.all(ref > 0)
    finally:
        S.close()


def test_noalloc_rejects_batches_beyond_its_slots():
    inst, m = problems.config(1, nx=4, nu=2, nint=32)
    bvp = solver.BvpDae(inst.ocp_text())
    B = 8193
    S = solver.CollocationSolver(bvp, m, n_max=2 * B, batch=B)
    try:
        S.set_batch_mesh([np.array([0.0, 1.0])] * B)
        out = np.zeros(B)
        rc = native.lib().cs_compute_f_q_noalloc(S.h, 0.1, native.ptr(out))
        assert rc == native.CS_E_SHAPE
        assert b"8192" in native.lib().cs_last_error()
    finally:
        S.close()
