"""Handles driven concurrently from several host threads (one HIP stream
each), as the batched bench runs its chunks.

Each handle owns its buffers, module and stream, and the library releases
the interpreter inside its calls, so chunks solved at the same time on
different handles must give exactly what the same chunks give when solved
one after the other: Newton iterations, norms, converged flags and the
solution arrays, bit for bit.  n_x = 20 family members at the bench's
settings (FP32 direction, pairing 1/16), plain guess + continuation, three
handles over six chunks; and the learned warm start (whose network buffers
are shared by the threads) on two handles.
"""
import os
import threading

import numpy as np
import pytest

from colloc_amd import learned, problems, solver   # learned imports torch before the library loads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def chunk_solve(S, insts):
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    cons = np.array([i.constant_values() for i in insts])
    hom = [np.array([i.constant_values(s=0.0) for i in insts])]
    r = solver.solve_batch(S, ts, [np.ones((len(t), S.ny)) for t in ts], [np.ones((len(t), S.nz)) for t in ts],
                           [np.ones(S.np) for _ in ts], cons, hom, tol=1e-8, max_iter=60)
    return r, S.get_state()


def run_threads(handles, groups, fn):
    out = {}
    errors = []

    def body(h):
        try:
            for k in groups[h]:
                out[k] = fn(handles[h], k)
        except BaseException as e:
            errors.append(e)
    threads = [threading.Thread(target=body, args=(h,)) for h in range(len(handles))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    return out


def same(a, b):
    ra, sa = a
    rb, sb = b
    for k in ("iterations", "norm", "converged"):
        np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)
    for x, y in zip(sa, sb):
        np.testing.assert_array_equal(x, y)


def test_concurrent_handles_equal_sequential():
    fam = problems.batch_family(1024)
    chunks = [[fam[i] for i in range(3 * k, 3 * k + 3)] for k in range(6)]
    bvp = solver.BvpDae(fam[0].ocp_text())
    n_max = max(sum(i.nodes for i in ch) for ch in chunks)
    handles = [solver.CollocationSolver(bvp, 5, n_max=n_max, batch=3) for _ in range(3)]
    for S in handles:
        S.set_direction_precision(32)
        S.set_babd_substitution(16)
    say("six chunks of three, one after the other")
    seq = {k: chunk_solve(handles[k % 3], chunks[k]) for k in range(6)}
    say("the same chunks on three threads")
    par = run_threads(handles, [[0, 3], [1, 4], [2, 5]], lambda S, k: chunk_solve(S, chunks[k]))
    for k in range(6):
        same(par[k], seq[k])
    assert sum(int(seq[k][0]["converged"].sum()) for k in range(6)) >= 1
    for S in handles:
        assert S.status() == 0
        S.close()


def test_concurrent_learned_warm_start():
    if not os.path.exists(learned.FAMILY_MODEL):
        pytest.skip("no trained model")
    ws = learned.WarmStart.load(learned.FAMILY_MODEL)
    fam = problems.batch_family(1024)
    chunks = [list(range(4 * k, 4 * k + 4)) for k in range(4)]
    bvp = solver.BvpDae(fam[0].ocp_text())
    n_max = max(sum(fam[i].nodes for i in ch) for ch in chunks)
    handles = [solver.CollocationSolver(bvp, 5, n_max=n_max, batch=4) for _ in range(2)]
    for S in handles:
        S.set_direction_precision(32)
    feats = learned.features([i.x0 for i in fam[:16]], [i.T for i in fam[:16]])

    def solve(S, k):
        ch = chunks[k]
        r = learned.warm_solve(ws, S, [np.linspace(0.0, fam[i].T, fam[i].nodes) for i in ch], feats[ch],
                               np.array([fam[i].constant_values() for i in ch]),
                               np.array([fam[i].constant_values(s=0.0) for i in ch]), max_iter=60)
        return r, S.get_state()
    say("learned warm start, four chunks in turn")
    seq = {k: solve(handles[k % 2], k) for k in range(4)}
    say("the same on two threads")
    par = run_threads(handles, [[0, 2], [1, 3]], solve)
    for k in range(4):
        same(par[k], seq[k])
    for S in handles:
        assert S.status() == 0
        S.close()
