"""Launch shapes and the autotuner (autotune.py, cs_set_launch).

* the series condensation as a persistent grid taking intervals from a
  work queue gives the results of one workgroup per interval, bit for bit:
  the condensed rows, the Newton direction, and whole batched solves, in
  the host-driven and the device-resident (graph) Newton loop, float64 and
  FP32 directions (n_x = 4 and n_x = 20 family batches, uneven meshes);
* every residual workgroup size gives the same norms and residual arrays;
* the autotuner measures the candidates, applies the fastest, keeps only
  pairing rules whose direction agrees with the Householder-only one, and
  a second call with the same shape is served from its cache.
"""
import numpy as np
import pytest

from colloc_amd import autotune, native, problems, solver

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def batch(nx, nu, count, nint_range, seed=1):
    insts = problems.batch_family(count, seed=seed, nx=nx, nu=nu, nint_range=nint_range)
    ts = [np.linspace(0.0, it.T, it.nodes) for it in insts]
    bvp = solver.BvpDae(insts[0].ocp_text())
    return insts, ts, bvp


def solve(S, insts, ts):
    cons = np.array([i.constant_values() for i in insts])
    hom = [np.array([i.constant_values(s=0.0) for i in insts])]
    Nt = sum(len(t) for t in ts)
    S.set_batch_mesh(ts)
    S.form_initial_input(np.ones((Nt, S.ny)), np.ones((Nt, S.nz)), np.ones((len(ts), S.np)))
    r = solver.newton_stages(S, cons, hom, 0.1, S.bvp.tolerance, S.bvp.maximum_newton_iterations)
    return r, S.get_state()


def test_persistent_condensation_equals_one_workgroup_per_interval():
    insts, ts, bvp = batch(4, 2, 6, (40, 120))
    Nt = sum(len(t) for t in ts)
    S = solver.CollocationSolver(bvp, 5, n_max=Nt, batch=len(insts))
    cus = S.get_launch(native.CS_LAUNCH_COMPUTE_UNITS)
    say("compute units %d" % cus)
    assert cus > 0
    S.set_batch_mesh(ts)
    S.set_batch_constants(np.array([i.constant_values() for i in insts]))
    rng = np.random.default_rng(3)
    nst = (Nt - len(ts)) * 5
    S.set_state(rng.normal(scale=0.5, size=(Nt, S.ny)), rng.normal(scale=0.5, size=(nst, S.ny)),
                rng.normal(scale=0.5, size=(nst, S.nz)), rng.normal(scale=0.5, size=(len(ts), S.np)))
    out = {}
    for bits in (64, 32):
        S.set_direction_precision(bits)
        for grid in (0, 7, 2 * cus):
            S.set_launch(native.CS_LAUNCH_CONDENSE_GRID, grid)
            say("direction fp%d, condensation grid %d" % (bits, grid))
            S.compute_f_q(0.1, outputs=False)
            S.construct_jacobian(0.1, outputs=False)
            A, C, H, b = S.reduce_jacobian()
            S.solve_babd()
            S.recover_delta_k()
            out[(bits, grid)] = (A, H, b) + tuple(S.get_direction())
        for grid in (7, 2 * cus):
            for x, y in zip(out[(bits, grid)], out[(bits, 0)]):
                np.testing.assert_array_equal(x, y)
    assert S.status() == 0
    S.close()


@pytest.mark.parametrize("graph", [False, True])
def test_persistent_condensation_whole_solves(graph):
    """n_x = 20 members, FP32 direction (the matrix-core kernel) and the
    bench's pairing: identical solves with either launch, host loop or graphs."""
    insts, ts, bvp = batch(20, 10, 4, (64, 200))
    Nt = sum(len(t) for t in ts)
    S = solver.CollocationSolver(bvp, 5, n_max=Nt, batch=len(insts))
    S.set_direction_precision(32)
    S.set_babd_substitution(16)
    S.set_newton_graph(graph)
    res = {}
    for grid in (0, 3 * S.get_launch(native.CS_LAUNCH_COMPUTE_UNITS)):
        S.set_launch(native.CS_LAUNCH_CONDENSE_GRID, grid)
        say("n_x=20 solves, graph %s, condensation grid %d" % (graph, grid))
        res[grid] = solve(S, insts, ts)
    a, b = list(res.values())
    for k in ("iterations", "norm", "converged"):
        np.testing.assert_array_equal(a[0][k], b[0][k], err_msg=k)
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)
    assert a[0]["converged"].sum() >= 1
    S.close()


def test_residual_workgroup_sizes_agree():
    insts, ts, bvp = batch(4, 2, 5, (30, 90))
    Nt = sum(len(t) for t in ts)
    S = solver.CollocationSolver(bvp, 5, n_max=Nt, batch=len(insts))
    S.set_batch_mesh(ts)
    S.set_batch_constants(np.array([i.constant_values() for i in insts]))
    rng = np.random.default_rng(5)
    nst = (Nt - len(ts)) * 5
    S.set_state(rng.normal(size=(Nt, S.ny)), rng.normal(size=(nst, S.ny)), rng.normal(size=(nst, S.nz)),
                rng.normal(size=(len(ts), S.np)))
    out = {}
    for blk in (64, 128, 256):
        S.set_launch(native.CS_LAUNCH_RESIDUAL_BLOCK, blk)
        out[blk] = S.compute_f_q(0.1)
    for blk in (64, 128):
        for x, y in zip(out[blk], out[256]):
            np.testing.assert_array_equal(x, y)
    with pytest.raises(native.ColloError):
        S.set_launch(native.CS_LAUNCH_RESIDUAL_BLOCK, 96)
    S.close()


def test_autotuner_picks_measured_choices_and_caches(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE_DIR", str(tmp_path))
    insts, ts, bvp = batch(20, 10, 6, (200, 400))
    Nt = sum(len(t) for t in ts)
    S = solver.CollocationSolver(bvp, 5, n_max=Nt, batch=len(insts))
    S.set_direction_precision(32)
    r, _ = solve(S, insts, ts)
    say("tuning on the solved batch")
    rec = autotune.tune(S, log=say)
    ch = rec["choices"]
    assert not rec["cached"]
    assert ch["condense_grid"] in [int(k) for k in rec["timings_ms"]["condense_grid"]]
    tg = rec["timings_ms"]["condense_grid"]
    assert tg[str(ch["condense_grid"])] == min(tg.values())
    assert ch["residual_block"] in (64, 128, 256)
    errs = rec["timings_ms"]["babd_direction_rel_diff"]
    assert errs["0"] == 0.0 and errs[str(ch["babd_spans"])] <= 1e-7
    assert S.get_launch(native.CS_LAUNCH_CONDENSE_GRID) == ch["condense_grid"]
    assert S.get_launch(native.CS_LAUNCH_RESIDUAL_BLOCK) == ch["residual_block"]
    S.set_launch(native.CS_LAUNCH_CONDENSE_GRID, 0)
    again = autotune.tune(S)
    assert again["cached"] and again["choices"] == ch
    assert S.get_launch(native.CS_LAUNCH_CONDENSE_GRID) == ch["condense_grid"]
    # the tuned settings solve the batch as before (same converged set)
    r2, _ = solve(S, insts, ts)
    np.testing.assert_array_equal(r["converged"], r2["converged"])
    S.close()
