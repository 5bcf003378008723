"""The device-resident Newton loop (cs_set_newton_graph): the per-instance
decisions of solver:109-174 made by kernels, the launch list built on the
device, each iteration a replay of captured HIP graphs.

Its arithmetic is the host-driven loop's (the same kernels on the same
intervals; a trial residual runs over the iterating instances' intervals
masked to the searching ones, which changes no instance's value), so every
result must be identical, bit for bit: iteration counts, norms, norm
histories, solutions, for
* a batch of n_x = 4 family members with uneven meshes (one of a single
  interval), the plain guess and the continuation, float64 direction;
* the stall and progress tests (cs_newton_stall / cs_newton_progress) on
  n_x = 20 members from the plain guess, and the norm histories of
  cs_newton;
* the FP32 direction with the bench's BABD pairing on n_x = 20 members.
The kernel timings recorded inside the graphs must add up like the
host-driven loop's (the same launches, timed at every replay).
"""
import numpy as np
import pytest

from colloc_amd import native, problems, solver

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def both_ways(S, run):
    out = {}
    for graph in (False, True):
        S.set_newton_graph(graph)
        out[graph] = run()
    S.set_newton_graph(False)
    return out[False], out[True]


def solve(S, insts, ts, hom_scales=(0.0,), patience=0, min_ratio=0.0, stall=None):
    bvp = S.bvp
    Y = [np.ones((len(t), S.ny)) for t in ts]
    Z = [np.ones((len(t), S.nz)) for t in ts]
    P = [np.ones(S.np) for _ in ts]
    cons = np.array([i.constant_values() for i in insts])
    hom = [np.array([i.constant_values(s=s) for i in insts]) for s in hom_scales]
    S.set_batch_mesh(ts)
    S.form_initial_input(np.concatenate(Y), np.concatenate(Z), np.stack(P))
    r = solver.newton_stages(S, cons, hom, 0.1, bvp.tolerance, bvp.maximum_newton_iterations,
                             patience=patience, min_ratio=min_ratio, stall=stall)
    y, yd, zt, p = S.get_state()
    return r, (y, yd, zt, p)


def assert_same(a, b):
    ra, sa = a
    rb, sb = b
    for k in ("iterations", "norm", "max_residual", "converged"):
        np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)
    for x, y in zip(sa, sb):
        np.testing.assert_array_equal(x, y)


def test_graph_newton_equals_host_loop_small_batch():
    insts = problems.batch_family(5, seed=1, nx=4, nu=2, nint_range=(96, 160))
    rng = np.random.default_rng(7)
    ts = []
    for k, inst in enumerate(insts):
        n = 2 if k == 3 else inst.nodes
        ts.append(np.sort(np.concatenate([[0.0, inst.T], rng.uniform(0.0, inst.T, n - 2)])))
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(len(t) for t in ts), batch=len(insts))
    say("n_x=4 batch: host loop, then graphs")
    def timed():
        S.profile(True)
        r = solve(S, insts, ts)
        return r, {k: S.profile_get(k) for k in ("residual", "jacobian", "condense", "babd", "recover")}
    host, graph = both_ways(S, timed)
    S.profile(False)
    S.close()
    say("iterations %s, converged %s" % (host[0][0]["iterations"].tolist(), host[0][0]["converged"].tolist()))
    say("kernel groups (ms, launches): host %s" % host[1])
    say("                             graph %s" % graph[1])
    assert_same(host[0], graph[0])
    # timings: the same groups of launches timed at every replay (each
    # Newton solve's last replay also runs the groups over an empty list)
    for k, (ms, n) in graph[1].items():
        hms, hn = host[1][k]
        assert hn <= n <= hn + 2 * 4, k
        assert 0.2 * hms < ms < 5.0 * hms, k


def test_graph_stall_and_progress_tests_equal_host_loop():
    insts = problems.batch_family(8, seed=1, nint_range=(48, 80))
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(i.nodes for i in insts), batch=len(insts))
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    for stall, hom in (((8, 5, 2.0), ()), ((6, 1, 1.5), ()), ((10, 5, 2.0), (0.0,))):
        say("stall test %s, continuation %s" % (stall, hom))
        host, graph = both_ways(S, lambda: solve(S, insts, ts, hom_scales=hom, stall=stall))
        say("  iterations %s, converged %s" % (host[0]["iterations"].tolist(), host[0]["converged"].tolist()))
        assert_same(host, graph)
    for patience, ratio in ((3, 2.0), (2, 1.2)):
        say("progress test %d, %.1f" % (patience, ratio))
        host, graph = both_ways(S, lambda: solve(S, insts, ts, patience=patience, min_ratio=ratio))
        say("  iterations %s, converged %s" % (host[0]["iterations"].tolist(), host[0]["converged"].tolist()))
        assert_same(host, graph)
    # histories of the reference loop through the C ABI, graph and host
    S.set_batch_mesh(ts)
    Nt = sum(len(t) for t in ts)
    cons = np.array([i.constant_values() for i in insts])
    hists = []
    for graph in (False, True):
        S.set_newton_graph(graph)
        S.form_initial_input(np.ones((Nt, S.ny)), np.ones((Nt, S.nz)), np.ones((len(insts), S.np)))
        S.set_batch_constants(cons)
        it = np.zeros(S.B, dtype=np.int32)
        nrm = np.zeros(S.B)
        hist = np.full((S.B, 40), -7.0)
        native.check(native.lib().cs_newton(S.h, 0.1, bvp.tolerance, 40, native.iptr(it), native.ptr(nrm),
                                            native.ptr(hist), 40))
        hists.append((it, nrm, hist))
    S.close()
    for x, y in zip(*hists):
        np.testing.assert_array_equal(x, y)


def test_graph_fp32_direction_bench_pairing():
    fam = problems.batch_family(1024)
    insts = [fam[i] for i in (0, 1, 2, 5, 8, 13)]
    bvp = solver.BvpDae(insts[0].ocp_text())
    S = solver.CollocationSolver(bvp, 5, n_max=sum(i.nodes for i in insts), batch=len(insts))
    S.set_direction_precision(32)
    S.set_babd_substitution(16)
    ts = [np.linspace(0.0, i.T, i.nodes) for i in insts]
    say("n_x=20 FP32 direction, Nint %s" % [len(t) - 1 for t in ts])
    host, graph = both_ways(S, lambda: solve(S, insts, ts))
    S.close()
    say("  iterations %s, converged %s" % (host[0]["iterations"].tolist(), host[0]["converged"].tolist()))
    assert_same(host, graph)
