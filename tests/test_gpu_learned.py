"""The learned warm start on the GPU (libcolloc_mlp.so through ctypes).

Kernels against the CPU oracle (oracle/mlp.py), which rounds to bf16 at the
points where the kernels store bf16 and sums exactly: products and
epilogues to fp32-accumulation accuracy, whole-network gradients per layer,
Adam, and the grid-to-mesh interpolation against a float64 Lagrange
restatement.  Then the method end to end on the n_x=4 family: a small
network trained on solver-generated solutions must give a guess from which
Newton converges more often than from the plain guess with continuation,
in fewer iterations, to the same solutions.
"""
import numpy as np
import pytest
import torch

from colloc_amd import learned, problems, solver
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def bf16_tensor(x):
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (96, 80, 64), (512, 3264, 1024), (33, 7, 96)])
def test_gemm_epilogues(M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = omlp.bf16(rng.normal(size=(M, K)))
    B = omlp.bf16(rng.normal(size=(N, K)) / np.sqrt(K))
    bias = rng.normal(size=N).astype(np.float32)
    aux = rng.normal(size=(M, N)).astype(np.float32)
    Ad, Bd = bf16_tensor(A), bf16_tensor(B)
    ref = A @ B.T
    say("gemm %dx%dx%d" % (M, N, K))
    for act in (learned.ACT_NONE, learned.ACT_GELU, learned.ACT_GELU_BWD):
        Cf = torch.zeros((M, N), device="cuda")
        Cb = torch.zeros((M, N), dtype=torch.bfloat16, device="cuda")
        learned.gemm(Ad, Bd, M, N, K, bias=torch.from_numpy(bias).cuda(), aux=torch.from_numpy(aux).cuda(), act=act,
                     Cf=Cf, Cb=Cb)
        torch.cuda.synchronize()
        v = ref + bias
        if act == learned.ACT_GELU_BWD:
            v = v * omlp.gelu_grad(aux.astype(np.float64))
        assert np.max(np.abs(Cf.cpu().numpy() - v)) <= 2e-5 * max(1.0, np.max(np.abs(v)))
        out = omlp.gelu(v) if act == learned.ACT_GELU else v
        cb = Cb.float().cpu().numpy()
        # one bf16 rounding of an fp32 value within 2e-5 of out
        assert np.max(np.abs(cb - out) / np.maximum(np.abs(out), 1e-3)) <= 2.0 ** -8 + 1e-4


def test_transpose_and_col_sum():
    rng = np.random.default_rng(1)
    x = omlp.bf16(rng.normal(size=(70, 45)))
    xd = bf16_tensor(x)
    out = torch.zeros((45, 70), dtype=torch.bfloat16, device="cuda")
    learned.transpose(xd, out)
    f = torch.from_numpy(rng.normal(size=(37, 29)).astype(np.float32)).cuda()
    s = torch.zeros(29, device="cuda")
    learned._chk(learned.mlp_lib().cm_col_sum(37, 29, learned._p(f), 29, learned._p(s), learned._stream()))
    torch.cuda.synchronize()
    assert np.array_equal(out.float().cpu().numpy(), x.T)
    assert np.allclose(s.cpu().numpy(), f.cpu().numpy().astype(np.float64).sum(axis=0), rtol=1e-5, atol=1e-5)


def small_net(seed=3):
    net = learned.WarmStartNet(21, 100, width=64, depth=2, seed=seed)
    # non-zero biases, so the bias paths are exercised
    g = torch.Generator(device="cpu").manual_seed(seed)
    for b in net.b:
        b.copy_((torch.randn(b.shape, generator=g) * 0.1).cuda())
    return net


def oracle_params(net):
    Ws = [W.cpu().numpy().astype(np.float64) for W in net.W]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    return Ws, bs


def test_network_gradients_against_oracle():
    net = small_net()
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (64, 21)).astype(np.float32)
    T = rng.normal(size=(64, 100)).astype(np.float32)
    say("forward + backward, dims %s" % net.dims)
    loss, B = net.gradients(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), scale=1.0 / (64 * 100))
    torch.cuda.synchronize()
    Ws, bs = oracle_params(net)
    Xp = np.zeros((64, net.k_in))
    Xp[:, :21] = X
    Tp = np.zeros((64, net.n_outp))
    Tp[:, :100] = T
    ref_loss, gW, gb = omlp.loss_and_grads(Ws, bs, Xp, Tp, n_real=100, scale=1.0 / (64 * 100), bf16_storage=True)
    Y, _ = omlp.forward(Ws, bs, Xp, bf16_storage=True)
    assert rel(net.forward(torch.from_numpy(X).cuda()).cpu().numpy(), Y) < 1e-5
    assert abs(loss.item() / (64 * 100) - ref_loss) <= 1e-5 * ref_loss
    for l in range(len(Ws)):
        assert rel(B["gW"][l].cpu().numpy(), gW[l]) < 1e-3, l
        assert rel(B["gb"][l].cpu().numpy(), gb[l]) < 1e-3, l
        # padding rows / columns receive no gradient
        assert np.all(B["gW"][l].cpu().numpy()[gW[l] == 0.0] == 0.0)


def test_adam_against_oracle():
    net = small_net()
    rng = np.random.default_rng(6)
    X = torch.from_numpy(rng.uniform(-1, 1, (32, 21)).astype(np.float32)).cuda()
    T = torch.from_numpy(rng.normal(size=(32, 100)).astype(np.float32)).cuda()
    Ws, bs = oracle_params(net)
    m1 = [np.zeros_like(w) for w in Ws]
    m2 = [np.zeros_like(w) for w in Ws]
    for step in (1, 2):
        net.gradients(X, T)
        g = [gw.cpu().numpy().astype(np.float64) for gw in net._grads["gW"]]
        net.adam(1e-3)
        torch.cuda.synchronize()
        for l in range(len(Ws)):
            omlp.adam_step(Ws[l], g[l], m1[l], m2[l], step, 1e-3)
            assert np.max(np.abs(net.W[l].cpu().numpy() - Ws[l])) < 1e-6
            assert np.array_equal(net.Wb[l].float().cpu().numpy(), omlp.bf16(net.W[l].cpu().numpy()))
            assert np.array_equal(net.WTb[l].float().cpu().numpy(), omlp.bf16(net.W[l].cpu().numpy()).T)
        Ws = [W.cpu().numpy().astype(np.float64) for W in net.W]


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (4096, 21)).astype(np.float32)
    P = rng.normal(size=(21, 100)) / 5
    T = np.sin(X @ P).astype(np.float32)
    say("training 40 epochs")
    net, mean, sd, loss = learned.train(X, T, epochs=40, batch=256, lr=2e-3, width=128, depth=2)
    assert loss < 0.05    # standardised targets: loss 1 is the mean predictor


def test_grid_to_mesh_interpolation():
    rng = np.random.default_rng(8)
    ng, ny, nz = learned.GRID, 3, 2
    nc = ny + nz
    meshes = [np.linspace(0.0, 1.3, 40), np.sort(np.concatenate([[0.2, 2.0], rng.uniform(0.2, 2.0, 11)]))]
    noff = np.array([0, 40, 53], dtype=np.int32)
    G = rng.normal(size=(2, 64 * 52)).astype(np.float32)   # row stride > ng * nc
    mean = rng.normal(size=ng * nc).astype(np.float32)
    sd = rng.uniform(0.5, 2.0, ng * nc).astype(np.float32)
    t = np.concatenate(meshes)
    Y = torch.zeros((53, ny), dtype=torch.float64, device="cuda")
    Z = torch.zeros((53, nz), dtype=torch.float64, device="cuda")
    td, Gd = torch.from_numpy(t).cuda(), torch.from_numpy(G).cuda()
    nd = torch.from_numpy(noff).cuda()
    md, sdd = torch.from_numpy(mean).cuda(), torch.from_numpy(sd).cuda()   # kept alive until the kernel has run
    learned._chk(learned.mlp_lib().cm_grid_to_mesh(2, learned._p(nd), 53, learned._p(td), ng, nc, ny,
                                                   learned._p(Gd), G.shape[1], learned._p(md), learned._p(sdd),
                                                   None, None, learned._p(Y), learned._p(Z), learned._stream()))
    torch.cuda.synchronize()
    for b, tm in enumerate(meshes):
        vals = (mean.astype(np.float64) + sd.astype(np.float64) * G[b, :ng * nc].astype(np.float64)).reshape(ng, nc)
        for j, tv in enumerate(tm):
            s = (tv - tm[0]) / (tm[-1] - tm[0]) * (ng - 1)
            k0 = min(max(int(np.floor(s)) - 1, 0), ng - 4)
            ks = np.arange(k0, k0 + 4)
            w = np.array([np.prod([(s - ks[q]) / (ks[a] - ks[q]) for q in range(4) if q != a]) for a in range(4)])
            ref = w @ vals[ks]
            n = noff[b] + j
            assert np.allclose(Y[n].cpu().numpy(), ref[:ny], rtol=1e-12, atol=1e-12)
            assert np.allclose(Z[n].cpu().numpy(), ref[ny:], rtol=1e-12, atol=1e-12)
    # with base values (here the output itself, in place): the sum
    Y0, Z0 = Y.clone(), Z.clone()
    learned._chk(learned.mlp_lib().cm_grid_to_mesh(2, learned._p(nd), 53, learned._p(td), ng, nc, ny,
                                                   learned._p(Gd), G.shape[1], learned._p(md), learned._p(sdd),
                                                   learned._p(Y), learned._p(Z), learned._p(Y), learned._p(Z),
                                                   learned._stream()))
    torch.cuda.synchronize()
    assert torch.allclose(Y, 2 * Y0, rtol=1e-14, atol=0) and torch.allclose(Z, 2 * Z0, rtol=1e-14, atol=0)


def test_warm_start_gains_over_plain_guess():
    """n_x=4 family: train on 3072 solver-generated solutions (Nint=64), then
    solve 64 held-out instances (meshes Nint~U{96..160}) from the network's
    guess at the final constants, and from the plain guess with the
    continuation; compare iterations, convergence and solutions."""
    m = 5
    fam = problems.batch_family(3072 + 64, seed=1, nx=4, nu=2, nint_range=(96, 160))
    held, train_insts = fam[:64], fam[64:]
    bvp = solver.BvpDae(held[0].ocp_text())
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    say("generating %d training solutions" % len(train_insts))
    feats, targets, ok = learned.generate_solutions(bvp, m, train_insts, nint=64, base="lq", log=say)
    say("training on %d converged solutions" % ok.sum())
    net, mean, sd, loss = learned.train(feats[ok], targets[ok], epochs=150, batch=256, lr=2e-3, width=256, depth=4,
                                        log=say)
    ws = learned.WarmStart(net, mean, sd, ny, nz, nx=4, base="lq")
    ts = [np.linspace(0.0, i.T, i.nodes) for i in held]
    cons = np.array([i.constant_values() for i in held])
    S = solver.CollocationSolver(bvp, m, n_max=sum(len(t) for t in ts), batch=len(held))
    say("plain guess + continuation")
    Y1 = [np.ones((len(t), ny)) for t in ts]
    plain = solver.solve_batch(S, ts, Y1, [np.ones((len(t), nz)) for t in ts], [np.ones(np_) for _ in ts], cons,
                               [np.array([i.constant_values(s=0.0) for i in held])], tol=bvp.tolerance,
                               max_iter=bvp.maximum_newton_iterations)
    sol_plain = solver.batch_solutions(S)
    say("LQ solve + network correction")
    warm = learned.warm_solve(ws, S, ts, learned.features([i.x0 for i in held], [i.T for i in held]), cons,
                              np.array([i.constant_values(s=0.0) for i in held]), tol=bvp.tolerance,
                              max_iter=bvp.maximum_newton_iterations, max_iter_warm=bvp.maximum_newton_iterations)
    sol_warm = solver.batch_solutions(S)
    assert np.all(warm["lq_iterations"] == 1)   # the LQ problem is linear: one Newton step
    # Newton convergence (the mesh acceptance test does not depend on the guess)
    cp, cw = plain["norm"] < bvp.tolerance, warm["norm"] < bvp.tolerance
    both = cp & cw
    it_plain = plain["iterations"].sum(axis=0)
    it_warm = warm["iterations"].sum(axis=0) + warm["lq_iterations"]
    say("Newton converged: plain %d, warm %d of %d; mean iterations (both converged) plain %.2f, warm %.2f"
        % (cp.sum(), cw.sum(), len(held), it_plain[both].mean(), it_warm[both].mean()))
    # this small network (width 256, ~2.5k samples) reaches a standardised MSE
    # of ~5e-3: more instances converge, in at least 20% fewer iterations
    assert cw.sum() > cp.sum()
    assert it_warm[both].mean() < 0.8 * it_plain[both].mean()
    # The necessary conditions of this nonconvex OCP can have several
    # solutions: both starts must reach the same one on almost every
    # instance (any other is still a solution: its residual is below tol).
    same = [all(rel(sol_warm[k][key], sol_plain[k][key]) < 1e-6 for key in ("y", "z", "p"))
            for k in np.nonzero(both)[0]]
    say("same solution from both starts: %d of %d" % (sum(same), len(same)))
    assert sum(same) >= 0.9 * len(same)
    # a tightly capped warm stage, alone and with the continuation of the
    # instances it leaves (cs_newton_stall from where they stopped): the
    # continuation leaves the converged instances untouched and adds others
    feats = learned.features([i.x0 for i in held], [i.T for i in held])
    lq = np.array([i.constant_values(s=0.0) for i in held])
    capped = learned.warm_solve(ws, S, ts, feats, cons, lq, tol=bvp.tolerance, max_iter=bvp.maximum_newton_iterations,
                                max_iter_warm=3, patience=2, min_ratio=2.0, cont=None)
    sol_capped = solver.batch_solutions(S)
    cont = learned.warm_solve(ws, S, ts, feats, cons, lq, tol=bvp.tolerance, max_iter=bvp.maximum_newton_iterations,
                              max_iter_warm=3, patience=2, min_ratio=2.0, cont=(10, 5, 2.0))
    sol_cont = solver.batch_solutions(S)
    say("capped warm stage: %d converged; with the continuation: %d"
        % (capped["converged"].sum(), cont["converged"].sum()))
    np.testing.assert_array_equal(cont["warm_converged"], capped["converged"])
    assert np.all(cont["converged"][capped["converged"]])
    for k in np.nonzero(capped["converged"])[0]:
        for key in ("y", "z", "p"):
            np.testing.assert_array_equal(sol_cont[k][key], sol_capped[k][key])
    assert cont["converged"].sum() > capped["converged"].sum()
    assert cont["iterations"].shape[0] == 2
