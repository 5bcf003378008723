"""Learned warm start: a network maps instance parameters to the solution.

The reference solver starts Newton from the problem's initial guess (all
ones, OCP2ABVP.py:392-398) and, for the families here, needs a
continuation in the tanh scale to converge at all.  This module trains a
multilayer perceptron on solutions the solver itself produced:

  input   (x0, T) of an instance, scaled to [-1, 1]   (21 values)
  hidden  4 layers of width 1024, GELU
  output  y and z at 65 uniform normalised times      (65 x 50 values),
          standardised per output by the training set's mean and spread

Weights and activations are bf16 and every product accumulates in fp32 on
the matrix cores; the forward, backward and Adam kernels are this
project's own (libcolloc_mlp.so, include/colloc_mlp.h).  PyTorch is used
only to hold device memory and to order work on its current stream.

The network predicts the solution relative to a base: with base "lq" it
learns the difference between the instance's solution and the solution of
its linear-quadratic problem (tanh scale 0), which one Newton iteration
finds exactly from any guess (the first stage of the plain continuation).
The base carries the large multipliers of short horizons exactly, so the
network only models the nonlinear part.

A warm-started solve (warm_solve) runs that one LQ iteration, interpolates
the network's grid output onto each instance's mesh on the GPU (cubic
Lagrange) and adds it to the LQ solution at the nodes, sets the boundary
multipliers from r = 0 of the family's boundary conditions (kappa_i =
-lambda(0), kappa_f = lambda(T)), and runs Newton at the instance's own
constants, for at most a few iterations.  Instances whose Newton solve
does not converge fall back to the plain guess with continuation
(solver.solve_batch), which is also the judge of the network's error: both
must reach the same solution.
"""
import ctypes
import os
import threading
import time

import numpy as np
# PyTorch-ROCm ships its own HIP runtime under the same soname as the
# system's (libamdhip64.so.7): the first one loaded serves the process, and
# torch only works with its own.  Importing torch here, before this
# package's libraries are loaded, lets both share torch's runtime.
import torch

from . import build, native, solver

GRID = 65   # normalised-time grid of the network's output


# ------------------------------------------------------------- native layer
_mlp = None
P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
LL = ctypes.c_longlong
MLP_SIGNATURES = {
    "cm_last_error": (ctypes.c_char_p, []),
    "cm_gemm_nt": (I, [I, I, I, P, I, P, I, P, P, I, I, P, I, P, I, P]),
    "cm_transpose_bf16": (I, [I, I, P, I, P, I, P]),
    "cm_col_sum": (I, [I, I, P, I, P, P]),
    "cm_mse_grad": (I, [I, I, I, P, I, P, I, F, P, P, I, P, P]),
    "cm_adam": (I, [LL, P, P, P, P, F, F, F, F, I, P, P, I, I, P]),
    "cm_cast_bf16": (I, [LL, P, P, P]),
    "cm_grid_to_mesh": (I, [I, P, I, P, I, I, I, P, I, P, P, P, P, P, P, P]),
}
ACT_NONE, ACT_GELU, ACT_GELU_BWD = 0, 1, 2


def mlp_lib():
    """libcolloc_mlp.so with its ctypes signatures; raises if it is not built."""
    global _mlp
    if _mlp is None:
        if not os.path.exists(build.MLP_LIB_PATH):
            raise ImportError("%s is missing: build it with __graft_entry__.build()" % build.MLP_LIB_PATH)
        L = ctypes.CDLL(build.MLP_LIB_PATH)
        for name, (res, args) in MLP_SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _mlp = L
    return _mlp


def _chk(rc):
    if rc != 0:
        raise RuntimeError("colloc_mlp error %d: %s" % (rc, mlp_lib().cm_last_error().decode()))


def _torch():
    if not torch.cuda.is_available():
        if native._lib is not None:
            raise RuntimeError("torch sees no GPU: libcolloc.so was loaded before torch, so the process runs the "
                               "system's HIP runtime, which this torch build cannot use; import the learned module "
                               "(or torch) before the solver's first library call")
        raise RuntimeError("the learned warm start runs on the GPU (no device visible)")
    return torch


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _up(n, k):
    return (n + k - 1) // k * k


def gemm(A, B, M, N, K, bias=None, aux=None, act=ACT_NONE, Cf=None, Cb=None):
    """acc = A[M,K] B[N,K]^T (+bias) with the epilogue act (see colloc_mlp.h)."""
    _chk(mlp_lib().cm_gemm_nt(M, N, K, _p(A), A.stride(0), _p(B), B.stride(0), _p(bias), _p(aux),
                              aux.stride(0) if aux is not None else 0, act, _p(Cf),
                              Cf.stride(0) if Cf is not None else 0, _p(Cb), Cb.stride(0) if Cb is not None else 0,
                              _stream()))


def transpose(x, out):
    R, C = x.shape
    _chk(mlp_lib().cm_transpose_bf16(R, C, _p(x), x.stride(0), _p(out), out.stride(0), _stream()))


# ------------------------------------------------------------------ network
class WarmStartNet:
    """The perceptron, its fp32 master weights, bf16 copies and Adam state.

    sizes: [n_in, hidden..., n_out]; the input is zero-padded to a multiple
    of 32 features and the output to a multiple of 64 (the padding weights
    stay zero: their gradients are zero)."""

    def __init__(self, n_in, n_out, width=1024, depth=4, seed=0):
        torch = _torch()
        self.n_in, self.n_out = n_in, n_out
        self.k_in = _up(n_in, 32)
        self.n_outp = _up(n_out, 64)
        dims = [self.k_in] + [width] * depth + [self.n_outp]
        real_in = [n_in] + [width] * depth
        real_out = [width] * depth + [n_out]
        rng = np.random.default_rng(seed)
        dev = "cuda"
        self.W, self.b, self.Wb, self.WTb, self.m, self.v = [], [], [], [], [], []
        for l in range(len(dims) - 1):
            fi, fo = dims[l], dims[l + 1]
            w = np.zeros((fo, fi), dtype=np.float32)
            gain = np.sqrt(2.0 / real_in[l]) if l < len(dims) - 2 else np.sqrt(1.0 / real_in[l])
            w[:real_out[l], :real_in[l]] = rng.normal(0.0, gain, (real_out[l], real_in[l]))
            W = torch.tensor(w, device=dev)
            self.W.append(W)
            self.b.append(torch.zeros(fo, device=dev))
            self.Wb.append(torch.empty((fo, fi), dtype=torch.bfloat16, device=dev))
            self.WTb.append(torch.empty((fi, fo), dtype=torch.bfloat16, device=dev))
            self.m.append([torch.zeros_like(W), torch.zeros(fo, device=dev)])
            self.v.append([torch.zeros_like(W), torch.zeros(fo, device=dev)])
        self.dims = dims
        self.step = 0
        self._bufs = {}
        self.refresh_copies()

    def refresh_copies(self):
        for W, Wb, WTb in zip(self.W, self.Wb, self.WTb):
            _chk(mlp_lib().cm_cast_bf16(W.numel(), _p(W), _p(Wb), _stream()))
            transpose(Wb, WTb)

    def _buffers(self, bt):
        if bt in self._bufs:
            return self._bufs[bt]
        torch = _torch()
        dev = "cuda"
        L = len(self.dims) - 1
        bf = torch.bfloat16
        b = {"X": torch.zeros((bt, self.k_in), dtype=bf, device=dev),
             "Z": [torch.empty((bt, self.dims[l + 1]), device=dev) for l in range(L)],
             "H": [torch.empty((bt, self.dims[l + 1]), dtype=bf, device=dev) for l in range(L - 1)],
             "dZf": [torch.empty((bt, self.dims[l + 1]), device=dev) for l in range(L)],
             "dZb": [torch.empty((bt, self.dims[l + 1]), dtype=bf, device=dev) for l in range(L)],
             "dZT": [torch.empty((self.dims[l + 1], bt), dtype=bf, device=dev) for l in range(L)],
             "inT": [torch.empty((self.dims[l], bt), dtype=bf, device=dev) for l in range(L)],
             "gW": [torch.empty_like(W) for W in self.W],
             "gb": [torch.empty_like(bb) for bb in self.b],
             "T": torch.zeros((bt, self.n_outp), device=dev),
             "loss": torch.zeros(1, dtype=torch.float64, device=dev)}
        self._bufs[bt] = b
        return b

    def forward(self, X):
        """X: (bt, n_in) float32 device tensor -> (bt, n_outp) float32 output
        (pre-activations of every layer kept for backward)."""
        bt = X.shape[0]
        if bt % 32:
            raise ValueError("batch of %d rows: a multiple of 32 is required" % bt)
        B = self._buffers(bt)
        B["X"][:, :self.n_in].copy_(X)
        h = B["X"]
        L = len(self.dims) - 1
        for l in range(L):
            last = l == L - 1
            gemm(h, self.Wb[l], bt, self.dims[l + 1], self.dims[l], bias=self.b[l],
                 act=ACT_NONE if last else ACT_GELU, Cf=B["Z"][l], Cb=None if last else B["H"][l])
            if not last:
                h = B["H"][l]
        return B["Z"][L - 1]

    def gradients(self, X, T, scale=None):
        """Forward + backward on one batch: fills the gradient buffers and
        returns (loss tensor (device), buffers)."""
        bt = X.shape[0]
        B = self._buffers(bt)
        Y = self.forward(X)
        B["T"][:, :self.n_out].copy_(T)
        B["loss"].zero_()
        L = len(self.dims) - 1
        scale = 1.0 / (bt * self.n_out) if scale is None else scale
        _chk(mlp_lib().cm_mse_grad(bt, self.n_outp, self.n_out, _p(Y), Y.stride(0), _p(B["T"]), B["T"].stride(0),
                                   scale, _p(B["dZf"][L - 1]), _p(B["dZb"][L - 1]), self.n_outp, _p(B["loss"]),
                                   _stream()))
        for l in range(L - 1, -1, -1):
            fi, fo = self.dims[l], self.dims[l + 1]
            xin = B["X"] if l == 0 else B["H"][l - 1]
            transpose(B["dZb"][l], B["dZT"][l])
            transpose(xin, B["inT"][l])
            gemm(B["dZT"][l], B["inT"][l], fo, fi, bt, Cf=B["gW"][l])
            _chk(mlp_lib().cm_col_sum(bt, fo, _p(B["dZf"][l]), fo, _p(B["gb"][l]), _stream()))
            if l > 0:
                gemm(B["dZb"][l], self.WTb[l], bt, fi, fo, aux=B["Z"][l - 1], act=ACT_GELU_BWD,
                     Cf=B["dZf"][l - 1], Cb=B["dZb"][l - 1])
        self._grads = B
        return B["loss"], B

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step with the gradients of the last gradients() call."""
        self.step += 1
        B = self._grads
        for l in range(len(self.W)):
            W, Wb, WTb = self.W[l], self.Wb[l], self.WTb[l]
            _chk(mlp_lib().cm_adam(W.numel(), _p(W), _p(B["gW"][l]), _p(self.m[l][0]), _p(self.v[l][0]), lr, b1, b2,
                                   eps, self.step, _p(Wb), _p(WTb), W.shape[0], W.shape[1], _stream()))
            bb = self.b[l]
            _chk(mlp_lib().cm_adam(bb.numel(), _p(bb), _p(B["gb"][l]), _p(self.m[l][1]), _p(self.v[l][1]), lr, b1,
                                   b2, eps, self.step, None, None, 1, bb.numel(), _stream()))

    def train_step(self, X, T, lr):
        loss, _ = self.gradients(X, T)
        self.adam(lr)
        return loss

    # -- persistence -------------------------------------------------------
    def state(self):
        return {"dims": np.array(self.dims), "n_in": self.n_in, "n_out": self.n_out, "step": self.step,
                **{"W%d" % l: W.cpu().numpy() for l, W in enumerate(self.W)},
                **{"b%d" % l: b.cpu().numpy() for l, b in enumerate(self.b)}}

    @classmethod
    def from_state(cls, st):
        dims = [int(d) for d in st["dims"]]
        net = cls.__new__(cls)
        torch = _torch()
        net.n_in, net.n_out = int(st["n_in"]), int(st["n_out"])
        net.k_in, net.n_outp = dims[0], dims[-1]
        net.dims = dims
        net.step = int(st["step"])
        net._bufs = {}
        net.W = [torch.tensor(st["W%d" % l], device="cuda") for l in range(len(dims) - 1)]
        net.b = [torch.tensor(st["b%d" % l], device="cuda") for l in range(len(dims) - 1)]
        net.Wb = [torch.empty(W.shape, dtype=torch.bfloat16, device="cuda") for W in net.W]
        net.WTb = [torch.empty((W.shape[1], W.shape[0]), dtype=torch.bfloat16, device="cuda") for W in net.W]
        net.m = [[torch.zeros_like(W), torch.zeros_like(b)] for W, b in zip(net.W, net.b)]
        net.v = [[torch.zeros_like(W), torch.zeros_like(b)] for W, b in zip(net.W, net.b)]
        net.refresh_copies()
        return net


# ------------------------------------------------------- family interface
def features(x0, T, T_range=(0.5, 2.0)):
    """Network input of instances with initial states x0 (B, nx) and
    horizons T (B,): x0 as is (already in [-1, 1]) and T scaled to [-1, 1]."""
    x0 = np.atleast_2d(np.asarray(x0, dtype=np.float64))
    T = np.asarray(T, dtype=np.float64).reshape(-1, 1)
    t = 2.0 * (T - T_range[0]) / (T_range[1] - T_range[0]) - 1.0
    return np.hstack([x0, t]).astype(np.float32)


def grid_targets(sol, ny, nz):
    """(GRID * (ny + nz),) target row of a solution on a uniform mesh whose
    interval count is a multiple of GRID - 1: y and z at the grid nodes,
    node-major."""
    y, z = sol["y"], sol["z"]
    n = y.shape[0] - 1
    if n % (GRID - 1):
        raise ValueError("mesh of %d intervals: a multiple of %d is needed" % (n, GRID - 1))
    step = n // (GRID - 1)
    return np.hstack([y[::step], z[::step]]).reshape(-1)


class WarmStart:
    """A trained network with its output standardisation, for one family.
    base: "none" (the network predicts the solution) or "lq" (the
    difference from the instance's linear-quadratic solution)."""

    def __init__(self, net, mean, sd, ny, nz, nx, base="none"):
        torch = _torch()
        self.net = net
        self.ny, self.nz, self.nx = ny, nz, nx
        self.base = base
        self.mean = torch.tensor(np.asarray(mean, dtype=np.float32), device="cuda")
        self.sd = torch.tensor(np.asarray(sd, dtype=np.float32), device="cuda")
        # the network's activation buffers are shared: one guess at a time
        # when several host threads drive handles
        self._lock = threading.Lock()

    def save(self, path):
        st = self.net.state()
        np.savez(path, mean=self.mean.cpu().numpy(), sd=self.sd.cpu().numpy(), ny=self.ny, nz=self.nz, nx=self.nx,
                 base=self.base, **st)

    @classmethod
    def load(cls, path):
        st = dict(np.load(path))
        net = WarmStartNet.from_state(st)
        base = str(st["base"]) if "base" in st else "none"
        return cls(net, st["mean"], st["sd"], int(st["ny"]), int(st["nz"]), int(st["nx"]), base)

    def predict_grid(self, feats):
        """Standardised network output (B, n_outp) for a feature matrix; the
        batch is padded to a multiple of 32 rows."""
        torch = _torch()
        B = feats.shape[0]
        bt = _up(B, 32)
        X = torch.zeros((bt, feats.shape[1]), device="cuda")
        X[:B] = torch.from_numpy(feats).cuda()
        return self.net.forward(X)

    def guess_on_meshes(self, solv, feats):
        """Initial guess of every instance of solv's current batch (its meshes
        set) from the network, written with cs_form_initial_input_device.
        With base "lq" the solver's current state must be the instances' LQ
        solution on these meshes: the network's output is added to it.
        Safe to call from several threads (one at a time: the network's
        buffers are shared; the work ends with a stream synchronisation)."""
        with self._lock:
            return self._guess_on_meshes(solv, feats)

    def _guess_on_meshes(self, solv, feats):
        torch = _torch()
        B = solv.B
        G = self.predict_grid(feats)
        t = torch.from_numpy(np.ascontiguousarray(solv.t_span)).cuda()
        noff = torch.from_numpy(solv.node_offsets().astype(np.int32)).cuda()
        N = solv.N
        Y = torch.empty((N, self.ny), dtype=torch.float64, device="cuda")
        Z = torch.empty((N, max(self.nz, 1)), dtype=torch.float64, device="cuda")
        nc = self.ny + self.nz
        yb = zb = None
        p_base = np.zeros((B, solv.np))
        if self.base == "lq":
            yp, zp, pp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
            native.check(native.lib().cs_node_solution_device(solv.h, ctypes.byref(yp), ctypes.byref(zp),
                                                              ctypes.byref(pp)))
            yb, zb = yp, zp
            _, _, _, p_host = solv.get_state_parts(p=True)
            p_base = np.asarray(p_host).reshape(B, solv.np)
        _chk(mlp_lib().cm_grid_to_mesh(B, _p(noff), N, _p(t), GRID, nc, self.ny, _p(G), G.stride(0), _p(self.mean),
                                       _p(self.sd), yb, zb, _p(Y), _p(Z), _stream()))
        # boundary multipliers from r = 0: kappa_i = -lambda(0), kappa_f = lambda(T)
        g = G[:B, :GRID * nc].double() * self.sd.double() + self.mean.double()
        g = g.view(B, GRID, nc)
        nx = self.nx
        Pm = torch.cat([-g[:, 0, nx:2 * nx], g[:, GRID - 1, nx:2 * nx]], dim=1)
        Pm = (Pm + torch.from_numpy(p_base).cuda()).contiguous()
        torch.cuda.current_stream().synchronize()
        native.check(native.lib().cs_form_initial_input_device(solv.h, _p(Y), _p(Z), _p(Pm)))
        return Y, Z, Pm


def warm_solve(ws, solv, t_spans, feats, constants, lq_constants, tol=1e-8, max_iter=200, max_iter_warm=12,
               alpha=0.1, patience=8, min_ratio=2.0, cont=(8, 5, 2.0)):
    """Warm-started batched solve of the instances (meshes t_spans): for an
    "lq" network, one Newton solve of the LQ problems from the all-ones
    guess (converged after one iteration), then the network's guess and at
    most max_iter_warm Newton iterations at the instances' constants; from
    iteration `patience` on, an instance whose norm falls by less than a
    factor min_ratio in an iteration stops (patience=0 disables the test).
    cont = (start, window, ratio): the instances that stage leaves
    unconverged continue damped Newton from where they stopped, for the
    rest of the max_iter budget, each stopping when its norm fell by less
    than `ratio` over the last `window` iterations from iteration `start` on
    (cs_newton_stall); None: no continuation (a caller's fallback takes
    over).  Returns the solve_batch-style result dict ("iterations" has one
    row per stage run; "warm_converged" is the acceptance after the capped
    stage); "lq_iterations" holds the iterations of the LQ stage."""
    torch = _torch()
    B = len(t_spans)
    solv.set_batch_mesh(t_spans)
    lq_its = np.zeros(B, dtype=int)
    if ws.base == "lq":
        N = solv.N
        ones_y = torch.ones((N, solv.ny), dtype=torch.float64, device="cuda")
        ones_z = torch.ones((N, max(solv.nz, 1)), dtype=torch.float64, device="cuda")
        ones_p = torch.ones((B, max(solv.np, 1)), dtype=torch.float64, device="cuda")
        torch.cuda.current_stream().synchronize()
        native.check(native.lib().cs_form_initial_input_device(solv.h, _p(ones_y), _p(ones_z), _p(ones_p)))
        solv.set_batch_constants(lq_constants)
        it, _, _ = solv.newton(alpha, tol, max_iter)
        lq_its = np.atleast_1d(it)
    ws.guess_on_meshes(solv, feats)
    r = solver.newton_stages(solv, constants, [], alpha, tol, max_iter_warm, patience=patience, min_ratio=min_ratio)
    r["lq_iterations"] = lq_its
    r["warm_converged"] = r["converged"].copy()
    left = max_iter - max_iter_warm
    if cont is not None and left >= 2 and not r["converged"].all():
        it, nrm, _ = solv.newton(alpha, tol, left, int(cont[0]), float(cont[2]), window=int(cont[1]))
        it, nrm = np.atleast_1d(it), np.atleast_1d(nrm)
        _, mx = solv.segment_residual(alpha, tol, 1, True, False)
        mx = np.atleast_1d(mx)
        r["iterations"] = np.vstack([r["iterations"], it[None, :]])
        r["norm"] = nrm
        r["max_residual"] = mx
        r["converged"] = (nrm < tol) & (mx <= 1.0) & (it < left - 1)
    return r


# ------------------------------------------------------ data and training
def generate_solutions(bvp, m, insts, nint, schedule=(0.0, 0.25), chunk_nodes=320000, base="lq", log=None):
    """Solve instances of a batched family on uniform meshes of nint
    intervals (plain guess + continuation in the tanh scale, whose first
    stage, scale 0, is the LQ problem), batched.  Returns (features, targets,
    mask of the instances whose Newton solve converged); with base "lq" the
    targets are the solution minus the LQ solution.  The mesh-acceptance
    test is not applied: a training target only needs the collocation
    solution at the grid nodes, which a coarse mesh already gives far below
    the network's own error."""
    if base == "lq" and (not schedule or schedule[0] != 0.0):
        raise ValueError("an LQ base needs a continuation that starts at scale 0")
    ny, nz, np_ = bvp.size_y, bvp.size_z, bvp.size_p
    per = max(1, min(len(insts), chunk_nodes // (nint + 1)))
    S = solver.CollocationSolver(bvp, m, n_max=per * (nint + 1), batch=per)
    feats = features([i.x0 for i in insts], [i.T for i in insts])
    targets = np.zeros((len(insts), GRID * (ny + nz)), dtype=np.float32)
    ok = np.zeros(len(insts), dtype=bool)
    t0 = time.time()
    for c0 in range(0, len(insts), per):
        ch = insts[c0:c0 + per]
        ts = [np.linspace(0.0, i.T, nint + 1) for i in ch]
        Nt = len(ch) * (nint + 1)
        S.set_batch_mesh(ts)
        S.form_initial_input(np.ones((Nt, ny)), np.ones((Nt, nz)), np.ones((len(ch), np_)))
        cons = np.array([i.constant_values() for i in ch])
        hom = [np.array([i.constant_values(s=s) for i in ch]) for s in schedule]
        base_rows = np.zeros((len(ch), targets.shape[1]))
        if base == "lq":
            S.set_batch_constants(hom[0])
            S.newton(0.1, bvp.tolerance, bvp.maximum_newton_iterations)
            base_rows = np.array([grid_targets(s, ny, nz) for s in solver.batch_solutions(S)])
            hom = hom[1:]
        r = solver.newton_stages(S, cons, hom, 0.1, bvp.tolerance, bvp.maximum_newton_iterations)
        sols = solver.batch_solutions(S)
        conv = r["norm"] < bvp.tolerance
        for k, s in enumerate(sols):
            if conv[k]:
                targets[c0 + k] = grid_targets(s, ny, nz) - base_rows[k]
        ok[c0:c0 + per] = conv
        if log:
            log("solutions %d/%d: %d converged (%.1f s)" % (c0 + len(ch), len(insts), int(ok[:c0 + per].sum()),
                                                           time.time() - t0))
    S.close()
    return feats, targets, ok


def train(feats, targets, epochs=60, batch=512, lr=1e-3, seed=0, width=1024, depth=4, log=None):
    """Train a WarmStartNet on standardised targets with Adam (cosine-decayed
    learning rate).  Returns (net, mean, sd, the last epoch's mean squared
    error of the standardised targets)."""
    torch = _torch()
    mean = targets.mean(axis=0)
    sd = targets.std(axis=0) + 1e-6
    Tn = ((targets - mean) / sd).astype(np.float32)
    net = WarmStartNet(feats.shape[1], targets.shape[1], width=width, depth=depth, seed=seed)
    Xd = torch.from_numpy(feats).cuda()
    Td = torch.from_numpy(Tn).cuda()
    n = (feats.shape[0] // batch) * batch
    rng = np.random.default_rng(seed)
    steps = epochs * (n // batch)
    k = 0
    ep_loss = 0.0
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(feats.shape[0])[:n]).cuda()
        tot = torch.zeros(1, dtype=torch.float64, device="cuda")
        for s in range(0, n, batch):
            idx = perm[s:s + batch]
            lr_k = lr * 0.5 * (1.0 + np.cos(np.pi * k / steps))
            tot += net.train_step(Xd[idx], Td[idx], lr_k)
            k += 1
        ep_loss = float(tot.item()) / (n // batch) / (batch * targets.shape[1])   # mean squared error
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log("epoch %d: mean loss %.3e" % (ep, ep_loss))
    return net, mean, sd, ep_loss


# ------------------------------------------------------- the batched family
MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_model")
FAMILY_MODEL = os.path.join(MODEL_DIR, "warmstart_tanh20_m5.npz")
TRAIN_OFFSET = 100000   # training instances are family members TRAIN_OFFSET + i (held-out: 0 .. 99999)


def family_training_set(n, seed=1, nx=20, nu=10):
    """Members TRAIN_OFFSET .. TRAIN_OFFSET + n - 1 of the batched family (the
    benchmark and the evaluation use members 0, 1, ...)."""
    from . import problems
    A, B, C, _ = problems.draw_family(nx, nu, seed, True)
    out = []
    for i in range(TRAIN_OFFSET, TRAIN_OFFSET + n):
        rng = np.random.default_rng([seed, i])
        x0 = rng.uniform(-1.0, 1.0, nx)
        T = rng.uniform(0.5, 2.0)
        out.append(problems.FamilyInstance(nx=nx, nu=nu, A=A, B=B, C=C, x0=x0, s=0.5, q=1.0, c_u=0.1, T=T,
                                           nodes=257, tolerance=1e-8, instance_x0=True, homotopy=True))
    return out


def fit_family(n_train=50000, nint=256, epochs=200, batch=512, lr=1e-3, width=1024, depth=4, path=None, base="lq",
               log=None):
    """Generate n_train solutions of the n_x=20 batched family with the
    plain batched solver, train the network on the converged ones and save
    it (with the timing and fit figures) to path."""
    from . import problems
    m = 5
    insts = family_training_set(n_train)
    bvp = solver.BvpDae(problems.batch_family(1)[0].ocp_text())
    t0 = time.time()
    feats, targets, ok = generate_solutions(bvp, m, insts, nint, base=base, log=log)
    t_gen = time.time() - t0
    t0 = time.time()
    net, mean, sd, loss = train(feats[ok], targets[ok], epochs=epochs, batch=batch, lr=lr, width=width, depth=depth,
                                log=log)
    _torch().cuda.synchronize()
    t_train = time.time() - t0
    ws = WarmStart(net, mean, sd, bvp.size_y, bvp.size_z, nx=20, base=base)
    path = path or FAMILY_MODEL
    os.makedirs(os.path.dirname(path), exist_ok=True)
    ws.save(path)
    import hashlib
    with open(path, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    info = {"n_train": int(n_train), "converged": int(ok.sum()), "nint": nint, "epochs": epochs, "batch": batch,
            "width": width, "depth": depth, "base": base, "seconds_generate": t_gen, "seconds_train": t_train,
            "final_mse": loss, "family_seed": 1, "members": [TRAIN_OFFSET, TRAIN_OFFSET + int(n_train)],
            "training_seed": 0, "lr": lr, "sha256": digest}
    with open(path + ".json", "w") as f:
        import json
        json.dump(info, f, indent=1)
    return ws, info

