"""CPU checks of the warm-start network's oracle and host helpers.

The oracle's gradients are pinned by central finite differences of its own
loss (float64, no rounding); its bf16 rounding by torch's float32 ->
bfloat16 conversion (round to nearest even); GELU by its closed form at
special points and by finite differences of itself; Adam by the first step,
which moves every weight by lr against the sign of its gradient."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from colloc_amd import build, learned
from oracle import mlp as omlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000), [0.0, -0.0, 1.0, 65504.0]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(omlp.bf16(x), ref)
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert omlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert omlp.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_gelu_closed_forms():
    assert omlp.gelu(np.array([0.0]))[0] == 0.0
    z = np.array([-30.0, 30.0])
    assert np.allclose(omlp.gelu(z), [0.0, 30.0])
    x = np.linspace(-4, 4, 33)
    h = 1e-6
    fd = (omlp.gelu(x + h) - omlp.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - omlp.gelu_grad(x))) < 1e-8
    assert abs(omlp.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15


def tiny_net(seed=0, dims=(5, 8, 7, 6)):
    rng = np.random.default_rng(seed)
    Ws = [rng.normal(size=(dims[l + 1], dims[l])) / np.sqrt(dims[l]) for l in range(len(dims) - 1)]
    bs = [rng.normal(size=dims[l + 1]) * 0.1 for l in range(len(dims) - 1)]
    X = rng.normal(size=(4, dims[0]))
    T = rng.normal(size=(4, dims[-1]))
    return Ws, bs, X, T


def test_oracle_gradients_by_finite_differences():
    Ws, bs, X, T = tiny_net()
    n_real = 5   # the last output column is padding and does not count
    loss, gW, gb = omlp.loss_and_grads(Ws, bs, X, T, n_real=n_real, scale=0.3)
    h = 1e-6
    for l in range(len(Ws)):
        for idx in [(0, 0), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1), (1, 2)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (omlp.loss_and_grads(Wp, bs, X, T, n_real, 0.3)[0] - omlp.loss_and_grads(Wm, bs, X, T, n_real, 0.3)[0]) / (2 * h)
            assert abs(fd - gW[l][idx]) < 1e-6 * max(1.0, abs(fd)), (l, idx)
        for j in (0, bs[l].shape[0] - 1):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][j] += h
            bm[l][j] -= h
            fd = (omlp.loss_and_grads(Ws, bp, X, T, n_real, 0.3)[0] - omlp.loss_and_grads(Ws, bm, X, T, n_real, 0.3)[0]) / (2 * h)
            assert abs(fd - gb[l][j]) < 1e-6 * max(1.0, abs(fd)), (l, j)
    # padding column: no gradient reaches the last layer's row for it
    assert np.all(gW[-1][n_real:] == 0.0) and np.all(gb[-1][n_real:] == 0.0)


def test_adam_first_step_is_a_signed_step():
    w = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    m1, m2 = np.zeros(3), np.zeros(3)
    omlp.adam_step(w, g, m1, m2, 1, lr=0.01, eps=0.0)
    assert np.allclose(w, [0.99, -1.99, 2.99], rtol=0, atol=1e-12)


def test_features_and_grid_targets():
    f = learned.features(np.array([[0.5, -1.0]]), np.array([2.0]))
    assert f.dtype == np.float32 and np.allclose(f, [[0.5, -1.0, 1.0]])
    f = learned.features(np.zeros((2, 1)), np.array([0.5, 1.25]))
    assert np.allclose(f[:, 1], [-1.0, 0.0])
    n = 2 * (learned.GRID - 1)
    y = np.arange((n + 1) * 2, dtype=float).reshape(n + 1, 2)
    z = -np.arange(n + 1, dtype=float).reshape(n + 1, 1)
    row = learned.grid_targets({"y": y, "z": z}, 2, 1).reshape(learned.GRID, 3)
    assert np.array_equal(row[:, :2], y[::2]) and np.array_equal(row[:, 2], z[::2, 0])
    with pytest.raises(ValueError):
        learned.grid_targets({"y": y[:-1], "z": z[:-1]}, 2, 1)


def test_mlp_library_exports_header():
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "colloc_mlp.h")).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(cm_[a-z_0-9]+)\s*\(", text)))
    assert names == sorted(learned.MLP_SIGNATURES)
    if not os.path.exists(build.MLP_LIB_PATH):
        build.build_mlp_library()
    L = ctypes.CDLL(build.MLP_LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
