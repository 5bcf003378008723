"""CPU oracle of the warm-start network (test infrastructure only).

A multilayer perceptron y = W_L g(... g(W_1 x + b_1) ...) + b_L with the
exact GELU g(z) = z Phi(z), trained on the squared error.  The reference
solver has no learned component, so there is no reference file to follow;
the oracle is pinned by central finite differences of its own loss
(tests/test_oracle_mlp.py) and by the closed forms of GELU and its
derivative.

bf16=True rounds to bfloat16 (round to nearest even) where the device
kernels store bf16: the layer inputs, the weights, and the back-propagated
gradients fed to the products; sums are exact (float64), which the device's
fp32 accumulation matches to ~1e-6 relative.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    if nan.any():
        out[nan] = np.nan
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(Ws, bs, X, bf16_storage=False):
    """Returns (Y, cache): Y the output, cache the layer inputs and
    pre-activations for backward."""
    h = _r(X, bf16_storage)
    ins, pre = [], []
    L = len(Ws)
    for l, (W, b) in enumerate(zip(Ws, bs)):
        ins.append(h)
        z = h @ _r(W, bf16_storage).T + b
        pre.append(z)
        if l < L - 1:
            h = _r(gelu(z), bf16_storage)
        else:
            h = z
    return h, (ins, pre)


def loss_and_grads(Ws, bs, X, T, n_real=None, scale=1.0, bf16_storage=False):
    """Squared-error loss sum over the first n_real output columns and its
    gradients: dL/dW_l, dL/db_l for L = scale * sum (Y - T)^2."""
    Y, (ins, pre) = forward(Ws, bs, X, bf16_storage)
    n_real = Y.shape[1] if n_real is None else n_real
    d = np.zeros_like(Y)
    d[:, :n_real] = Y[:, :n_real] - T[:, :n_real]
    loss = float(np.sum(d[:, :n_real] ** 2))
    dz = scale * 2.0 * d
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dzr = _r(dz, bf16_storage)
        gW[l] = dzr.T @ ins[l]
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dz = (dzr @ _r(Ws[l], bf16_storage)) * gelu_grad(pre[l - 1])
    return scale * loss, gW, gb


def adam_step(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Bias-corrected Adam (in place on w, m1, m2)."""
    m1[...] = b1 * m1 + (1 - b1) * g
    m2[...] = b2 * m2 + (1 - b2) * g * g
    w[...] = w - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return w
