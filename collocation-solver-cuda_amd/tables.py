"""Collocation and quadrature tables used by the kernels.

The reference hard-codes Lobatto IIIA tableaux (collocation_coefficients.py:10-635),
the Lagrange basis and its integral on the Lobatto nodes (compute_L / compute_I,
:638-788) and Gauss-Legendre points (gauss_coefficients.py:6-131) as decimal
literals.  Here they are computed from their definitions in 40-digit
arithmetic and rounded once to float64:
  Lobatto nodes c = {0, 1} U {(1 + x)/2 : P'_{m-1}(x) = 0},
  A[j, k] = int_0^{c_j} l_k(s) ds,  b = A[m-1, :],
  Gauss nodes (1 + x)/2 with P_n(x) = 0, weights w_n / 2.
"""
from functools import lru_cache

import mpmath
import numpy as np

_DPS = 40


def _legendre_coeffs(n):
    """P_n coefficients, highest power first (explicit sum formula)."""
    co = [mpmath.mpf(0)] * (n + 1)
    for k in range(n // 2 + 1):
        co[2 * k] = ((-1) ** k * mpmath.factorial(2 * n - 2 * k)
                     / (2 ** n * mpmath.factorial(k) * mpmath.factorial(n - k) * mpmath.factorial(n - 2 * k)))
    return co


def _deriv(co):
    n = len(co) - 1
    return [co[i] * (n - i) for i in range(n)]


def _roots(co):
    if len(co) < 2:
        return []
    return sorted(mpmath.re(r) for r in mpmath.polyroots(co, maxsteps=500, extraprec=300))


@lru_cache(maxsize=None)
def _lobatto(m):
    with mpmath.workdps(_DPS):
        inner = _roots(_deriv(_legendre_coeffs(m - 1)))
        c = [mpmath.mpf(0)] + [(1 + x) / 2 for x in inner] + [mpmath.mpf(1)]
        basis = []
        for k in range(m):
            # l_k as coefficients (highest first) via products of (s - c_q)/(c_k - c_q)
            poly = [mpmath.mpf(1)]
            for q in range(m):
                if q != k:
                    poly = [a - b for a, b in zip(poly + [0], [0] + [v * c[q] for v in poly])]
                    poly = [v / (c[k] - c[q]) for v in poly]
            basis.append(poly)
        integ = [[v / (len(pl) - i) for i, v in enumerate(pl)] + [mpmath.mpf(0)] for pl in basis]
        A = [[mpmath.polyval(integ[k], c[j]) for k in range(m)] for j in range(m)]
        return c, basis, integ, A


@lru_cache(maxsize=None)
def _gauss(n):
    with mpmath.workdps(_DPS):
        co = _legendre_coeffs(n)
        xs = _roots(co)
        dco = _deriv(co)
        ws = [2 / ((1 - x * x) * mpmath.polyval(dco, x) ** 2) for x in xs]
        return [(1 + x) / 2 for x in xs], [w / 2 for w in ws]


class CollocationTables:
    """Everything the kernels need for m Lobatto IIIA stages."""

    def __init__(self, m):
        if m < 2:
            raise ValueError("need m >= 2")
        c, basis, integ, A = _lobatto(m)
        gt, gw = _gauss(m + 1)
        f = lambda v: float(v)
        self.m = m
        self.c = np.array([f(v) for v in c])
        self.A = np.array([[f(v) for v in row] for row in A])
        self.b = self.A[m - 1].copy()
        self.gt = np.array([f(v) for v in gt])
        self.gw = np.array([f(v) for v in gw])
        with mpmath.workdps(_DPS):
            self.gL = np.array([[f(mpmath.polyval(basis[k], tau)) for k in range(m)] for tau in gt])
            self.gI = np.array([[f(mpmath.polyval(integ[k], tau)) for k in range(m)] for tau in gt])
            self.L0 = np.array([f(mpmath.polyval(basis[k], 0)) for k in range(m)])
            self.L1 = np.array([f(mpmath.polyval(basis[k], 1)) for k in range(m)])


@lru_cache(maxsize=None)
def collocation_tables(m):
    return CollocationTables(m)
