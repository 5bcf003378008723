"""Collocation and quadrature tables, computed from their closed forms.

Restates ``collocation_coefficients.lobatto`` (collocation_coefficients.py:10-635),
``compute_L`` / ``compute_I`` (collocation_coefficients.py:638-788) and
``gauss_coefficients.gauss`` (gauss_coefficients.py:6-131).  The reference
prints these as decimal literals; here they are derived in 50-digit arithmetic
and rounded once to double, so they agree with the reference's literals to the
literals' own precision (checked in tests against a fixture of those literals).
"""
from functools import lru_cache

import mpmath
import numpy as np

_DPS = 50


def _legendre_poly(n):
    """Coefficients (low -> high) of the Legendre polynomial P_n on [-1, 1]."""
    p0, p1 = [mpmath.mpf(1)], [mpmath.mpf(0), mpmath.mpf(1)]
    if n == 0:
        return p0
    for k in range(1, n):
        # (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1}
        xp = [mpmath.mpf(0)] + p1
        pm = p0 + [mpmath.mpf(0)] * (len(xp) - len(p0))
        p0, p1 = p1, [((2 * k + 1) * a - k * b) / (k + 1) for a, b in zip(xp, pm)]
    return p1


def _poly_roots(coeffs):
    return sorted(mpmath.polyroots(list(reversed(coeffs)), maxsteps=400, extraprec=200))


@lru_cache(maxsize=None)
def _lobatto_mp(m):
    with mpmath.workdps(_DPS):
        pn = _legendre_poly(m - 1)
        dp = [k * pn[k] for k in range(1, len(pn))]          # P'_{m-1}
        inner = _poly_roots(dp) if len(dp) > 1 else []
        xs = [mpmath.mpf(-1)] + [mpmath.re(r) for r in inner] + [mpmath.mpf(1)]
        c = [(x + 1) / 2 for x in xs]
        A = [[mpmath.mpf(0)] * m for _ in range(m)]
        for k in range(m):
            # Lagrange basis l_k on nodes c, integrated from 0 to c_j
            poly = [mpmath.mpf(1)]
            den = mpmath.mpf(1)
            for q in range(m):
                if q == k:
                    continue
                poly = [mpmath.mpf(0)] + poly
                for i in range(len(poly) - 1):
                    poly[i] -= c[q] * poly[i + 1]
                den *= c[k] - c[q]
            ipoly = [mpmath.mpf(0)] + [poly[i] / (i + 1) for i in range(len(poly))]
            for j in range(m):
                A[j][k] = mpmath.polyval(list(reversed(ipoly)), c[j]) / den
        return c, A


class lobatto:
    """Lobatto IIIA tableau with m stages (collocation_coefficients.py:10-22).

    ``A[j, k] = int_0^{c_j} l_k(s) ds``; ``b = A[m-1, :]`` because c_{m-1} = 1;
    the first row of A is zero because c_0 = 0.
    """

    def __init__(self, m):
        if m < 2:
            raise ValueError("Lobatto IIIA needs m >= 2 stages")
        c, A = _lobatto_mp(m)
        self.m = m
        self.stages = m
        self.order = 2 * (m - 1)
        self.A = np.array([[float(v) for v in row] for row in A], dtype=np.float64)
        self.c = np.array([float(v) for v in c], dtype=np.float64)
        self.b = self.A[m - 1, :].copy()


@lru_cache(maxsize=None)
def _gauss_mp(n):
    with mpmath.workdps(_DPS):
        xs = [mpmath.re(r) for r in _poly_roots(_legendre_poly(n))]
        pn = _legendre_poly(n)
        dp = [k * pn[k] for k in range(1, len(pn))]
        ws = [2 / ((1 - x * x) * mpmath.polyval(list(reversed(dp)), x) ** 2) for x in xs]
        return [(x + 1) / 2 for x in xs], [w / 2 for w in ws]


class gauss:
    """Gauss-Legendre nodes t and weights w on [0, 1] (gauss_coefficients.py:6-131)."""

    def __init__(self, n):
        t, w = _gauss_mp(n)
        self.n = n
        self.t = np.array([float(v) for v in t], dtype=np.float64)
        self.w = np.array([float(v) for v in w], dtype=np.float64)


def compute_L(m, t):
    """Lagrange basis on the Lobatto nodes evaluated at t (collocation_coefficients.py:638-712)."""
    c = lobatto(m).c
    L = np.ones(m)
    for k in range(m):
        for q in range(m):
            if q != k:
                L[k] *= (t - c[q]) / (c[k] - c[q])
    return L


def compute_I(m, t):
    """Integral from 0 to t of the Lagrange basis (collocation_coefficients.py:715-788)."""
    c = lobatto(m).c
    I = np.zeros(m)
    for k in range(m):
        others = [c[q] for q in range(m) if q != k]
        poly = np.polynomial.Polynomial.fromroots(others)
        poly = poly / np.prod([c[k] - cq for cq in others])
        I[k] = poly.integ()(t)
    return I
