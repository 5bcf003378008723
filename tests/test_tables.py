"""Collocation / quadrature tables: closed-form identities and the reference's
printed literals (fixture tests/golden/reference_tables.json, extracted as
text from collocation_coefficients.py:24-635 and gauss_coefficients.py:12-131)."""
import json
import os

import numpy as np
import pytest

from oracle import tables as otab
from colloc_amd import tables as ptab

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_tables.json")))


@pytest.mark.parametrize("m", range(3, 11))
def test_lobatto_identities(m):
    L = otab.lobatto(m)
    A, b, c = L.A, L.b, L.c
    assert c[0] == 0.0 and c[-1] == 1.0
    assert np.all(A[0] == 0.0)                       # explicit first stage
    np.testing.assert_array_equal(b, A[m - 1])        # stiffly accurate
    np.testing.assert_allclose(A.sum(axis=1), c, atol=1e-15)   # row sums = c
    # collocation order conditions: sum_k A_jk c_k^(q-1) = c_j^q / q for q <= m
    for q in range(1, m + 1):
        np.testing.assert_allclose(A @ c ** (q - 1), c ** q / q, atol=2e-15)
    # Lobatto quadrature is exact to degree 2m-3
    for q in range(0, 2 * m - 2):
        assert abs(b @ c ** q - 1.0 / (q + 1)) < 1e-14
    # interior nodes are roots of P'_{m-1} on [0, 1]
    P = np.polynomial.legendre.Legendre.basis(m - 1).deriv()
    np.testing.assert_allclose(P(2 * c[1:-1] - 1), 0.0, atol=1e-11)


@pytest.mark.parametrize("m", range(3, 11))
def test_lobatto_vs_reference_literals(m):
    e = REF["lobatto"][str(m)]
    A = np.array([[np.nan if v is None else v for v in row] for row in e["A"]])
    L = otab.lobatto(m)
    # the reference's literals lose digits as m grows (collocation_coefficients.py)
    tol = {3: 1e-15, 4: 1e-14, 5: 1e-14, 6: 5e-14, 7: 1e-13, 8: 5e-13, 9: 5e-12, 10: 5e-11}[m]
    mask = ~np.isnan(A)
    assert np.max(np.abs(A[mask] - L.A[mask])) < tol
    np.testing.assert_allclose(e["b"], L.b, atol=tol)
    np.testing.assert_allclose(e["c"], L.c, atol=1e-15)
    if m == 10:
        # typo in the reference: A[9, 0] is assigned twice, A[8, 0] never
        assert [8, 0] not in e["A_assignments"]
        assert e["A_assignments"].count([9, 0]) == 2


@pytest.mark.parametrize("n", range(2, 12))
def test_gauss(n):
    G = otab.gauss(n)
    x, w = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(G.t, (x + 1) / 2, atol=1e-15)
    np.testing.assert_allclose(G.w, w / 2, atol=1e-15)
    e = REF["gauss"][str(n)]
    tol = 1e-10 if n == 11 else 1e-15     # n = 11 literals are shorter in the reference
    np.testing.assert_allclose(e["t"], G.t, atol=tol)
    np.testing.assert_allclose(e["w"], G.w, atol=1e-15)


@pytest.mark.parametrize("m", [3, 4, 5, 7])
def test_lagrange_basis(m):
    L = otab.lobatto(m)
    for j, cj in enumerate(L.c):
        np.testing.assert_allclose(otab.compute_L(m, cj), np.eye(m)[j], atol=1e-12)
        np.testing.assert_allclose(otab.compute_I(m, cj), L.A[j], atol=1e-13)
    np.testing.assert_allclose(otab.compute_I(m, 1.0), L.b, atol=1e-13)


@pytest.mark.parametrize("m", [3, 4, 5, 6, 8])
def test_product_tables_match_oracle(m):
    P = ptab.collocation_tables(m)
    O = otab.lobatto(m)
    np.testing.assert_allclose(P.A, O.A, atol=1e-16)
    np.testing.assert_allclose(P.c, O.c, atol=1e-16)
    G = otab.gauss(m + 1)
    np.testing.assert_allclose(P.gt, G.t, atol=1e-16)
    np.testing.assert_allclose(P.gw, G.w, atol=1e-16)
    for k, tau in enumerate(G.t):
        np.testing.assert_allclose(P.gL[k], otab.compute_L(m, tau), atol=1e-13)
        np.testing.assert_allclose(P.gI[k], otab.compute_I(m, tau), atol=1e-13)
    np.testing.assert_allclose(P.L0, np.eye(m)[0], atol=1e-15)
    np.testing.assert_allclose(P.L1, np.eye(m)[m - 1], atol=1e-15)
