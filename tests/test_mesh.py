"""Mesh deletion/insertion (collocation_solver_parallel_2d_shared.py:2106-2311):
the product's remesh against a hand-computed case and against the oracle's
line-by-line restatement on random residual patterns."""
import numpy as np
import pytest

from colloc_amd import mesh
from oracle import solver as osolver


def test_hand_case():
    t = np.arange(7, dtype=float)
    y = np.stack([t, 2 * t], axis=1)
    z = (10 * t)[:, None]
    r = np.array([1e-3] * 5 + [5.0, 0.5])
    # nodes 0..4 all <= 1e-2: keep 0, 2, 4; then 5, 6.  Interval [5, 6] has
    # residual 5 (1 < r <= 10): one midpoint.
    N, tn, yn, zn = mesh.remesh(2, 1, 7, t, y, z, r)
    assert N == 6
    np.testing.assert_array_equal(tn, [0, 2, 4, 5, 5.5, 6])
    np.testing.assert_array_equal(yn[:, 1], 2 * tn)
    np.testing.assert_array_equal(zn[:, 0], 10 * tn)


def test_three_nodes_for_large_residual():
    t = np.array([0.0, 1.0, 2.0])
    y = t[:, None].copy()
    z = t[:, None].copy()
    N, tn, _, _ = mesh.remesh(1, 1, 3, t, y, z, np.array([20.0, 0.5, 0.0]))
    np.testing.assert_array_equal(tn, [0, 0.25, 0.5, 0.75, 1.0, 2.0])


@pytest.mark.parametrize("seed", range(8))
def test_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(5, 60))
    t = np.sort(rng.uniform(0, 1, N))
    y, z = rng.normal(size=(N, 3)), rng.normal(size=(N, 2))
    # mixture of tiny, moderate and large residuals, runs of small ones
    r = np.where(rng.uniform(size=N) < 0.6, 1e-3, 10 ** rng.uniform(-1, 2, N))
    a = mesh.remesh(3, 2, N, t, y, z, r)
    b = osolver.remesh(3, 2, N, t, y, z, r)
    assert a[0] == b[0]
    for u, v in zip(a[1:], b[1:]):
        np.testing.assert_allclose(u, v, rtol=0, atol=1e-15)


def test_sanity_check():
    assert mesh.mesh_sanity_check(10, 0, 10, 2000, 3) is False
    assert mesh.mesh_sanity_check(2, 0, 10, 2000, 3) is True
    assert mesh.mesh_sanity_check(3000, 0, 10, 2000, 3) is True
    assert mesh.mesh_sanity_check(10, 11, 10, 2000, 3) is True
