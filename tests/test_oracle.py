"""Pins of the CPU oracle by mathematics that does not restate its formulas:
finite differences, the fully assembled Newton matrix, the LQ closed form and
hand-computed cases of the segment residual and remesh."""
import numpy as np
import pytest

from oracle import babd, families, solver, tables
from colloc_amd import problems


def small_tanh(nx=3, nu=2, seed=5, c_u=0.1):
    A, B, C, x0 = problems.draw_family(nx, nu, seed, True)
    return families.TanhFamily(A, B, x0, C=C, s=0.5, q=1.0, c_u=c_u)


def fd_jac(fun, x, h=1e-6):
    f0 = fun(x)
    J = np.zeros((f0.size, x.size))
    for k in range(x.size):
        e = np.zeros_like(x); e[k] = h
        J[:, k] = (fun(x + e) - fun(x - e)) / (2 * h)
    return J


def test_family_jacobians_by_finite_differences():
    prob = small_tanh()
    rng = np.random.default_rng(0)
    ny, nz, np_ = prob.size_y, prob.size_z, prob.size_p
    y, z, p = rng.normal(size=ny), rng.normal(size=nz), rng.normal(size=np_)
    v = np.concatenate([y, z, p])
    split = lambda w: (w[None, :ny], w[None, ny:ny + nz], w[ny + nz:])
    Jf = fd_jac(lambda w: prob.f(*split(w))[0], v)
    Jg = fd_jac(lambda w: prob.g(*split(w))[0], v)
    np.testing.assert_allclose(prob.Df(y[None], z[None], p)[0], Jf, atol=1e-8)
    np.testing.assert_allclose(prob.Dg(y[None], z[None], p)[0], Jg, atol=1e-8)
    y1 = rng.normal(size=ny)
    w = np.concatenate([y, y1, p])
    Jr = fd_jac(lambda u: prob.r(u[:ny], u[ny:2 * ny], u[2 * ny:]), w)
    np.testing.assert_allclose(prob.Dr(y, y1, p), Jr, atol=1e-8)


def test_family_is_hamiltonian_system():
    """f = [H_lambda; -H_x] and g = H_u for H = q x'x + c_u u'u + lambda'(Ax+Bu+s tanh(Cx)),
    checked by differentiating H numerically."""
    prob = small_tanh()
    nx, nu = prob.nx, prob.nu
    rng = np.random.default_rng(1)
    x, lam, u = rng.normal(size=nx), rng.normal(size=nx), rng.normal(size=nu)

    def H(x, lam, u):
        return prob.q * x @ x + prob.c_u * u @ u + lam @ (prob.A @ x + prob.B @ u + prob.s * np.tanh(prob.C @ x))
    gx = fd_jac(lambda w: np.array([H(w, lam, u)]), x)[0]
    gl = fd_jac(lambda w: np.array([H(x, w, u)]), lam)[0]
    gu = fd_jac(lambda w: np.array([H(x, lam, w)]), u)[0]
    Y, Z = np.concatenate([x, lam])[None], u[None]
    np.testing.assert_allclose(prob.f(Y, Z, None)[0], np.concatenate([gl, -gx]), atol=1e-8)
    np.testing.assert_allclose(prob.g(Y, Z, None)[0], gu, atol=1e-8)


def assemble_newton_matrix(prob, m, N, rk, t, y, y_tilde, z_tilde, p, alpha):
    """Dense Jacobian of f_q with respect to [y_0..y_{N-1}, k (stage order), p]."""
    ny, nz, np_ = prob.size_y, prob.size_z, prob.size_p
    nint = N - 1
    K = m * (ny + nz)
    J, V, D, W, B0, Bn, Vn = solver.construct_jacobian(prob, m, N, rk, t, y, y_tilde, z_tilde, p, alpha)
    nrow = nint * (K + ny) + ny + np_
    ncol = N * ny + nint * K + np_
    Jm = np.zeros((nrow, ncol))
    ky = lambda i: slice(i * ny, (i + 1) * ny)
    kk = lambda i: slice(N * ny + i * K, N * ny + (i + 1) * K)
    kp = slice(N * ny + nint * K, ncol)
    for i in range(nint):
        ra = slice(i * (K + ny), i * (K + ny) + K)
        rb = slice(i * (K + ny) + K, (i + 1) * (K + ny))
        Jm[ra, ky(i)] = J[i]; Jm[ra, kk(i)] = W[i]; Jm[ra, kp] = V[i]
        Jm[rb, ky(i)] = -np.eye(ny); Jm[rb, ky(i + 1)] = np.eye(ny); Jm[rb, kk(i)] = -D[i]
    rr = slice(nint * (K + ny), nrow)
    Jm[rr, ky(0)] = B0; Jm[rr, ky(N - 1)] = Bn; Jm[rr, kp] = Vn
    return Jm


def random_state(prob, m, N, seed):
    rng = np.random.default_rng(seed)
    y = rng.normal(scale=0.5, size=(N, prob.size_y))
    yd = rng.normal(scale=0.5, size=((N - 1) * m, prob.size_y))
    zt = rng.normal(scale=0.5, size=((N - 1) * m, prob.size_z))
    p = rng.normal(scale=0.5, size=prob.size_p)
    return y, yd, zt, p


@pytest.mark.parametrize("m,nint", [(4, 16), (5, 9), (3, 24)])
def test_condensed_direction_equals_dense_newton(m, nint):
    prob = small_tanh()
    N = nint + 1
    rk = tables.lobatto(m)
    t = np.sort(np.concatenate([[0.0, 1.3], np.random.default_rng(3).uniform(0, 1.3, N - 2)]))
    y, yd, zt, p = random_state(prob, m, N, 7)
    f_q, y_tilde, f_a, f_b, r_bc = solver.compute_f_q(prob, m, N, rk, t, y, yd, zt, p, 0.1)
    Jm = assemble_newton_matrix(prob, m, N, rk, t, y, y_tilde, zt, p, 0.1)
    # the assembled matrix is the derivative of f_q (finite differences in a few directions)
    ny, nz, np_ = prob.size_y, prob.size_z, prob.size_p
    K = m * (ny + nz)

    def unpack(q):
        yy = q[:N * ny].reshape(N, ny)
        kk = q[N * ny:N * ny + (N - 1) * K].reshape(N - 1, m, ny + nz)
        return yy, kk[:, :, :ny].reshape(-1, ny), kk[:, :, ny:].reshape(-1, nz), q[N * ny + (N - 1) * K:]

    q0 = np.concatenate([y.ravel(), np.concatenate([yd.reshape(N - 1, m, ny), zt.reshape(N - 1, m, nz)], 2).ravel(), p])
    rng = np.random.default_rng(11)
    for _ in range(3):
        d = rng.normal(size=q0.size)
        h = 1e-6
        fp = solver.compute_f_q(prob, m, N, rk, t, *unpack(q0 + h * d), 0.1)[0]
        fm = solver.compute_f_q(prob, m, N, rk, t, *unpack(q0 - h * d), 0.1)[0]
        np.testing.assert_allclose((fp - fm) / (2 * h), Jm @ d, atol=5e-7)
    delta = np.linalg.solve(Jm, -f_q)
    dy_ref, dyd_ref, dz_ref, dp_ref = unpack(delta)
    dy, dyd, dz, dp = solver.newton_direction(prob, m, N, rk, t, y, y_tilde, zt, p, 0.1, f_a, f_b, r_bc, M=2 if nint < 16 else 8)
    scale = np.max(np.abs(delta))
    np.testing.assert_allclose(dy, dy_ref, atol=1e-10 * scale)
    np.testing.assert_allclose(dyd, dyd_ref, atol=1e-10 * scale)
    np.testing.assert_allclose(dz, dz_ref, atol=1e-10 * scale)
    np.testing.assert_allclose(dp, dp_ref, atol=1e-10 * scale)


def test_householder_qr_invariants():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(12, 6))
    q, r = babd.householder_qr(a)
    np.testing.assert_allclose(q.T @ q, np.eye(12), atol=1e-14)
    np.testing.assert_allclose(q @ r, a, atol=1e-13)
    np.testing.assert_allclose(np.tril(r[:6, :6], -1), 0.0, atol=1e-13)
    # sign convention: diagonal of R has the opposite sign of the pivot column's lead
    assert np.sign(r[0, 0]) == -np.sign(a[0, 0])


def test_lq_newton_matches_closed_form():
    inst, m = problems.config(0)
    prob = families.TanhFamily(inst.A, inst.B, inst.x0, C=None, q=inst.q, c_u=inst.c_u)
    N = inst.nodes
    t = np.linspace(0, inst.T, N)
    res = solver.collocation_solver(prob, m, t, np.ones((N, prob.size_y)), np.ones((N, prob.size_z)),
                                    np.ones(prob.size_p), tol=1e-8)
    assert res["success"] == 1
    # linear problem: one Newton step reaches the tolerance
    assert res["history"][0]["iterations"] == 1
    Ys, Us, P = prob.lq_exact(inst.T, t)
    # Lobatto IIIA with m = 4 is order 6: error ~ C h^6 at h = 1/128
    assert np.max(np.abs(res["y"] - Ys)) < 1e-9
    assert np.max(np.abs(res["z"] - Us)) < 1e-9
    assert np.max(np.abs(res["p"] - P)) < 1e-9


def test_segment_residual_quirk_by_hand():
    """Scalar ODE y' = a y (no z), constant-slope state: the in-loop dt scaling
    multiplies the accumulator after every component, so for ny = 1 each
    Gauss point contributes w_i |res_i| and the total is scaled once per point."""
    a = 0.7

    class Scalar:
        size_y, size_z, size_p = 1, 0, 0
        f = staticmethod(lambda Y, Z, P: a * Y)
        g = staticmethod(lambda Y, Z, P, alpha=0.0: np.zeros((Y.shape[0], 0)))
        r = staticmethod(lambda y0, y1, p: np.zeros(1))

    m, N = 3, 3
    t = np.array([0.0, 0.5, 1.5])
    y = np.array([[1.0], [2.0], [3.0]])
    yd = np.array([[1.0], [2.0], [3.0], [0.5], [0.5], [0.5]])
    zt = np.zeros((6, 0))
    res, mx = solver.compute_segment_residual(Scalar, m, N, t, y, yd, zt, np.zeros(0), 0.0, 1e-8)
    gc = tables.gauss(m + 1)
    for i in range(2):
        h = t[i + 1] - t[i]
        rho = 0.0
        for gi in range(m + 1):
            L = tables.compute_L(m, gc.t[gi]); I = tables.compute_I(m, gc.t[gi])
            ydt = L @ yd[i * m:(i + 1) * m, 0]
            yt = y[i, 0] + h * (I @ yd[i * m:(i + 1) * m, 0])
            rho = (rho + gc.w[gi] * abs(a * yt - ydt)) * h
        assert abs(res[i] - rho / 1e-8) <= 1e-9 * abs(res[i])
    assert res[2] == 0.0 and mx == max(res)


def test_remesh_by_hand():
    N = 12
    t = np.linspace(0, 1, N)
    y = np.arange(N, dtype=float)[:, None] * np.array([[1.0, 2.0]])
    z = np.arange(N, dtype=float)[:, None]
    res = np.array([0.001] * 6 + [5.0, 0.5, 50.0, 0.5, 0.5, 0.5])
    N2, t2, y2, z2 = solver.remesh(2, 1, N, t, y, z, res)
    # nodes 0..4 all <= 1e-2 -> keep 0, 2, 4; node 5 kept; 6 (5.0) adds a midpoint;
    # 8 (50) adds three; the rest are kept
    kept = [0, 2, 4, 5, 6, 6.5, 7, 8, 8.25, 8.5, 8.75, 9, 10, 11]
    np.testing.assert_allclose(z2[:, 0], kept)
    np.testing.assert_allclose(t2, np.array(kept) / (N - 1))
    assert N2 == len(kept)


class _TwoComp:
    """ny = 2, nz = 1 test system for the acceptance residual variants."""
    size_y, size_z, size_p = 2, 1, 1
    f = staticmethod(lambda Y, Z, P: np.stack([Y[:, 1] * Z[:, 0], -Y[:, 0]], axis=1))
    g = staticmethod(lambda Y, Z, P, alpha=0.0: (Z[:, 0] - Y[:, 0] ** 2)[:, None])
    r = staticmethod(lambda y0, y1, p: np.array([y0[0] - 1.0, y1[1], p[0]]))


def _residual_by_loops(m, t, y, yd, zt, p, tol, rtype, scale, relative):
    """The acceptance residual written out with scalar loops; the stage
    polynomials are rebuilt with numpy.polynomial (fit through the Lobatto
    nodes, integrated with polyint), independent of the oracle's tables."""
    c = tables.lobatto(m).c
    gc = tables.gauss(m + 1)
    N = len(t)
    out = np.zeros(N)
    for i in range(N - 1):
        h = t[i + 1] - t[i]
        st = slice(i * m, (i + 1) * m)
        pts = []
        for gi in range(m + 1):
            tau = gc.t[gi]
            ydt = np.array([np.polyval(np.polyfit(c, yd[st, k], m - 1), tau) for k in range(2)])
            yt = np.array([y[i, k] + h * np.polyval(np.polyint(np.polyfit(c, yd[st, k], m - 1)), tau) for k in range(2)])
            z = np.array([np.polyval(np.polyfit(c, zt[st, 0], m - 1), tau)])
            hr = _TwoComp.f(yt[None], z[None], p)[0] - ydt
            gr = _TwoComp.g(yt[None], z[None], p)[0]
            pts.append((ydt, z, hr, gr, gc.w[gi]))
        if relative:
            ydm = np.max([np.abs(q[0]) for q in pts], axis=0)
            zm = np.max([np.abs(q[1]) for q in pts], axis=0)
            rh = max(abs(q[2][k]) / (1 + ydm[k]) for q in pts for k in range(2))
            rg = max(abs(q[3][0]) / (1 + zm[0]) for q in pts)
            out[i] = max(rh, rg) / tol
            continue
        rh = rg = 0.0
        for ydt, z, hr, gr, w in pts:
            for k in range(2):
                rh = rh + w * hr[k] ** 2 if rtype == 2 else (rh + w * abs(hr[k]) if rtype == 1 else max(rh, w * abs(hr[k])))
                rh = rh * h if scale else rh
            for k in range(1):
                rg = rg + w * gr[k] ** 2 if rtype == 2 else (rg + w * abs(gr[k]) if rtype == 1 else max(rg, w * abs(hr[k])))
                rg = rg * h if scale else rg
        out[i] = (np.sqrt(rh + rg) if rtype == 2 else (abs(rh) + abs(rg) if rtype == 1 else max(rh, rg))) / tol
    out[N - 1] = np.max(np.abs(_TwoComp.r(y[0], y[-1], p))) / tol
    return out


@pytest.mark.parametrize("rtype, scale, relative", [(0, True, False), (1, True, False), (2, True, False),
                                                    (0, False, False), (2, False, False), (1, True, True)])
def test_segment_residual_variants_by_loops(rtype, scale, relative):
    rng = np.random.default_rng(11 + rtype)
    m, N = 4, 5
    t = np.sort(np.concatenate([[0.0, 1.3], rng.uniform(0, 1.3, N - 2)]))
    y = rng.normal(size=(N, 2))
    yd = rng.normal(size=((N - 1) * m, 2))
    zt = rng.normal(size=((N - 1) * m, 1))
    p = np.array([0.3])
    if relative:
        res, mx = solver.compute_segment_residual2(_TwoComp, m, N, t, y, yd, zt, p, 0.0, 1e-6)
    else:
        res, mx = solver.compute_segment_residual(_TwoComp, m, N, t, y, yd, zt, p, 0.0, 1e-6,
                                                  residual_type=rtype, scale_by_time=scale)
    ref = _residual_by_loops(m, t, y, yd, zt, p, 1e-6, rtype, scale, relative)
    np.testing.assert_allclose(res, ref, rtol=1e-11)
    assert mx == np.max(res)
