"""Closed-form BVP-DAE functions for the synthetic OCP families.

The reference turns an OCP into a BVP-DAE by forming the Hamiltonian
H = L + lambda'f and emitting f = [H_lambda; -H_x], g = H_u and the boundary
conditions r = [Gamma; lambda + Gamma_x' kappa_i; Psi; lambda - phi_x - Psi_x' kappa_f]
(OCP2ABVP.py:41-182).  For the two benchmark families those expressions are
derived here by hand, so that they pin the product's emitted device code
(straight-line and polynomial-table forms) and the symbolic problem of
tests/symbolic_problem.py.

Family ``tanh``:  x' = A x + B u + s * tanh(C x),
                  cost int (q x'x + c_u u'u) dt,  x(0) = x0,  x(T) = 0.
Family ``lq`` is the same with s = 0 (no C term).

Variables: y = [x, lambda] (ny = 2 nx), z = u (nz = nu), p = [kappa_i, kappa_f]
(np = 2 nx).  Jacobian column order is [y | z | p] for Df/Dg and
[y0 | y1 | p] for Dr (OCP2ABVP.py:264-326).
"""
import numpy as np


class TanhFamily:
    """Vectorised f, g, r and Jacobians of the tanh OCP family.

    All stage functions take Y (S, ny), Z (S, nz), P (np,) and return arrays
    with a leading stage axis.
    """

    def __init__(self, A, B, x0, C=None, s=0.5, q=1.0, c_u=1.0):
        self.A = np.asarray(A, dtype=np.float64)
        self.B = np.asarray(B, dtype=np.float64)
        self.nx = self.A.shape[0]
        self.nu = self.B.shape[1]
        self.C = None if C is None else np.asarray(C, dtype=np.float64)
        self.s = s if C is not None else 0.0
        self.q = q
        self.c_u = c_u
        self.x0 = np.asarray(x0, dtype=np.float64)
        self.size_y = 2 * self.nx
        self.size_z = self.nu
        self.size_p = 2 * self.nx

    # ---- helpers -------------------------------------------------------
    def _split(self, Y, Z):
        nx = self.nx
        return Y[:, :nx], Y[:, nx:2 * nx], Z

    def _tanh(self, X):
        if self.C is None:
            return np.zeros_like(X), np.zeros_like(X)
        W = X @ self.C.T
        T = np.tanh(W)
        return T, 1.0 - T * T          # tanh, sech^2

    # ---- stage functions ----------------------------------------------
    def f(self, Y, Z, P):
        X, L, U = self._split(Y, Z)
        T, S2 = self._tanh(X)
        fx = X @ self.A.T + U @ self.B.T + self.s * T
        # -H_x = -(2 q x + A' lambda + s C'(sech^2(Cx) * lambda))
        fl = -(2.0 * self.q * X + L @ self.A + (self.s * S2 * L) @ (self.C if self.C is not None else 0.0 * self.A))
        return np.concatenate([fx, fl], axis=1)

    def g(self, Y, Z, P, alpha=0.0):
        X, L, U = self._split(Y, Z)
        return 2.0 * self.c_u * U + L @ self.B

    def Df(self, Y, Z, P):
        nx, nu, np_ = self.nx, self.nu, self.size_p
        S = Y.shape[0]
        X, L, U = self._split(Y, Z)
        T, S2 = self._tanh(X)
        out = np.zeros((S, 2 * nx, 2 * nx + nu + np_))
        # d(fx)/dx = A + s diag(sech^2) C ; d(fx)/du = B
        out[:, :nx, :nx] = self.A[None]
        out[:, :nx, 2 * nx:2 * nx + nu] = self.B[None]
        # d(fl)/dlambda = -(A' + s C' diag(sech^2))
        out[:, nx:2 * nx, nx:2 * nx] = -self.A.T[None]
        out[:, nx:2 * nx, :nx] += -2.0 * self.q * np.eye(nx)[None]
        if self.C is not None:
            C = self.C
            out[:, :nx, :nx] += self.s * S2[:, :, None] * C[None]
            out[:, nx:2 * nx, nx:2 * nx] += -self.s * (C.T[None] * S2[:, None, :])
            # d/dx of -s C'(sech^2(w) lambda): d sech^2/dw = -2 tanh sech^2
            dS2 = -2.0 * T * S2                       # (S, nx) over k
            wk = self.s * dS2 * L                     # weight per k
            out[:, nx:2 * nx, :nx] += -np.einsum('ki,sk,kj->sij', C, wk, C)
        return out

    def Dg(self, Y, Z, P, alpha=0.0):
        nx, nu, np_ = self.nx, self.nu, self.size_p
        S = Y.shape[0]
        out = np.zeros((S, nu, 2 * nx + nu + np_))
        out[:, :, nx:2 * nx] = self.B.T[None]
        out[:, :, 2 * nx:2 * nx + nu] = 2.0 * self.c_u * np.eye(nu)[None]
        return out

    # ---- boundary conditions (host functions in the reference) ----------
    def r(self, y0, y1, p):
        nx = self.nx
        ki, kf = p[:nx], p[nx:2 * nx]
        return np.concatenate([y0[:nx] - self.x0, y0[nx:2 * nx] + ki,
                               y1[:nx], y1[nx:2 * nx] - kf])

    def Dr(self, y0, y1, p):
        nx = self.nx
        ny, np_ = 2 * nx, 2 * nx
        D = np.zeros((ny + np_, 2 * ny + np_))
        I = np.eye(nx)
        D[0:nx, 0:nx] = I                      # x(0) - x0
        D[nx:2 * nx, nx:2 * nx] = I            # lambda(0) + kappa_i
        D[nx:2 * nx, 2 * ny:2 * ny + nx] = I
        D[2 * nx:3 * nx, ny:ny + nx] = I       # x(T)
        D[3 * nx:4 * nx, ny + nx:ny + 2 * nx] = I   # lambda(T) - kappa_f
        D[3 * nx:4 * nx, 2 * ny + nx:2 * ny + 2 * nx] = -I
        return D

    # ---- LQ closed-form optimal solution (independent of the reference) ----
    def lq_exact(self, T, t):
        """Exact state/costate/control of the LQ problem (s = 0).

        Necessary conditions: x' = A x + B u, lambda' = -(2 q x + A' lambda),
        2 c_u u + B' lambda = 0.  With Phi = expm(t M),
        M = [[A, -B B'/(2 c_u)], [-2 q I, -A']], x(T) = 0 fixes lambda(0).
        """
        from scipy.linalg import expm
        if self.C is not None:
            raise ValueError("closed form holds only for the linear-quadratic family")
        nx = self.nx
        M = np.block([[self.A, -self.B @ self.B.T / (2.0 * self.c_u)],
                      [-2.0 * self.q * np.eye(nx), -self.A.T]])
        Phi = expm(T * M)
        lam0 = -np.linalg.solve(Phi[:nx, nx:], Phi[:nx, :nx] @ self.x0)
        s0 = np.concatenate([self.x0, lam0])
        Ys = np.array([expm(ti * M) @ s0 for ti in t])
        Us = -(Ys[:, nx:] @ self.B) / (2.0 * self.c_u)
        return Ys, Us, np.concatenate([-lam0, Ys[-1, nx:]])
