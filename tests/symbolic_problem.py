"""Numerical problem object (f, g, Df, Dg, r, Dr vectorised over stages) built
from an OCP's symbolic BVP-DAE, for oracle runs on problems that have no
hand-derived family.  The expressions are the emitter's (pinned against
finite differences in test_emitter.py); the oracle's linear algebra and
Newton loop are independent of the HIP path."""
import numpy as np
import sympy as sp

from colloc_amd import emitter, ocp as ocpmod


def _expand(hoister, exprs):
    defs = hoister.needed(exprs)
    sub = {}
    for s, d in defs:
        sub[s] = d.xreplace(sub)
    return [sp.sympify(e).xreplace(sub) for e in exprs]


class SymbolicProblem:
    def __init__(self, ocp_text):
        sym = emitter.BvpDaeSymbolic(ocpmod.parse_ocp(ocp_text))
        self.sym = sym
        self.size_y, self.size_z, self.size_p = sym.size_y, sym.size_z, sym.size_p
        consts = {s: v for s, v in sym.constants.items()}
        args = sym.y + sym.z + sym.p + [sym.alpha]
        H = sym.hoist_fg

        def lam(exprs):
            ex = [e.xreplace(consts) for e in _expand(H, exprs)]
            return sp.lambdify(args, ex, modules="numpy")
        self._f = lam(sym.f_exprs)
        self._g = lam(sym.g_exprs)
        self._Df = lam([v for row in sym.Df_exprs for v in row])
        self._Dg = lam([v for row in sym.Dg_exprs for v in row])
        yp = sym.y + sym.p
        self._ri = sp.lambdify(yp, [e.xreplace(consts) for e in _expand(sym.hoist_ri, sym.ri_exprs)], "numpy")
        self._rf = sp.lambdify(yp, [e.xreplace(consts) for e in _expand(sym.hoist_rf, sym.rf_exprs)], "numpy")
        self._Dri = sp.lambdify(yp, [e.xreplace(consts) for row in sym.Dri_exprs for e in _expand(sym.hoist_ri, row)], "numpy")
        self._Drf = sp.lambdify(yp, [e.xreplace(consts) for row in sym.Drf_exprs for e in _expand(sym.hoist_rf, row)], "numpy")

    def _call(self, fn, Y, Z, P, alpha, shape):
        S = Y.shape[0]
        cols = [Y[:, k] for k in range(self.size_y)] + [Z[:, k] for k in range(self.size_z)]
        cols += [np.full(S, P[k]) for k in range(self.size_p)] + [np.full(S, alpha)]
        out = fn(*cols)
        return np.stack([np.broadcast_to(np.asarray(o, dtype=np.float64), (S,)) for o in out], axis=1).reshape((S,) + shape)

    def f(self, Y, Z, P):
        return self._call(self._f, Y, Z, P, 0.0, (self.size_y,))

    def g(self, Y, Z, P, alpha=0.0):
        return self._call(self._g, Y, Z, P, alpha, (self.size_z,))

    def Df(self, Y, Z, P):
        nj = self.size_y + self.size_z + self.size_p
        return self._call(self._Df, Y, Z, P, 0.0, (self.size_y, nj))

    def Dg(self, Y, Z, P, alpha=0.0):
        nj = self.size_y + self.size_z + self.size_p
        return self._call(self._Dg, Y, Z, P, alpha, (self.size_z, nj))

    def r(self, y0, y1, p):
        return np.concatenate([np.asarray(self._ri(*y0, *p), float).ravel(),
                               np.asarray(self._rf(*y1, *p), float).ravel()])

    def Dr(self, y0, y1, p):
        ny, np_ = self.size_y, self.size_p
        Di = np.asarray(self._Dri(*y0, *p), float).reshape(-1, ny + np_)
        Df = np.asarray(self._Drf(*y1, *p), float).reshape(-1, ny + np_)
        D = np.zeros((ny + np_, 2 * ny + np_))
        ni = Di.shape[0]
        D[:ni, :ny], D[:ni, 2 * ny:] = Di[:, :ny], Di[:, ny:]
        D[ni:, ny:2 * ny], D[ni:, 2 * ny:] = Df[:, :ny], Df[:, ny:]
        return D
