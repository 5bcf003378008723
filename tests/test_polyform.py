"""The polynomial form of the BVP-DAE (polyform.py) against the symbolic form
(emitter.BvpDaeSymbolic, sympy), both restating OCP2ABVP.py:41-182.

At seeded points, the table programs' NumPy evaluation (the same walk as
the device evaluator) of f, g, the stage Jacobian block [Df; Dg], r and Dr
equals the sympy expressions' values, on the n_x = 4 tanh family, the
reference's constrained examples (inequality smoothing: sqrt nodes; a free
final time; a parameter in the dynamics) and the parameter test problem.
The two forms are built independently (chain rule over hoisted sympy
functions against sparse-polynomial gradients), so a dropped term, a wrong
sign or a wrong column order in either fails here; the sympy form is pinned
by finite differences in test_emitter.py.  The n_x = 64 family's form
builds in seconds (sympy's takes minutes).
"""
import time

import numpy as np
import pytest
import sympy as sp

from colloc_amd import emitter, ocp, polyform, problems


def sym_values(S, exprs, hoister, vals):
    syms = S.y + S.z + S.p + list(S.instance_constants) + [S.alpha]
    consts = {s: v for s, v in S.constants.items() if s not in S.instance_constants}
    sub = {s: sp.Float(float(v)) for s, v in zip(syms, vals)}
    sub.update({s: sp.Float(v) for s, v in consts.items()})
    hv = {}
    for s, d in hoister.defs:
        hv[s] = sp.Float(float(d.xreplace(sub).xreplace(hv).evalf()))
    out = []
    for e in exprs:
        v = sp.sympify(e).xreplace(sub).xreplace(hv)
        out.append(float(v.evalf()) if v != 0 else 0.0)
    return np.array(out)


def rel_err(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1.0, np.max(np.abs(b)))


CASES = [("tanh4", lambda: problems.config(1, nx=4, nu=2, nint=32)[0].ocp_text()),
         ("tanh4_batch", lambda: problems.batch_family(1, nx=4, nu=2, nint_range=(16, 64))[0].ocp_text()),
         ("param_osc", lambda: problems.PARAM_OSC_OCP)] + \
        [(k, (lambda k=k: problems.CONSTRAINED_EXAMPLES[k])) for k in sorted(problems.CONSTRAINED_EXAMPLES)]


@pytest.mark.parametrize("name,text", CASES, ids=[c[0] for c in CASES])
def test_polynomial_form_equals_symbolic_form(name, text):
    o = ocp.parse_ocp(text())
    P = polyform.BvpDaePoly(o)
    S = emitter.BvpDaeSymbolic(o)
    ny, nz, np_ = S.size_y, S.size_z, S.size_p
    assert (P.size_y, P.size_z, P.size_p) == (ny, nz, np_)
    assert P.active_p_columns() == emitter.active_p_columns(S)
    progs, mask, vals = emitter.table_programs(P)
    rng = np.random.default_rng(0)
    base = rng.uniform(0.2, 1.0, (3, P.nb))
    fg, _ = progs["fg"].evaluate(base)
    jd, _ = progs["jac"].evaluate(base)
    blk = [e for row in S.Df_exprs for e in row] + [e for row in S.Dg_exprs for e in row]
    ndr = 2 * ny + np_
    for q in range(3):
        ref = sym_values(S, S.f_exprs + S.g_exprs, S.hoist_fg, base[q])
        assert rel_err(fg[q, :ny + nz], ref) < 1e-13
        refj = sym_values(S, blk, S.hoist_fg, base[q])
        got = np.array(vals, dtype=float)
        dyn = np.flatnonzero(np.array(mask) == 0)
        got[dyn] = jd[q, dyn] if jd.shape[1] > 0 else 0.0
        assert rel_err(got, refj) < 1e-13
        # boundary conditions: the initial rows at y0 = base y, the final at y1
        y0, y1 = base[q, :ny], base[(q + 1) % 3, :ny]
        b0 = base[q].copy()
        b1 = base[q].copy()
        b1[:ny] = y1
        r0, _ = progs["ri"].evaluate(b0[None])
        r1, _ = progs["rf"].evaluate(b1[None])
        d0, _ = progs["dri"].evaluate(b0[None])
        d1, _ = progs["drf"].evaluate(b1[None])
        nri = len(S.ri_exprs)
        vi = base[q].copy()
        ref_ri = sym_values(S, S.ri_exprs, S.hoist_ri, vi)
        vf = base[q].copy()
        vf[:ny] = y1
        ref_rf = sym_values(S, S.rf_exprs, S.hoist_rf, vf)
        assert rel_err(r0[0, :nri], ref_ri) < 1e-13
        assert rel_err(r1[0, nri:nri + len(S.rf_exprs)], ref_rf) < 1e-13
        dri = sym_values(S, [e for row in S.Dri_exprs for e in row], S.hoist_ri, vi).reshape(nri, ny + np_)
        drf = sym_values(S, [e for row in S.Drf_exprs for e in row], S.hoist_rf, vf).reshape(-1, ny + np_)
        Dr = np.zeros((ny + np_) * ndr)
        Dr[:d0.shape[1]] += d0[0]
        Dr[:d1.shape[1]] += d1[0]
        Dr = Dr.reshape(ny + np_, ndr)
        assert rel_err(Dr[:nri, :ny], dri[:, :ny]) < 1e-13
        assert rel_err(Dr[:nri, 2 * ny:], dri[:, ny:]) < 1e-13
        assert np.all(Dr[:nri, ny:2 * ny] == 0)
        assert rel_err(Dr[nri:, ny:2 * ny], drf[:, :ny]) < 1e-13
        assert rel_err(Dr[nri:, 2 * ny:], drf[:, ny:]) < 1e-13
        assert np.all(Dr[nri:, :ny] == 0)


def test_node_derivatives_by_finite_differences():
    """Chain rule through every node kind: d/dx of a polynomial in the nodes
    against central differences of its value."""
    x, w = sp.symbols("x w")
    funcs = [sp.tanh(2 * x + w), sp.exp(x * w), sp.sin(x) * sp.cos(w * x), sp.tan(x / 3), sp.sinh(x) + sp.cosh(w),
             sp.sqrt(x * x + w), sp.log(1 + x * w), sp.asin(x / 2), sp.acos(x / 3), sp.atan(x * w),
             sp.Abs(x - 0.1) * w, (x + w) ** 2.5, (x * w + 1) ** -1, x ** -3 * w]
    for e in funcs:
        E = polyform.Engine(2)
        p = E.from_sympy(e, {}, {x: 0, w: 1})
        g = E.grad(p)
        prog = polyform.Program(E, [p, g.get(0, {}), g.get(1, {})], [0, 1, 2])
        pt = np.array([[0.37, 0.61]])
        v, _ = prog.evaluate(pt)
        for a in (0, 1):
            h = 1e-6
            up, dn = pt.copy(), pt.copy()
            up[0, a] += h
            dn[0, a] -= h
            fd = (prog.evaluate(up)[0][0, 0] - prog.evaluate(dn)[0][0, 0]) / (2 * h)
            assert abs(v[0, 1 + a] - fd) < 1e-7 * max(1.0, abs(fd)), (e, a)
        assert abs(v[0, 0] - float(e.subs({x: 0.37, w: 0.61}))) < 1e-13


def test_large_family_builds_in_seconds():
    A, B, C, x0 = problems.draw_family(64, 32, 1, True)
    inst = problems.FamilyInstance(nx=64, nu=32, A=A, B=B, C=C, x0=x0, s=0.5, q=1.0, c_u=0.1, T=2.0, nodes=65,
                                   homotopy=True)
    t0 = time.time()
    P = polyform.BvpDaePoly(ocp.parse_ocp(inst.ocp_text()))
    progs, mask, _ = emitter.table_programs(P)
    assert time.time() - t0 < 120
    assert (P.size_y, P.size_z, P.size_p) == (128, 32, 128)
    # the costate block d(-H_x)/dx: 64 x 64 entries of 128 terms (two monomials per tanh node)
    assert progs["jac"].terms > 64 * 64 * 128
    assert mask.count(0) == 3 * 64 * 64
