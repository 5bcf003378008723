"""The symbolic BVP-DAE the emitter generates code from (restating
OCP2ABVP.make_abvp, OCP2ABVP.py:41-182).

1. For the tanh family it equals the hand-derived f, g, r and Jacobians of
   oracle.families (which test_oracle pins by finite differences).
2. For a general OCP (parameters, terminal penalty, inequality and
   state-inequality constraints with the alpha smoothing, sin/exp/sqrt/abs)
   the chain-rule Jacobians over hoisted terms equal central differences.
3. The emitted C text is well formed (balanced, every output written).
"""
import numpy as np
import pytest
import sympy as sp

from colloc_amd import emitter, ocp as ocpmod, problems
from oracle import families


def evaluate(hoister, exprs, env):
    """Numeric values of exprs (which may contain hoisted symbols)."""
    env = dict(env)
    for s, d in hoister.needed(exprs):
        env[s] = sp.Float(float(d.xreplace(env).evalf()), 30)
    out = []
    for e in exprs:
        v = sp.sympify(e).xreplace(env)
        out.append(float(v.evalf()))
    return np.array(out)


def base_env(sym, const_overrides=None):
    env = {s: sp.Float(v, 30) for s, v in sym.constants.items()}
    for s, v in (const_overrides or {}).items():
        env[s] = sp.Float(v, 30)
    return env


def bind(env, syms, vals):
    env = dict(env)
    env.update({s: sp.Float(float(v), 30) for s, v in zip(syms, vals)})
    return env


@pytest.fixture(scope="module")
def tanh_case():
    inst, m = problems.config(1, nx=3, nu=2, nint=8)
    sym = emitter.BvpDaeSymbolic(ocpmod.parse_ocp(inst.ocp_text()))
    fam = families.TanhFamily(inst.A, inst.B, inst.x0, C=inst.C, s=inst.s, q=inst.q, c_u=inst.c_u)
    S = [s for s in sym.instance_constants if s.name == "S"]
    env = base_env(sym, {S[0]: inst.s})
    return sym, fam, env


def test_tanh_family_functions(tanh_case):
    sym, fam, env = tanh_case
    rng = np.random.default_rng(3)
    for _ in range(3):
        y = rng.normal(size=sym.size_y)
        z = rng.normal(size=sym.size_z)
        p = rng.normal(size=sym.size_p)
        e = bind(bind(bind(env, sym.y, y), sym.z, z), sym.p, p)
        e[sym.alpha] = sp.Float(0.1, 30)
        H = sym.hoist_fg
        f = evaluate(H, sym.f_exprs, e)
        g = evaluate(H, sym.g_exprs, e)
        Df = evaluate(H, [v for row in sym.Df_exprs for v in row], e).reshape(sym.size_y, -1)
        Dg = evaluate(H, [v for row in sym.Dg_exprs for v in row], e).reshape(sym.size_z, -1)
        Y, Z = y[None], z[None]
        np.testing.assert_allclose(f, fam.f(Y, Z, p)[0], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(g, fam.g(Y, Z, p)[0], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(Df, fam.Df(Y, Z, p)[0], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(Dg, fam.Dg(Y, Z, p)[0], rtol=1e-13, atol=1e-13)


def test_tanh_family_boundary_conditions(tanh_case):
    sym, fam, env = tanh_case
    rng = np.random.default_rng(4)
    y0, y1, p = rng.normal(size=sym.size_y), rng.normal(size=sym.size_y), rng.normal(size=sym.size_p)
    e0 = bind(bind(env, sym.y, y0), sym.p, p)
    e1 = bind(bind(env, sym.y, y1), sym.p, p)
    r = np.concatenate([evaluate(sym.hoist_ri, sym.ri_exprs, e0), evaluate(sym.hoist_rf, sym.rf_exprs, e1)])
    np.testing.assert_allclose(r, fam.r(y0, y1, p), atol=1e-14)
    ny, np_ = sym.size_y, sym.size_p
    Dri = evaluate(sym.hoist_ri, [v for row in sym.Dri_exprs for v in row], e0).reshape(-1, ny + np_)
    Drf = evaluate(sym.hoist_rf, [v for row in sym.Drf_exprs for v in row], e1).reshape(-1, ny + np_)
    Dr = np.zeros((ny + np_, 2 * ny + np_))
    ni = Dri.shape[0]
    Dr[:ni, :ny], Dr[:ni, 2 * ny:] = Dri[:, :ny], Dri[:, ny:]
    Dr[ni:, ny:2 * ny], Dr[ni:, 2 * ny:] = Drf[:, :ny], Drf[:, ny:]
    np.testing.assert_allclose(Dr, fam.Dr(y0, y1, p), atol=1e-14)


GENERAL = """
Constants = [c1 = 0.7, c2 = 1.3];
StateVariables = [x1, x2];
ControlVariables = [u1, u2];
ParameterVariables = [tf];
CostFunctional = tf*(x1^2 + c1*u1^2 + u2^2 + exp(0.1*x2));
TerminalPenalty = tf^2 + sin(x1);
DifferentialEquations = [tf*(x2 + sin(u1)), tf*(-c2*x1 + u2*sqrt(1 + x1^2) + abs(x2)*0.01)];
EqualityConstraints = [u1 - 0.5*u2 - x1];
InequalityConstraints = [u2 - 2];
StateVariableInequalityConstraints = [x2 - 3];
InitialConstraints = [x1 - 1, x2];
TerminalConstraints = [x1 + x2];
"""


def central_jacobian(fun, v, eps=1e-6):
    f0 = fun(v)
    J = np.zeros((f0.size, v.size))
    for k in range(v.size):
        d = np.zeros_like(v)
        d[k] = eps
        J[:, k] = (fun(v + d) - fun(v - d)) / (2 * eps)
    return J


def test_general_ocp_jacobians_match_finite_differences():
    sym = emitter.BvpDaeSymbolic(ocpmod.parse_ocp(GENERAL))
    assert (sym.size_y, sym.size_z, sym.size_p) == (2 + 2 + 1, 2 + 1 + 1 + 1 + 1 + 1, 1 + 2 + 1)
    env = base_env(sym)
    env[sym.alpha] = sp.Float(0.05, 30)
    rng = np.random.default_rng(7)
    ny, nz, np_ = sym.size_y, sym.size_z, sym.size_p
    v = rng.uniform(0.2, 1.0, ny + nz + np_)
    H = sym.hoist_fg
    allv = sym.y + sym.z + sym.p

    def fg(w):
        e = bind(env, allv, w)
        return np.concatenate([evaluate(H, sym.f_exprs, e), evaluate(H, sym.g_exprs, e)])
    e = bind(env, allv, v)
    Df = evaluate(H, [x for row in sym.Df_exprs for x in row], e).reshape(ny, -1)
    Dg = evaluate(H, [x for row in sym.Dg_exprs for x in row], e).reshape(nz, -1)
    J = central_jacobian(fg, v)
    np.testing.assert_allclose(np.vstack([Df, Dg]), J, rtol=1e-6, atol=1e-7)
    # the smoothed complementarity rows: mu + nu - sqrt(mu^2 + nu^2 + 2 alpha)
    g = evaluate(H, sym.g_exprs, e)
    mu, nu = v[ny + 3], v[ny + 5]
    assert abs(g[5] - (mu + nu - np.sqrt(mu ** 2 + nu ** 2 + 0.1))) < 1e-14

    yp = sym.y + sym.p
    for hz, exprs, D in ((sym.hoist_ri, sym.ri_exprs, sym.Dri_exprs), (sym.hoist_rf, sym.rf_exprs, sym.Drf_exprs)):
        w0 = rng.uniform(0.2, 1.0, ny + np_)
        fun = lambda w: evaluate(hz, exprs, bind(env, yp, w))
        Dn = evaluate(hz, [x for row in D for x in row], bind(env, yp, w0)).reshape(len(exprs), -1)
        np.testing.assert_allclose(Dn, central_jacobian(fun, w0), rtol=1e-6, atol=1e-7)


def _table(text, name):
    """Values of an emitted table `name[N] = {...};`."""
    i = text.index(name + "[")
    body = text[text.index("{", i) + 1:text.index("};", i)]
    return [float(v) for v in body.replace("\n", " ").split(",")]


def _body(text, fn):
    i = text.index("void %s(" % fn)
    return text[i:text.index("\n}\n", i)]


def test_emitted_code_is_well_formed():
    sym = emitter.BvpDaeSymbolic(ocpmod.parse_ocp(GENERAL))
    text = emitter.emit_problem(sym)
    assert text.count("{") == text.count("}")
    ny, nz, np_ = sym.size_y, sym.size_z, sym.size_p
    nj = ny + nz + np_
    for name, n in (("_f", ny), ("_g", nz)):
        for i in range(n):
            assert ("%s[%d] =" % (name, i)) in text, (name, i)
    # every entry of the stage Jacobian block [Df; Dg] is either written by
    # exactly one of the jac_part<k> functions or static: tabulated with its
    # value, and then a function of the compile-time constants only
    mask = _table(text, "jac_static_mask")
    vals = _table(text, "jac_static_val")
    assert len(mask) == len(vals) == (ny + nz) * nj
    block = [e for row in sym.Df_exprs for e in row] + [e for row in sym.Dg_exprs for e in row]
    # the parts hand their values on in the order of jac_dyn_idx
    idx = [int(v) for v in _table(text, "jac_dyn_idx")]
    assert sorted(idx) == idx and len(set(idx)) == len(idx)
    nparts = int(text.split("constexpr int JAC_PARTS = ")[1].split(";")[0])
    puts = sum(_body(text, "jac_part%d" % q).count("_sink.template put<") for q in range(nparts))
    assert puts >= -(-len(idx) // emitter.JAC_CHUNK)
    consts = {s: v for s, v in sym.constants.items() if s not in sym.instance_constants}
    for k, e in enumerate(block):
        assert (k in idx) != bool(mask[k]), k
        name, i = ("_Df", k) if k < ny * nj else ("_Dg", k - ny * nj)
        if mask[k]:
            assert not (e.free_symbols - set(consts)), (name, i)
            assert vals[k] == float(e.subs(consts)), (name, i)
    assert any(mask) and not all(mask)
    assert "K_c1" in text and "K_c2" in text
    # deterministic: the code text (and so the code-object cache key) is stable
    assert text == emitter.emit_problem(emitter.BvpDaeSymbolic(ocpmod.parse_ocp(GENERAL)))


def test_jacobian_entries_partition():
    """Dynamic entries are split into contiguous groups of about equal
    operation count: every entry in exactly one group, in order, one group
    per JAC_OPS_PER_PART operations up to JAC_MAX_PARTS."""
    x = sp.symbols("x0:40", real=True)
    heavy = [(k, sum(sp.tanh(x[j]) * x[(j + k) % 40] for j in range(40))) for k in range(400)]
    light = [(400 + k, x[k % 40] * 2) for k in range(100)]
    groups = emitter.partition_entries(light + heavy)
    flat = [k for g in groups for k, _ in g]
    assert flat == [k for k, _ in light + heavy]
    total = sum(sp.count_ops(e) + 1 for _, e in light + heavy)
    assert len(groups) == min(emitter.JAC_MAX_PARTS, -(-total // emitter.JAC_OPS_PER_PART))
    cost = [sum(sp.count_ops(e) + 1 for _, e in g) for g in groups]
    assert max(cost) < 1.3 * total / len(groups)
    assert emitter.partition_entries(light[:3]) == [light[:3]]   # a small block stays whole
