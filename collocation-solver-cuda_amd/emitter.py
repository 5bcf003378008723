"""OCP -> BVP-DAE conversion and HIP code emission.

Restates OCP2ABVP.make_abvp (OCP2ABVP.py:41-182): the Hamiltonian
H = L + lambda'f (+ alpha u'u/2 when inequalities exist) + theta'c + mu'd + xi's,
f = [H_lambda; -H_x; H_p], g = [H_u; c; smoothed complementarity], and the
boundary conditions r = [Gamma; lambda + Gamma_x' kappa_i; gamma;
Psi; lambda - phi_x - Psi_x' kappa_f; gamma + phi_p + Psi_p' kappa_f].
Variable order, Jacobian column order ([y | z | p] for Df/Dg, [y0 | y1 | p]
for Dr) and the boundary multiplier convention are the reference's.

Differences from the reference's emitter (OCP2ABVP.py:184-468), by design:
 * nonlinear function applications (tanh, sin, sqrt, ...) are hoisted into
   their own symbols before differentiation and derivatives are taken by the
   chain rule over those symbols, so that sympy never differentiates a
   nested argument twice and the emitted code evaluates each such term once;
 * constants stay symbolic until code generation (emitted as constexpr);
 * the callbacks are HIP __device__ functions (r and Dr too), and Jacobians
   are written densely (structural zeros included) so callers need not clear
   them.
"""
import hashlib

import numpy as np
import sympy as sp
from sympy.printing.c import C99CodePrinter

from .ocp import OCP

# -------------------------------------------------------------- hoisting
class Hoister:
    """Holds hoisted nonlinear subexpressions h_k = F(arg_k) and differentiates
    expressions that contain them by the chain rule."""

    def __init__(self, prefix="_h"):
        self.prefix = prefix
        self.defs = []          # list of (symbol, definition) in creation order
        self.defn = {}          # symbol -> definition
        self.key = {}           # definition -> symbol
        self.deps = {}          # symbol -> frozenset of base symbols it depends on
        self._dcache = {}

    # -- structure -------------------------------------------------------
    def _intern(self, d):
        s = self.key.get(d)
        if s is not None:
            return s
        s = sp.Symbol("%s%d" % (self.prefix, len(self.defs)), real=True)
        self.key[d] = s
        self.defn[s] = d
        self.defs.append((s, d))
        self.deps[s] = self.base_deps(d)
        return s

    def base_deps(self, e):
        out = set()
        for s in e.free_symbols:
            if s in self.deps:
                out |= self.deps[s]
            else:
                out.add(s)
        return frozenset(out)

    def hoist(self, e):
        e = sp.sympify(e)
        if e.is_Atom:
            return e
        if isinstance(e, sp.Function):
            args = [self.hoist(a) for a in e.args]
            if all(a.is_number for a in args):
                return e.func(*args)
            return self._intern(e.func(*args))
        if isinstance(e, sp.Pow):
            b = self.hoist(e.base)
            x = self.hoist(e.exp)
            if x.is_Integer or not (b.free_symbols or x.free_symbols):
                return b ** x
            return self._intern(b ** x)
        return e.func(*[self.hoist(a) for a in e.args])

    # -- differentiation ---------------------------------------------------
    def deriv(self, e, v):
        """d e / d v where e may contain hoisted symbols."""
        if v not in self.base_deps(e):
            return sp.Integer(0)
        res = sp.diff(e, v)
        for h in e.free_symbols:
            if h in self.defn and v in self.deps[h]:
                dh = self.dh(h, v)
                if dh != 0:
                    res = res + sp.diff(e, h) * dh
        return res

    def dh(self, h, v):
        ck = (h, v)
        if ck in self._dcache:
            return self._dcache[ck]
        d = self.defn[h]
        if isinstance(d, sp.Pow):
            b, x = d.base, d.exp
            if v not in self.base_deps(x):
                out = x * h / b * self.deriv(b, v)
            else:
                lb = self._intern(sp.log(b))
                out = h * (self.deriv(x, v) * lb + x * self.deriv(b, v) / b)
        else:
            out = sp.Integer(0)
            dummies = [sp.Dummy("a%d" % i, real=True) for i in range(len(d.args))]
            fd = d.func(*dummies)
            for k, a in enumerate(d.args):
                da = self.deriv(a, v)
                if da == 0:
                    continue
                fp = sp.diff(fd, dummies[k])
                if fp.has(sp.Derivative) or fp.has(sp.Subs):
                    raise ValueError("cannot differentiate %s" % d.func)
                fp = fp.subs(dict(zip(dummies, d.args)))
                out = out + self.hoist(fp) * da
        self._dcache[ck] = out
        return out

    def needed(self, exprs):
        """Hoisted symbols reachable from exprs, in definition order."""
        need = set()
        stack = [s for e in exprs for s in e.free_symbols if s in self.defn]
        while stack:
            s = stack.pop()
            if s in need:
                continue
            need.add(s)
            stack.extend(t for t in self.defn[s].free_symbols if t in self.defn)
        return [(s, d) for s, d in self.defs if s in need]


# ------------------------------------------------------------ BVP-DAE form
class BvpDaeSymbolic:
    """The BVP-DAE of an OCP: symbols, expressions and problem settings."""

    def __init__(self, ocp):
        o = ocp
        self.ocp = o
        nx, npar = len(o.x), len(o.p)
        nu, nc, nd, ns = len(o.u), len(o.c), len(o.d), len(o.s)
        ngam, npsi = len(o.Gamma), len(o.Psi)
        R = dict(real=True)
        lam = [sp.Symbol("_lambda%d" % (i + 1), **R) for i in range(nx)]
        gam = [sp.Symbol("_gamma%d" % (i + 1), **R) for i in range(npar)]
        theta = [sp.Symbol("_theta%d" % (i + 1), **R) for i in range(nc)]
        mu = [sp.Symbol("_mu%d" % (i + 1), **R) for i in range(nd)]
        nuv = [sp.Symbol("_nu%d" % (i + 1), **R) for i in range(nd)]
        xi = [sp.Symbol("_xi%d" % (i + 1), **R) for i in range(ns)]
        sig = [sp.Symbol("_sigma%d" % (i + 1), **R) for i in range(ns)]
        kap_i = [sp.Symbol("_kappa_i%d" % (i + 1), **R) for i in range(ngam)]
        kap_f = [sp.Symbol("_kappa_f%d" % (i + 1), **R) for i in range(npsi)]
        alpha = sp.Symbol("_alpha", **R)
        # user symbols are made real so that |.| differentiates to sign(.)
        realmap = {}
        for s in list(o.x) + list(o.u) + list(o.p) + list(o.constants):
            realmap[s] = sp.Symbol(s.name, real=True)
        sub = lambda e: sp.sympify(e).xreplace(realmap)
        xs = [realmap[s] for s in o.x]
        us = [realmap[s] for s in o.u]
        ps = [realmap[s] for s in o.p]
        self.constants = {realmap[s]: v for s, v in o.constants.items()}
        self.instance_constants = [realmap[s] for s in o.instance_constants]

        # variable vectors (OCP2ABVP.py:54-90)
        self.y = xs + lam + gam
        self.z = us + theta + mu + xi + nuv + sig
        self.p = ps + kap_i + kap_f
        self.alpha = alpha
        self.size_y, self.size_z, self.size_p = len(self.y), len(self.z), len(self.p)
        self.size_inequality, self.size_sv_inequality = nd, ns

        hz = Hoister("_h")
        L, f = hz.hoist(sub(o.L)), [hz.hoist(sub(e)) for e in o.f]
        cc = [hz.hoist(sub(e)) for e in o.c]
        dd = [hz.hoist(sub(e)) for e in o.d]
        ss = [hz.hoist(sub(e)) for e in o.s]
        if len(f) != nx:
            raise ValueError("DifferentialEquations has %d entries for %d states" % (len(f), nx))
        # Hamiltonian (OCP2ABVP.py:95-111)
        H = L + sum(lam[i] * f[i] for i in range(nx))
        if nd > 0 or ns > 0:
            H = H + sum(alpha * u * u / 2 for u in us)
        H = H + sum(theta[i] * cc[i] for i in range(nc))
        H = H + sum(mu[i] * dd[i] for i in range(nd))
        H = H + sum(xi[i] * ss[i] for i in range(ns))
        # f = [H_lambda; -H_x; H_p] (OCP2ABVP.py:118-124)
        fe = [hz.deriv(H, lam[i]) for i in range(nx)]
        fe += [-hz.deriv(H, xs[i]) for i in range(nx)]
        fe += [hz.deriv(H, ps[i]) for i in range(npar)]
        # g (OCP2ABVP.py:130-144)
        ge = [sp.Integer(0)] * self.size_z
        for i in range(nu):
            ge[i] = hz.deriv(H, us[i])
        for i in range(nc):
            ge[nu + i] = hz.deriv(H, theta[i])
        for i in range(nd):
            ge[nu + nc + i] = hz.deriv(H, mu[i]) + nuv[i] - alpha * mu[i]
            ge[nu + nc + nd + ns + i] = mu[i] + nuv[i] - hz.hoist(sp.sqrt(mu[i] ** 2 + nuv[i] ** 2 + 2 * alpha))
        for i in range(ns):
            ge[nu + nc + nd + i] = hz.deriv(H, xi[i]) + sig[i] - alpha * xi[i]
            ge[nu + nc + nd + ns + nd + i] = xi[i] + sig[i] - hz.hoist(sp.sqrt(xi[i] ** 2 + sig[i] ** 2 + 2 * alpha))
        self.hoist_fg = hz
        self.f_exprs = fe
        self.g_exprs = ge
        yz = self.y + self.z + self.p
        self.Df_exprs = [[hz.deriv(e, v) for v in yz] for e in fe]
        self.Dg_exprs = [[hz.deriv(e, v) for v in yz] for e in ge]

        # boundary conditions (OCP2ABVP.py:146-180); separate hoisting because
        # the initial rows bind y to y(t_i) and the final rows to y(t_f)
        hi, hf = Hoister("_hi"), Hoister("_hf")
        Gam = [hi.hoist(sub(e)) for e in o.Gamma]
        Psi = [hf.hoist(sub(e)) for e in o.Psi]
        phi = hf.hoist(sub(o.phi))
        ri = list(Gam)
        for i in range(nx):
            ri.append(lam[i] + sum(hi.deriv(Gam[j], xs[i]) * kap_i[j] for j in range(ngam)))
        for i in range(npar):
            ri.append(gam[i])
        rf = list(Psi)
        for i in range(nx):
            rf.append(lam[i] - hf.deriv(phi, xs[i]) - sum(hf.deriv(Psi[j], xs[i]) * kap_f[j] for j in range(npsi)))
        for i in range(npar):
            rf.append(gam[i] + hf.deriv(phi, ps[i]) + sum(hf.deriv(Psi[j], ps[i]) * kap_f[j] for j in range(npsi)))
        if len(ri) + len(rf) != self.size_y + self.size_p:
            raise ValueError("boundary conditions: %d rows for %d unknowns (y %d + p %d)"
                             % (len(ri) + len(rf), self.size_y + self.size_p, self.size_y, self.size_p))
        self.hoist_ri, self.hoist_rf = hi, hf
        self.ri_exprs, self.rf_exprs = ri, rf
        yp = self.y + self.p
        self.Dri_exprs = [[hi.deriv(e, v) for v in yp] for e in ri]
        self.Drf_exprs = [[hf.deriv(e, v) for v in yp] for e in rf]

    # problem settings (the generated BvpDae.__init__, OCP2ABVP.py:329-447)
    @property
    def tolerance(self):
        return self.ocp.tolerance


# ---------------------------------------------------------------- printing
class _Printer(C99CodePrinter):
    def __init__(self, names):
        super().__init__()
        self._names = names

    def _print_Symbol(self, e):
        n = self._names.get(e)
        return n if n is not None else super()._print_Symbol(e)

    def _print_Pow(self, e):
        b, x = e.base, e.exp
        if x.is_Integer:
            k = int(x)
            if 1 <= k <= 4:
                s = "(" + self._print(b) + ")"
                return "(" + "*".join([s] * k) + ")"
            if -4 <= k <= -1:
                s = "(" + self._print(b) + ")"
                return "(1.0/(" + "*".join([s] * (-k)) + "))"
        if x == sp.Rational(1, 2):
            return "sqrt(%s)" % self._print(b)
        if x == sp.Rational(-1, 2):
            return "(1.0/sqrt(%s))" % self._print(b)
        return "pow(%s, %s)" % (self._print(b), self._print(x))

    def _print_Float(self, e):
        return repr(float(e))

    def _print_Rational(self, e):
        return "(%r)" % (float(e.p) / float(e.q))

    def _print_Integer(self, e):
        return "%d.0" % int(e)

    def _print_sign(self, e):
        a = self._print(e.args[0])
        return "((%s) > 0.0 ? 1.0 : ((%s) < 0.0 ? -1.0 : 0.0))" % (a, a)

    def _print_DiracDelta(self, e):
        # derivative of sign/abs: zero wherever it is defined
        return "0.0"

    def _print_Heaviside(self, e):
        a = self._print(e.args[0])
        return "((%s) > 0.0 ? 1.0 : ((%s) < 0.0 ? 0.0 : 0.5))" % (a, a)


class _TypedPrinter(_Printer):
    """Prints expressions evaluated in the C type `tc` (a template parameter
    such as TO): literals and compile-time constants are cast to it, so a
    float instantiation computes in float throughout."""

    def __init__(self, names, tc):
        super().__init__(names)
        self._tc = tc

    def _c(self, lit):
        return "((%s)%s)" % (self._tc, lit)

    def _print_Pow(self, e):
        b, x = e.base, e.exp
        if x.is_Integer and -4 <= int(x) <= -1:
            s = "(" + self._print(b) + ")"
            return "(%s/(%s))" % (self._c("1"), "*".join([s] * (-int(x))))
        if x == sp.Rational(-1, 2):
            return "(%s/sqrt(%s))" % (self._c("1"), self._print(b))
        return super()._print_Pow(e)

    def _print_Float(self, e):
        return self._c(repr(float(e)))

    def _print_Rational(self, e):
        return self._c(repr(float(e.p) / float(e.q)))

    def _print_Integer(self, e):
        return self._c("%d" % int(e))

    def _print_sign(self, e):
        a = self._print(e.args[0])
        return "((%s) > %s ? %s : ((%s) < %s ? %s : %s))" % (a, self._c(0), self._c(1), a, self._c(0), self._c(-1),
                                                             self._c(0))

    def _print_DiracDelta(self, e):
        return self._c(0)

    def _print_Heaviside(self, e):
        a = self._print(e.args[0])
        return "((%s) > %s ? %s : ((%s) < %s ? %s : %s))" % (a, self._c(0), self._c(1), a, self._c(0), self._c(0),
                                                             self._c("0.5"))


def _cname(sym):
    """C name of a symbol: generated symbols (leading '_') get prefix 'w', user
    symbols 'u_', constants are printed as 'K_<name>' (distinct namespaces)."""
    n = sym.name
    return "w" + n if n.startswith("_") else "u_" + n


def _emit_block(lines, exprs_out, hoister, bindings, constants, inst_consts, indent="  ", ctype=None, chunk=None):
    """Append C statements computing exprs_out = [(target, expr), ...]; with
    ctype (a C type name) every intermediate is computed in that type,
    else in double.  chunk = (C, flush): the targets are ignored, the values
    go to _v[0..C) in groups of C and each group is handed on by the
    statement flush(index of its first value, group size)."""
    td = ctype or "double"
    exprs = [e for _, e in exprs_out]
    hdefs = hoister.needed(exprs) if hoister is not None else []
    used = set()
    for e in exprs:
        used |= e.free_symbols
    for _, d in hdefs:
        used |= d.free_symbols
    names = {}
    cast = (lambda src: "(%s)(%s)" % (ctype, src)) if ctype else (lambda src: src)
    for s, src in bindings:
        names[s] = _cname(s)
        if s in used:
            lines.append("%sconst %s %s = %s;" % (indent, td, names[s], cast(src)))
    for k, s in enumerate(inst_consts):
        names[s] = _cname(s)
        if s in used:
            lines.append("%sconst %s %s = %s;" % (indent, td, names[s], cast("_c[%d]" % k)))
    for s in constants:
        if s not in names:
            names[s] = ("((%s)K_%s)" % (ctype, s.name)) if ctype else "K_" + s.name
    pr = _TypedPrinter(names, ctype) if ctype else _Printer(names)
    for s, d in hdefs:
        names[s] = _cname(s)
        lines.append("%sconst %s %s = %s;" % (indent, td, names[s], pr.doprint(d)))
    nonzero = [(tgt, e) for tgt, e in exprs_out if e != 0]
    if nonzero:
        reps, red = sp.cse([e for _, e in nonzero], symbols=sp.numbered_symbols("_t", real=True),
                           optimizations=None, order="none")
        for s, d in reps:
            names[s] = _cname(s)
            lines.append("%sconst %s %s = %s;" % (indent, td, names[s], pr.doprint(d)))
        if chunk is None:
            for (tgt, _), e in zip(nonzero, red):
                lines.append("%s%s = %s;" % (indent, tgt, pr.doprint(e)))
        else:
            C, flush = chunk
            if len(nonzero) != len(exprs_out):
                raise ValueError("chunked output needs non-zero expressions")
            lines.append("%s%s _v[%d];" % (indent, td, C))
            for q, e in enumerate(red):
                lines.append("%s_v[%d] = %s;" % (indent, q % C, pr.doprint(e)))
                if q % C == C - 1 or q == len(red) - 1:
                    lines.append(indent + flush(q - q % C, q % C + 1))
    for tgt, e in exprs_out:
        if chunk is not None:
            break
        if e == 0:
            lines.append("%s%s = %s;" % (indent, tgt, "(%s)0" % ctype if ctype else "0.0"))


def emit_problem(sym):
    """HIP source text of namespace prob { f, g, Df, Dg, r, Dr } for a BVP-DAE."""
    ny, nz, np_ = sym.size_y, sym.size_z, sym.size_p
    nj = ny + nz + np_
    ic = sym.instance_constants
    consts = [s for s in sym.constants if s not in ic]
    out = ["// Generated BVP-DAE callbacks (restating the _abvp_* contract of OCP2ABVP.py:203-327).",
           "namespace prob {"]
    for s in consts:
        out.append("constexpr double K_%s = %r;" % (s.name, float(sym.constants[s])))
    yb = [(s, "_y[%d]" % k) for k, s in enumerate(sym.y)]
    zb = [(s, "_z[%d]" % k) for k, s in enumerate(sym.z)]
    pb = [(s, "_p[%d]" % k) for k, s in enumerate(sym.p)]
    ab = [(sym.alpha, "_alpha")]
    sig_fg = "const double* __restrict__ _y, const double* __restrict__ _z, const double* __restrict__ _p, " \
             "const double* __restrict__ _c"
    H = sym.hoist_fg

    def fn(name, extra, target, rows, binds, hoister, out_type=False, chunk=None):
        # out_type: a template on the output type (the Jacobians are stored,
        # and evaluated, in float64, or float32 for the FP32 direction);
        # chunk: the values go to a sink object (template SINK) in groups
        if chunk is not None:
            lines = ["template <class TO, class SINK>",
                     "__device__ __forceinline__ void %s(%s%s, SINK& _sink) {" % (name, sig_fg, extra)]
        elif out_type:
            lines = ["template <class TO>",
                     "__device__ __forceinline__ void %s(%s%s, TO* __restrict__ %s) {" % (name, sig_fg, extra, target)]
        else:
            lines = ["__device__ __forceinline__ void %s(%s%s, double* __restrict__ %s) {" % (name, sig_fg, extra,
                                                                                            target)]
        lines.append("  (void)_y; (void)_z; (void)_p; (void)_c;")
        if "_alpha" in extra:
            lines.append("  (void)_alpha;")
        # the Jacobians compute in their stored type (float for the FP32
        # direction: an inexact-Newton direction), the residuals in double
        _emit_block(lines, rows, hoister, binds, consts, ic, ctype="TO" if out_type else None, chunk=chunk)
        lines.append("}")
        out.extend(lines)

    fn("f", "", "_f", [("_f[%d]" % i, e) for i, e in enumerate(sym.f_exprs)], yb + zb + pb, H)
    fn("g", ", double _alpha", "_g", [("_g[%d]" % i, e) for i, e in enumerate(sym.g_exprs)], yb + zb + pb + ab, H)
    # The stage Jacobian block [Df; Dg] splits into static entries (zeros and
    # expressions of the problem's compile-time constants only) and dynamic
    # ones (depending on the state, the parameters, the instance constants or
    # alpha).  jac_part<k> (JAC_PARTS groups) write the dynamic entries at
    # every Newton iteration; the static ones are tabulated (jac_static_val where
    # jac_static_mask is set) and written once per Jacobian buffer.
    cvals = {s: sym.constants[s] for s in consts}
    cset = set(consts)
    block = [e for row in sym.Df_exprs for e in row] + [e for row in sym.Dg_exprs for e in row]
    static = [not (e.free_symbols - cset) for e in block]
    # The dynamic entries are split into JAC_PARTS groups of about equal
    # operation count, each its own function (jac_part<k>), so that several
    # threads share a stage and none holds the whole block's intermediates.
    # Each part hands its values on in groups of JAC_CHUNK to a sink of the
    # kernel's (SINK::put<n>(first, values)), value q of the part being entry
    # jac_dyn_idx[offset of the part + q] of the stage block, so the kernel
    # can store a group for many stages at once.
    dyn = [(k, e) for k, (e, st) in enumerate(zip(block, static)) if not st]
    parts = partition_entries(dyn)
    offs, idx = [], []
    for q, group in enumerate(parts):
        offs.append(len(idx))
        idx += [k for k, _ in group]
        base = offs[-1]
        fn("jac_part%d" % q, ", double _alpha", "_J", [("_J[%d]" % k, e) for k, e in group], yb + zb + pb + ab, H,
           out_type=True, chunk=(JAC_CHUNK, lambda first, n, base=base:
                                 "_sink.template put<%d>(%d, _v);" % (n, base + first)))
    out.append("constexpr int JAC_PARTS = %d;   // threads per stage of the Jacobian kernel" % len(parts))
    out.append("constexpr int JAC_CHUNK = %d;   // values per sink call" % JAC_CHUNK)
    out.append(_table("__device__ const int jac_dyn_idx", ["%d" % k for k in idx] or ["0"]))
    out.append("template <class TO, class SINK>\n__device__ __forceinline__ void jac_dyn_part(int part, %s, "
               "double _alpha, SINK& _sink) {\n  switch (part) {" % sig_fg)
    for q in range(len(parts)):
        out.append("    case %d: jac_part%d<TO>(_y, _z, _p, _c, _alpha, _sink); break;" % (q, q))
    out.append("    default: break;\n  }\n}")
    out.append(_table("__device__ const unsigned char jac_static_mask", ["%d" % (1 if st else 0) for st in static]))
    out.append(_table("__device__ const double jac_static_val",
                      [repr(float(e.subs(cvals))) if st else "0.0" for e, st in zip(block, static)]))
    out.append("constexpr int JAC_DYNAMIC = %d;   // dynamic entries per stage block" % static.count(False))
    # r and Dr: initial rows see y(t_i) = _y0, final rows y(t_f) = _y1
    sig_r = "const double* __restrict__ _y0, const double* __restrict__ _y1, const double* __restrict__ _p, " \
            "const double* __restrict__ _c"
    nri = len(sym.ri_exprs)
    y0b = [(s, "_y0[%d]" % k) for k, s in enumerate(sym.y)]
    y1b = [(s, "_y1[%d]" % k) for k, s in enumerate(sym.y)]
    out.append("__device__ __forceinline__ void r(%s, double* __restrict__ _r) {" % sig_r)
    out.append("  (void)_y0; (void)_y1; (void)_p; (void)_c;")
    out.append("  {")
    _emit_block(out, [("_r[%d]" % i, e) for i, e in enumerate(sym.ri_exprs)], sym.hoist_ri, y0b + pb, consts, ic, "    ")
    out.append("  }")
    out.append("  {")
    _emit_block(out, [("_r[%d]" % (nri + i), e) for i, e in enumerate(sym.rf_exprs)], sym.hoist_rf, y1b + pb,
                consts, ic, "    ")
    out.append("  }")
    out.append("}")
    ndr = 2 * ny + np_
    out.append("__device__ __forceinline__ void Dr(%s, double* __restrict__ _Dr) {" % sig_r)
    out.append("  (void)_y0; (void)_y1; (void)_p; (void)_c;")
    rows = []
    for i, row in enumerate(sym.Dri_exprs):
        for j in range(ndr):
            col = j if j < ny else (None if j < 2 * ny else j - ny)
            rows.append(("_Dr[%d]" % (i * ndr + j), row[col] if col is not None else sp.Integer(0)))
    out.append("  {")
    _emit_block(out, rows, sym.hoist_ri, y0b + pb, consts, ic, "    ")
    out.append("  }")
    rows = []
    for i, row in enumerate(sym.Drf_exprs):
        for j in range(ndr):
            col = None if j < ny else (j - ny if j < 2 * ny else j - ny)
            rows.append(("_Dr[%d]" % ((nri + i) * ndr + j), row[col] if col is not None else sp.Integer(0)))
    out.append("  {")
    _emit_block(out, rows, sym.hoist_rf, y1b + pb, consts, ic, "    ")
    out.append("  }")
    out.append("}")
    out.append("}  // namespace prob")
    return "\n".join(out) + "\n"


JAC_OPS_PER_PART = 20000
JAC_MAX_PARTS = 8
JAC_CHUNK = 16


def partition_entries(dyn):
    """Split the dynamic Jacobian entries [(index, expr)] into contiguous
    groups of about equal operation count (sympy's count_ops), one group per
    JAC_OPS_PER_PART operations, at most JAC_MAX_PARTS groups."""
    if not dyn:
        return [[]]
    cost = [sp.count_ops(e) + 1 for _, e in dyn]
    total = sum(cost)
    n = max(1, min(JAC_MAX_PARTS, -(-total // JAC_OPS_PER_PART), len(dyn)))
    groups, cur, acc = [], [], 0.0
    for (k, e), c in zip(dyn, cost):
        cur.append((k, e))
        acc += c
        if acc >= total * (len(groups) + 1) / n and len(groups) < n - 1:
            groups.append(cur)
            cur = []
    groups.append(cur)
    return groups


def _table(decl, items, per_line=16):
    rows = [", ".join(items[k:k + per_line]) for k in range(0, len(items), per_line)]
    return "%s[%d] = {\n  %s};" % (decl, len(items), ",\n  ".join(rows))


def _carray(name, vals):
    return "__constant__ double %s[%d] = {%s};" % (name, len(vals), ", ".join(repr(float(v)) for v in vals))


def choose_waves(ny, nz, np_, m, npa=None):
    """Waves per workgroup of the register-resident eliminations.

    Condensation: the stage system (m-1) ny x ((m-1) ny + ny + npa + 1), npa
    being the parameters the stage functions depend on, held by the blocked
    Gauss-Jordan solver (lane = row, columns spread over the waves): eight
    waves (256 VGPRs each) when a thread's share is at most 96 doubles,
    otherwise sixteen (128 VGPRs) with at most 42.  Boundary system
    ((2 ny + np) x (2 ny + np + 1)): eight waves up to 96 doubles per
    thread, sixteen up to 48."""
    def rq(n):
        return (n + 63) // 64

    ns = (m - 1) * ny
    ncol = ns + ny + (np_ if npa is None else npa) + 1
    if rq(ns) * -(-ncol // 8) <= 96:
        cw = 8
    elif rq(ns) * -(-ncol // 16) <= 42:
        cw = 16
    else:
        raise ValueError("a %d x %d condensation does not fit in one workgroup's registers" % (ns, ncol))
    nf = 2 * ny + np_
    if rq(nf) * -(-(nf + 1) // 8) <= 96:
        fw = 8
    elif rq(nf) * -(-(nf + 1) // 16) <= 48:
        fw = 16
    else:
        raise ValueError("a %d x %d boundary system does not fit in one workgroup's registers" % (nf, nf + 1))
    return cw, fw


def active_p_columns(sym):
    """Parameters that the stage functions f, g depend on.

    A parameter column of [f_p; g_p] that is identically zero contributes an
    identically zero column to the condensed system (its H block is zero);
    the condensation skips such columns.  Returns the sorted list of the
    parameter indices whose Jacobian column is not structurally zero."""
    base = sym.size_y + sym.size_z
    out = []
    for j in range(sym.size_p):
        col = [row[base + j] for row in sym.Df_exprs] + [row[base + j] for row in sym.Dg_exprs]
        if any(sp.sympify(e) != 0 for e in col):
            out.append(j)
    return out


LARGE_NY = 64   # the register-resident condensation and cyclic-reduction kernels hold ny <= 64


def _unit_tail(ny, nz, np_, nc, m, cw, fw, pact, tables):
    parts = ["struct Cfg {\n  static constexpr int NY = %d, NZ = %d, NP = %d, NC = %d, M = %d, CW = %d, FW = %d;\n"
             "  // parameters the stage functions depend on (the other H columns are zero)\n"
             "  static constexpr int NPA = %d;\n  static constexpr int PACT[%d] = {%s};\n};"
             % (ny, nz, np_, nc, m, cw, fw, len(pact), max(len(pact), 1),
                ", ".join(str(j) for j in pact) if pact else "0")]
    parts.append(_carray("c_rkA", tables.A.ravel()))
    parts.append(_carray("c_rkb", tables.b))
    parts.append(_carray("c_rkc", tables.c))
    parts.append(_carray("c_gL", tables.gL.ravel()))
    parts.append(_carray("c_gI", tables.gI.ravel()))
    parts.append(_carray("c_gw", tables.gw))
    parts.append(_carray("c_L0", tables.L0))
    parts.append(_carray("c_L1", tables.L1))
    parts.append('extern "C" __device__ int colloc_meta[8] = {%d, %d, %d, %d, %d, %d, %d, %d};'
                 % (ny, nz, np_, m, nc, cw, fw, len(pact)))
    parts.append('#include "colloc_kernels.h"')
    return parts


def sym_static_mask(sym):
    """1 for the stage-block entries of [Df; Dg] that depend only on the
    problem's compile-time constants (emit_problem's static entries)."""
    cset = set(c for c in sym.constants if c not in sym.instance_constants)
    block = [e for row in sym.Df_exprs for e in row] + [e for row in sym.Dg_exprs for e in row]
    return [1 if not (sp.sympify(e).free_symbols - cset) else 0 for e in block]


def jac_gemm_section(sym, poly):
    """The matrix-core Jacobian's tables for a straight-line unit (the jac
    program of the problem's polynomial form, its coefficient matrix), or
    None when the polynomial form's state-dependent entries are not exactly
    the straight-line form's (then the straight-line Jacobian kernel is the
    only one)."""
    if poly is None:
        return None
    progs, mask, _ = table_programs(poly)
    if list(mask) != sym_static_mask(sym) or progs["jac"].nout == 0:
        return None
    return "\n".join(["namespace prob {",
                      "constexpr int TB_NB = %d;   // base atoms: y | z | p | instance constants | alpha" % poly.nb,
                      emit_program("jac", progs["jac"]), emit_jac_gemm(progs["jac"]),
                      "}  // namespace prob"])


def build_unit(sym, m, tables, poly=None):
    """Complete hiprtc translation unit: callbacks, sizes, tables, kernels.

    tables: object with lobatto A, b, c and gauss t, w plus the Lagrange basis
    L(tau) and its integral I(tau) at the Gauss points and L(0), L(1).
    poly: the problem's polyform.BvpDaePoly, if it has one: its Jacobian
    program then also evaluates the float32 stage Jacobians on matrix cores
    (COLLOC_JAC_GEMM)."""
    ny, nz, np_ = sym.size_y, sym.size_z, sym.size_p
    nc = len(sym.instance_constants)
    pact = active_p_columns(sym)
    cw, fw = choose_waves(ny, nz, np_, m, npa=len(pact))
    gemm = jac_gemm_section(sym, poly)
    parts = ["#define COLLOC_TABLE_FORM 0", "#define COLLOC_LARGE 0",
             "#define COLLOC_JAC_GEMM %d" % (1 if gemm else 0), '#include "colloc_common.h"',
             emit_problem(sym)]
    if gemm:
        parts.append(gemm)
    parts += _unit_tail(ny, nz, np_, nc, m, cw, fw, pact, tables)
    return "\n".join(parts) + "\n"


def unit_hash(text):
    return hashlib.sha256(text.encode()).hexdigest()[:24]


# ------------------------------------------------- polynomial (table) form
# Problems whose straight-line callbacks would be too large for the device
# compiler (and for the instruction cache) are emitted as tables of their
# polynomial form (polyform.py), walked by the evaluator of colloc_tables.h.
TABLE_TERMS = 100000   # Jacobian + f + g terms beyond which the table form is used


def _int_table(decl, vals, per_line=32):
    vals = [int(v) for v in vals] or [0]
    return _table(decl, [str(v) for v in vals], per_line)


def _dbl_table(decl, vals, per_line=8):
    vals = [float(v) for v in vals] or [0.0]
    return _table(decl, [repr(v) for v in vals], per_line)


def emit_program(name, prog):
    """C tables of one polyform.Program and a function returning its TabProg."""
    out = [
        _int_table("__device__ const int %s_mono_fac_off" % name, prog.mono_fac_off),
        _int_table("__device__ const int %s_mono_fac_atom" % name, prog.mono_fac_atom),
        _int_table("__device__ const signed char %s_mono_fac_pow" % name, prog.mono_fac_pow),
        _int_table("__device__ const int %s_mono_lev_off" % name, prog.mono_lev_off),
        _int_table("__device__ const int %s_node_slab_off" % name, prog.node_slab_off),
        _int_table("__device__ const unsigned char %s_node_kind" % name, prog.node_kind),
        _dbl_table("__device__ const double %s_node_r" % name, prog.node_r),
        _int_table("__device__ const int %s_slab_term_off" % name, prog.slab_term_off),
        _int_table("__device__ const int %s_slab_rows" % name, [v for row in prog.slab_rows for v in row]),
        _dbl_table("__device__ const double %s_coef" % name, prog.coef),
        _int_table("__device__ const int %s_mono" % name, prog.mono),
    ]
    nb = prog.natoms - prog.nnode
    out.append("__device__ __forceinline__ colloc::TabProg %s_prog() {\n"
               "  return colloc::TabProg{%d, %d, %d, %d, %d, %d, %s_mono_fac_off, %s_mono_fac_atom, "
               "%s_mono_fac_pow, %s_mono_lev_off, %s_node_slab_off, %s_node_kind, %s_node_r, %s_slab_term_off, "
               "%s_slab_rows, %s_coef, %s_mono};\n}"
               % ((name, nb, prog.natoms, prog.nmono, prog.nlev, prog.nslab, prog.out_slab_first) + (name,) * 11))
    out.append("constexpr int %s_NATOMS = %d, %s_NMONO = %d;" % (name.upper(), prog.natoms, name.upper(),
                                                                prog.nmono))
    return "\n".join(out)


JG_POINTS = 32   # stages per workgroup of the matrix-core Jacobian


def jac_gemm_matrix(prog):
    """The state-dependent stage-Jacobian entries of a jac program as a
    dense product: entry = sum_k C[row, k] mv[k] over the program's
    monomials mv.  Returns (C (rows padded to 16, columns to 4), dest per
    row (-1: padding)), rows ordered by destination."""
    import numpy as np
    from . import polyform
    dests = sorted({r for s in range(prog.out_slab_first, prog.nslab) for r in prog.slab_rows[s] if r >= 0})
    index = {r: i for i, r in enumerate(dests)}
    nr = (len(dests) + 15) // 16 * 16
    nk = (prog.nmono + 3) // 4 * 4
    C = np.zeros((max(nr, 16), max(nk, 4)))
    for s in range(prog.out_slab_first, prog.nslab):
        t0, t1 = prog.slab_term_off[s], prog.slab_term_off[s + 1]
        c = np.asarray(prog.coef[t0:t1]).reshape(-1, polyform.SLAB)
        m = np.asarray(prog.mono[t0:t1]).reshape(-1, polyform.SLAB)
        for q, r in enumerate(prog.slab_rows[s]):
            if r >= 0:
                np.add.at(C[index[r]], m[:, q], c[:, q])
    return C, dests + [-1] * (C.shape[0] - len(dests))


def emit_jac_gemm(prog):
    """Tables of the matrix-core evaluation of the float32 stage Jacobians
    (colloc_tables.h jacobian_gemm): the nonzero 16 x 4 blocks of C, row
    tile by row tile, as v_mfma_f32_16x16x4f32 A operands (lane l: row
    16 rt + l % 16, column 4 ks + l / 16), each block's k-step, and the
    destination of each row in the stage block."""
    import numpy as np
    C, dest = jac_gemm_matrix(prog)
    rt, ks = C.shape[0] // 16, C.shape[1] // 4
    lane = np.arange(64)
    off, kstep, A = [0], [], []
    for r in range(rt):
        for k in range(ks):
            blk = C[16 * r:16 * r + 16, 4 * k:4 * k + 4]
            if np.any(blk != 0):
                kstep.append(k)
                A.append(C[16 * r + (lane % 16), 4 * k + lane // 16])
        off.append(len(kstep))
    A = np.array(A) if A else np.zeros((1, 64))
    out = ["constexpr int JG_RT = %d, JG_KS = %d, JG_POINTS = %d;   // row tiles, k-steps, stages per workgroup"
           % (rt, ks, JG_POINTS),
           "constexpr int JG_KMAX = %d;   // nonzero blocks of a row tile, at most" %
           max(off[r + 1] - off[r] for r in range(rt)),
           _int_table("__device__ const int jg_blk_off", off),
           _int_table("__device__ const int jg_blk_ks", kstep),
           _table("__device__ const float jg_coef", [repr(float(np.float32(v))) + "f" for v in A.ravel()], 8),
           _int_table("__device__ const int jg_dest", dest)]
    return "\n".join(out)


def table_programs(P):
    """The programs of a BvpDaePoly: fg (f then g: destinations
    0 .. ny+nz-1), jac (the state-dependent entries of the stage block
    [Df; Dg]: destination = entry index), ri / rf (the initial / final
    boundary rows: destination = row of r), dri / drf (their rows of Dr:
    destination = entry of Dr, columns [y0 | y1 | p]).  Returns (programs,
    static mask, static values) of the stage block."""
    from . import polyform
    ny, nz, np_ = P.size_y, P.size_z, P.size_p
    fg = polyform.Program(P.engine_fg, P.f_polys + P.g_polys, list(range(ny + nz)))
    block = [e for row in P.Df_polys for e in row] + [e for row in P.Dg_polys for e in row]
    sval = [P.static_value(e) for e in block]
    dyn = [k for k, v in enumerate(sval) if v is None]
    jac = polyform.Program(P.engine_fg, [block[k] for k in dyn], dyn)
    ndr = 2 * ny + np_
    nri = len(P.ri_polys)
    ri = polyform.Program(P.engine_ri, list(P.ri_polys), list(range(nri)))
    rf = polyform.Program(P.engine_rf, list(P.rf_polys), [nri + i for i in range(len(P.rf_polys))])

    def dr_program(engine, rows, first, initial):
        outs, dest = [], []
        for i, row in enumerate(rows):
            for j in range(ndr):
                if j < ny:
                    e = row[j] if initial else {}
                elif j < 2 * ny:
                    e = {} if initial else row[j - ny]
                else:
                    e = row[ny + j - 2 * ny]
                outs.append(e)
                dest.append((first + i) * ndr + j)
        return polyform.Program(engine, outs, dest)

    dri = dr_program(P.engine_ri, P.Dri_polys, 0, True)
    drf = dr_program(P.engine_rf, P.Drf_polys, nri, False)
    mask = [0 if v is None else 1 for v in sval]
    vals = [0.0 if v is None else v for v in sval]
    return {"fg": fg, "jac": jac, "ri": ri, "rf": rf, "dri": dri, "drf": drf}, mask, vals


def emit_problem_tables(P):
    """HIP source of namespace prob for the polynomial (table) form."""
    progs, mask, vals = table_programs(P)
    out = ["// Generated BVP-DAE callbacks in polynomial table form (the _abvp_* contract of "
           "OCP2ABVP.py:203-327, evaluated by colloc_tables.h).",
           "namespace prob {",
           "constexpr bool TABLE_FORM = true;",
           "constexpr int TB_NB = %d;   // base atoms: y | z | p | instance constants | alpha" % P.nb]
    for name, prog in progs.items():
        out.append(emit_program(name, prog))
    out.append(_int_table("__device__ const unsigned char jac_static_mask", mask))
    out.append(_dbl_table("__device__ const double jac_static_val", vals))
    out.append("constexpr int JAC_DYNAMIC = %d;   // dynamic entries per stage block" % mask.count(0))
    out.append("constexpr int JAC_PARTS = 1;")
    out.append("}  // namespace prob")
    return "\n".join(out) + "\n"


def build_unit_tables(P, m, tables):
    """hiprtc translation unit of a problem in the polynomial (table) form."""
    ny, nz, np_ = P.size_y, P.size_z, P.size_p
    nc = P.n_inst
    pact = P.active_p_columns()
    large = ny > LARGE_NY
    cw, fw = (0, 0) if large else choose_waves(ny, nz, np_, m, npa=len(pact))
    parts = ["#define COLLOC_TABLE_FORM 1", "#define COLLOC_LARGE %d" % (1 if large else 0), "#define COLLOC_JAC_GEMM 0",
             '#include "colloc_common.h"', emit_problem_tables(P)]
    parts += _unit_tail(ny, nz, np_, nc, m, cw, fw, pact, tables)
    return "\n".join(parts) + "\n"
