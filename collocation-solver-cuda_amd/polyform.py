"""Sparse-polynomial form of the BVP-DAE of an OCP, for large problems.

The symbolic path (emitter.BvpDaeSymbolic, sympy) writes every callback as
straight-line code.  For the n_x = 64 tanh family that is about 5e5 terms in
the costate Jacobian alone: sympy needs tens of minutes to form it and the
device compiler far longer.  This module forms the same BVP-DAE
(OCP2ABVP.py:41-182: the Hamiltonian, f = [H_lambda; -H_x; H_p],
g = [H_u; c; smoothed complementarity], the boundary conditions) in a sparse
polynomial algebra instead:

 * every nonlinear function application F(arg) (tanh, exp, sqrt, ...) is a
   "node", an atom of its own whose value the device computes once per
   stage; the argument is itself a polynomial;
 * every expression is a polynomial over the atoms (base variables y, z, p,
   the instance constants, alpha, and the nodes), stored as
   {monomial: coefficient} with monomial = ((atom, power), ...); the
   problem's compile-time constants are folded into the coefficients;
 * derivatives follow the chain rule through the nodes
   (d F(a) / dv = F'(a) da/dv with F' again a polynomial in nodes, e.g.
   tanh' = 1 - tanh^2), one gradient pass per expression.

compile_program() turns a list of output polynomials into the tables the
device evaluator walks (csrc/kernels/colloc_tables.h): the monomials, the
nodes by level, and the polynomials as sliced ELL slabs (64 rows per slab,
rows sorted by term count, a slab's terms stored term-major so that a wave
reads them coalesced).  eval_program() is the same evaluation in NumPy.
"""
import math

import numpy as np
import sympy as sp

# node kinds (the device evaluator's switch, colloc_tables.h: node_value)
KINDS = {"tanh": 0, "exp": 1, "sin": 2, "cos": 3, "tan": 4, "sinh": 5, "cosh": 6, "sqrt": 7, "log": 8,
         "asin": 9, "acos": 10, "atan": 11, "abs": 12, "sign": 13, "pow": 14}
_SYMPY_KIND = {sp.tanh: "tanh", sp.exp: "exp", sp.sin: "sin", sp.cos: "cos", sp.tan: "tan", sp.sinh: "sinh",
               sp.cosh: "cosh", sp.log: "log", sp.asin: "asin", sp.acos: "acos", sp.atan: "atan",
               sp.Abs: "abs", sp.sign: "sign"}


class Unsupported(ValueError):
    """The expression uses a construct the polynomial form does not cover."""


# ------------------------------------------------------------ monomials
def mono_mul(a, b):
    if not a:
        return b
    if not b:
        return a
    d = dict(a)
    for k, p in b:
        q = d.get(k, 0) + p
        if q:
            d[k] = q
        else:
            del d[k]
    return tuple(sorted(d.items()))


def padd_into(acc, b, s=1.0):
    """acc += s * b (in place); exact zeros are dropped."""
    for m, c in b.items():
        v = acc.get(m, 0.0) + s * c
        if v == 0.0:
            acc.pop(m, None)
        else:
            acc[m] = v
    return acc


def pmul(a, b):
    out = {}
    for ma, ca in a.items():
        for mb, cb in b.items():
            m = mono_mul(ma, mb)
            v = out.get(m, 0.0) + ca * cb
            if v == 0.0:
                out.pop(m, None)
            else:
                out[m] = v
    return out


def pconst(c):
    return {(): float(c)} if c != 0 else {}


def patom(a, k=1):
    return {((a, k),): 1.0}


def pscale(a, s):
    return {m: c * s for m, c in a.items()} if s != 0 else {}


class Engine:
    """Atoms 0 .. nb-1 are base variables; atom nb + k is node k."""

    def __init__(self, nb):
        self.nb = nb
        self.nodes = []      # (kind, arg poly, r)
        self._key = {}
        self._ngrad = {}

    def node(self, kind, arg, r=0.0):
        if not arg:
            raise Unsupported("node of a zero argument")
        if len(arg) == 1 and () in arg:   # a constant argument: a number
            return pconst(_apply(kind, arg[()], r))
        key = (kind, tuple(sorted(arg.items())), float(r))
        a = self._key.get(key)
        if a is None:
            a = self.nb + len(self.nodes)
            self.nodes.append((kind, dict(arg), float(r)))
            self._key[key] = a
        return patom(a)

    def ppow(self, b, n):
        """b ** n for an integer n (a node of b for negative n unless b is
        a single monomial)."""
        if n == 0:
            return pconst(1.0)
        if n > 0:
            out = pconst(1.0)
            for _ in range(n):
                out = pmul(out, b)
            return out
        if len(b) == 1:
            (m, c), = b.items()
            return {tuple((a, k * n) for a, k in m): c ** n}
        return self.node("pow", b, float(n))

    # -- from sympy ----------------------------------------------------------
    def from_sympy(self, e, consts, atom_of):
        """Polynomial of a sympy expression: compile-time constants (consts:
        Symbol -> value) folded in, other symbols mapped by atom_of."""
        cache = {}

        def conv(x):
            r = cache.get(x)
            if r is not None:
                return r
            r = conv1(x)
            cache[x] = r
            return r

        def conv1(x):
            if x.is_Number:
                return pconst(float(x))
            if x.is_Symbol:
                if x in consts:
                    return pconst(consts[x])
                a = atom_of.get(x)
                if a is None:
                    raise Unsupported("symbol %s" % x)
                return patom(a)
            if x.is_Add:
                out = {}
                for t in x.args:
                    padd_into(out, conv(t))
                return out
            if x.is_Mul:
                out = pconst(1.0)
                for t in x.args:
                    out = pmul(out, conv(t))
                return out
            if x.is_Pow:
                b, ex = x.base, x.exp
                if ex.free_symbols:
                    if ex.free_symbols <= set(consts):
                        ex = sp.Float(float(ex.subs(consts)))
                    else:
                        raise Unsupported("symbolic exponent in %s" % x)
                pb = conv(b)
                if ex.is_Integer:
                    return self.ppow(pb, int(ex))
                r = float(ex)
                if r == int(r):
                    return self.ppow(pb, int(r))
                if r == 0.5:
                    return self.node("sqrt", pb)
                return self.node("pow", pb, r)
            if isinstance(x, sp.Function):
                kind = _SYMPY_KIND.get(x.func)
                if kind is None or len(x.args) != 1:
                    raise Unsupported("function %s" % x.func)
                return self.node(kind, conv(x.args[0]))
            raise Unsupported("expression %s" % type(x).__name__)

        return conv(sp.sympify(e))

    # -- differentiation -------------------------------------------------------
    def dF(self, k):
        """dF/darg of node k as a polynomial."""
        kind, arg, r = self.nodes[k]
        h = patom(self.nb + k)
        one = pconst(1.0)
        if kind == "tanh":
            return padd_into(dict(one), self.ppow(h, 2), -1.0)
        if kind == "exp":
            return h
        if kind == "sin":
            return self.node("cos", arg)
        if kind == "cos":
            return pscale(self.node("sin", arg), -1.0)
        if kind == "tan":
            return padd_into(dict(one), self.ppow(h, 2))
        if kind == "sinh":
            return self.node("cosh", arg)
        if kind == "cosh":
            return self.node("sinh", arg)
        if kind == "sqrt":
            return {((self.nb + k, -1),): 0.5}
        if kind == "log":
            return self.node("pow", arg, -1.0)
        if kind in ("asin", "acos"):
            w = self.node("pow", padd_into(dict(one), pmul(arg, arg), -1.0), -0.5)
            return w if kind == "asin" else pscale(w, -1.0)
        if kind == "atan":
            return self.node("pow", padd_into(dict(one), pmul(arg, arg)), -1.0)
        if kind == "abs":
            return self.node("sign", arg)
        if kind == "sign":
            return {}
        if kind == "pow":
            if r - 1.0 == 0.0:
                return pconst(r)
            if r - 1.0 == 1.0:
                return pscale(arg, r)
            return pscale(self.node("pow", arg, r - 1.0), r)
        raise Unsupported(kind)

    def node_grad(self, a):
        g = self._ngrad.get(a)
        if g is None:
            k = a - self.nb
            _, arg, _ = self.nodes[k]
            ga = self.grad(arg)
            df = self.dF(k)
            g = {v: pmul(df, dv) for v, dv in ga.items()} if df else {}
            g = {v: p for v, p in g.items() if p}
            self._ngrad[a] = g
        return g

    def grad(self, p):
        """{base atom: d p / d atom} (the chain rule through the nodes)."""
        out = {}
        for m, c in p.items():
            for i, (a, k) in enumerate(m):
                rest = m[:i] + m[i + 1:] if k == 1 else m[:i] + ((a, k - 1),) + m[i + 1:]
                base = {rest: c * k}
                if a < self.nb:
                    padd_into(out.setdefault(a, {}), base)
                else:
                    for v, dv in self.node_grad(a).items():
                        padd_into(out.setdefault(v, {}), pmul(base, dv))
        return {v: q for v, q in out.items() if q}


def _apply(kind, x, r=0.0):
    f = {"tanh": math.tanh, "exp": math.exp, "sin": math.sin, "cos": math.cos, "tan": math.tan,
         "sinh": math.sinh, "cosh": math.cosh, "sqrt": math.sqrt, "log": math.log, "asin": math.asin,
         "acos": math.acos, "atan": math.atan, "abs": abs,
         "sign": lambda v: (v > 0) - (v < 0), "pow": lambda v: v ** r}[kind]
    return float(f(x))


def node_value(kind, x, r):
    """NumPy restatement of the device's node evaluation."""
    with np.errstate(all="ignore"):
        if kind == "sign":
            return np.sign(x)
        if kind == "pow":
            return np.power(x, r)
        return {"tanh": np.tanh, "exp": np.exp, "sin": np.sin, "cos": np.cos, "tan": np.tan, "sinh": np.sinh,
                "cosh": np.cosh, "sqrt": np.sqrt, "log": np.log, "asin": np.arcsin, "acos": np.arccos,
                "atan": np.arctan, "abs": np.abs}[kind](x)


# ------------------------------------------------------------ BVP-DAE form
class BvpDaePoly:
    """The BVP-DAE of an OCP in polynomial form (OCP2ABVP.py:41-182).

    Atom layout (also the device's base vector): y (ny) | z (nz) | p (np) |
    instance constants (nc) | alpha.  Variable order, Jacobian column order
    ([y | z | p] for Df/Dg, [y0 | y1 | p] for Dr) and the boundary multiplier
    convention are the reference's, as in emitter.BvpDaeSymbolic."""

    def __init__(self, ocp):
        o = ocp
        nx, npar = len(o.x), len(o.p)
        nu, nc, nd, ns = len(o.u), len(o.c), len(o.d), len(o.s)
        ngam, npsi = len(o.Gamma), len(o.Psi)
        self.ocp = o
        self.size_y = 2 * nx + npar
        self.size_z = nu + nc + 2 * nd + 2 * ns
        self.size_p = npar + ngam + npsi
        self.size_inequality, self.size_sv_inequality = nd, ns
        ny, nz, np_ = self.size_y, self.size_z, self.size_p
        ic = list(o.instance_constants)
        self.n_inst = len(ic)
        consts = {s: v for s, v in o.constants.items() if s not in ic}
        y0, z0, p0, c0 = 0, ny, ny + nz, ny + nz + np_
        self.alpha_atom = c0 + len(ic)
        self.nb = self.alpha_atom + 1
        # atoms of the OCP's own symbols
        atom_of = {}
        for i, s in enumerate(o.x):
            atom_of[s] = y0 + i
        for i, s in enumerate(o.p):
            atom_of[s] = p0 + i
        for i, s in enumerate(o.u):
            atom_of[s] = z0 + i
        for i, s in enumerate(ic):
            atom_of[s] = c0 + i
        lam = [y0 + nx + i for i in range(nx)]
        gam = [y0 + 2 * nx + i for i in range(npar)]
        theta = [z0 + nu + i for i in range(nc)]
        mu = [z0 + nu + nc + i for i in range(nd)]
        xi = [z0 + nu + nc + nd + i for i in range(ns)]
        nuv = [z0 + nu + nc + nd + ns + i for i in range(nd)]
        sig = [z0 + nu + nc + nd + ns + nd + i for i in range(ns)]
        kap_i = [p0 + npar + i for i in range(ngam)]
        kap_f = [p0 + npar + ngam + i for i in range(npsi)]
        al = patom(self.alpha_atom)

        E = Engine(self.nb)
        conv = lambda e: E.from_sympy(e, consts, atom_of)
        L = conv(o.L)
        f = [conv(e) for e in o.f]
        if len(f) != nx:
            raise ValueError("DifferentialEquations has %d entries for %d states" % (len(f), nx))
        cc = [conv(e) for e in o.c]
        dd = [conv(e) for e in o.d]
        ss = [conv(e) for e in o.s]
        # Hamiltonian (OCP2ABVP.py:95-111)
        H = dict(L)
        for i in range(nx):
            padd_into(H, pmul(patom(lam[i]), f[i]))
        if nd > 0 or ns > 0:
            for i in range(nu):
                padd_into(H, pmul(al, patom(z0 + i, 2)), 0.5)
        for i in range(nc):
            padd_into(H, pmul(patom(theta[i]), cc[i]))
        for i in range(nd):
            padd_into(H, pmul(patom(mu[i]), dd[i]))
        for i in range(ns):
            padd_into(H, pmul(patom(xi[i]), ss[i]))
        gH = E.grad(H)
        d = lambda a: gH.get(a, {})
        # f = [H_lambda; -H_x; H_p] (OCP2ABVP.py:118-124)
        fe = [d(lam[i]) for i in range(nx)]
        fe += [pscale(d(y0 + i), -1.0) for i in range(nx)]
        fe += [d(p0 + i) for i in range(npar)]
        # g (OCP2ABVP.py:130-144)
        ge = [{} for _ in range(nz)]
        for i in range(nu):
            ge[i] = d(z0 + i)
        for i in range(nc):
            ge[nu + i] = d(theta[i])
        for i in range(nd):
            ge[nu + nc + i] = padd_into(padd_into(dict(d(mu[i])), patom(nuv[i])), pmul(al, patom(mu[i])), -1.0)
            sq = E.node("sqrt", padd_into(padd_into(patom(mu[i], 2), patom(nuv[i], 2)), al, 2.0))
            ge[nu + nc + nd + ns + i] = padd_into(padd_into(patom(mu[i]), patom(nuv[i])), sq, -1.0)
        for i in range(ns):
            ge[nu + nc + nd + i] = padd_into(padd_into(dict(d(xi[i])), patom(sig[i])), pmul(al, patom(xi[i])), -1.0)
            sq = E.node("sqrt", padd_into(padd_into(patom(xi[i], 2), patom(sig[i], 2)), al, 2.0))
            ge[nu + nc + nd + ns + nd + i] = padd_into(padd_into(patom(xi[i]), patom(sig[i])), sq, -1.0)
        self.engine_fg = E
        self.f_polys, self.g_polys = fe, ge
        ncol = ny + nz + np_
        self.Df_polys = [self._row(E.grad(e), ncol) for e in fe]
        self.Dg_polys = [self._row(E.grad(e), ncol) for e in ge]

        # boundary conditions (OCP2ABVP.py:146-180): the initial rows bind y
        # to y(t_i), the final rows to y(t_f), each with its own nodes
        Ei, Ef = Engine(self.nb), Engine(self.nb)
        Gam = [Ei.from_sympy(e, consts, atom_of) for e in o.Gamma]
        Psi = [Ef.from_sympy(e, consts, atom_of) for e in o.Psi]
        phi = Ef.from_sympy(o.phi, consts, atom_of)
        gG = [Ei.grad(e) for e in Gam]
        gP = [Ef.grad(e) for e in Psi]
        gphi = Ef.grad(phi)
        ri = list(Gam)
        for i in range(nx):
            e = patom(lam[i])
            for j in range(ngam):
                padd_into(e, pmul(gG[j].get(y0 + i, {}), patom(kap_i[j])))
            ri.append(e)
        for i in range(npar):
            ri.append(patom(gam[i]))
        rf = list(Psi)
        for i in range(nx):
            e = padd_into(patom(lam[i]), gphi.get(y0 + i, {}), -1.0)
            for j in range(npsi):
                padd_into(e, pmul(gP[j].get(y0 + i, {}), patom(kap_f[j])), -1.0)
            rf.append(e)
        for i in range(npar):
            e = padd_into(patom(gam[i]), gphi.get(p0 + i, {}))
            for j in range(npsi):
                padd_into(e, pmul(gP[j].get(p0 + i, {}), patom(kap_f[j])))
            rf.append(e)
        if len(ri) + len(rf) != ny + np_:
            raise ValueError("boundary conditions: %d rows for %d unknowns (y %d + p %d)"
                             % (len(ri) + len(rf), ny + np_, ny, np_))
        self.engine_ri, self.engine_rf = Ei, Ef
        self.ri_polys, self.rf_polys = ri, rf
        yp = list(range(ny)) + [p0 + j for j in range(np_)]
        self.Dri_polys = [[g.get(a, {}) for a in yp] for g in (Ei.grad(e) for e in ri)]
        self.Drf_polys = [[g.get(a, {}) for a in yp] for g in (Ef.grad(e) for e in rf)]

    def _row(self, g, ncol):
        return [g.get(a, {}) for a in range(ncol)]

    @property
    def tolerance(self):
        return self.ocp.tolerance

    def static_value(self, p):
        """The value of a polynomial without atoms (a constant), else None."""
        if not p:
            return 0.0
        if len(p) == 1 and () in p:
            return p[()]
        return None

    def active_p_columns(self):
        base = self.size_y + self.size_z
        return [j for j in range(self.size_p)
                if any(row[base + j] for row in self.Df_polys) or any(row[base + j] for row in self.Dg_polys)]

    def term_count(self):
        n = sum(len(p) for p in self.f_polys + self.g_polys)
        n += sum(len(p) for row in self.Df_polys + self.Dg_polys for p in row)
        return n


# ------------------------------------------------------------ tables
SLAB = 64   # rows per ELL slab: one wave


class Program:
    """Device tables of a set of output polynomials over one engine's atoms."""

    def __init__(self, engine, outputs, dest):
        """outputs: list of polynomials; dest[q]: where output q goes (an
        index the kernel interprets)."""
        E = engine
        nb = E.nb
        # the nodes reachable from the outputs (their arguments included)
        need = set()
        stack = [a for p in outputs for m in p for a, _ in m if a >= nb]
        while stack:
            a = stack.pop()
            if a in need:
                continue
            need.add(a)
            _, arg, _ = E.nodes[a - nb]
            stack.extend(b for m in arg for b, _ in m if b >= nb)
        level = {}

        def atom_level(a):
            if a < nb:
                return 0
            lv = level.get(a)
            if lv is None:
                _, arg, _ = E.nodes[a - nb]
                lv = 1 + max([atom_level(b) for m in arg for b, _ in m] or [0])
                level[a] = lv
            return lv

        nodes = sorted(need, key=lambda a: (atom_level(a), a))
        self.nlev = 1 + max([atom_level(a) for a in nodes] or [0])
        # atom numbering on the device: base atoms, then the needed nodes
        self.atom_index = {a: a for a in range(nb)}
        for k, a in enumerate(nodes):
            self.atom_index[a] = nb + k
        self.natoms = nb + len(nodes)
        mono_level = lambda m: max([atom_level(a) for a, _ in m] or [0])
        monos = {(): 0}
        allpolys = [E.nodes[a - nb][1] for a in nodes] + list(outputs)
        for p in allpolys:
            for m in p:
                if m not in monos:
                    monos[m] = len(monos)
        mlist = sorted(monos, key=lambda m: (mono_level(m), monos[m]))
        self.mono_index = {m: k for k, m in enumerate(mlist)}
        self.nmono = len(mlist)
        # evaluation order: the monomials of level L, then the nodes of level
        # L + 1 (their arguments use monomials of levels <= L), L = 0, 1, ...
        self.mono_lev_off = [0] * (self.nlev + 1)
        for m in mlist:
            self.mono_lev_off[mono_level(m) + 1] += 1
        for L in range(self.nlev):
            self.mono_lev_off[L + 1] += self.mono_lev_off[L]
        # nodes of level L (>= 1) are [node_lev_off[L - 1], node_lev_off[L])
        self.node_lev_off = [0] + [sum(1 for a in nodes if atom_level(a) <= L) for L in range(1, self.nlev)]
        self.node_lev_off.append(len(nodes))
        self.mono_fac_off = [0]
        self.mono_fac_atom, self.mono_fac_pow = [], []
        for m in mlist:
            for a, k in m:
                self.mono_fac_atom.append(self.atom_index[a])
                self.mono_fac_pow.append(k)
            self.mono_fac_off.append(len(self.mono_fac_atom))
        self.node_kind = [KINDS[E.nodes[a - nb][0]] for a in nodes]
        self.node_r = [E.nodes[a - nb][2] for a in nodes]
        self.node_names = [E.nodes[a - nb][0] for a in nodes]
        # polynomial rows: the nodes' arguments (row k = node k), then the outputs
        self.nnode = len(nodes)
        self.nout = len(outputs)
        rows = [E.nodes[a - nb][1] for a in nodes] + list(outputs)
        self.row_dest = list(range(len(nodes))) + list(dest)
        # the node rows and the output rows are sliced separately (a slab
        # never mixes them), each sorted by term count, longest first
        def slabs(first, count):
            order = sorted(range(first, first + count), key=lambda r: -len(rows[r]))
            out = []
            for s in range(0, count, SLAB):
                out.append(order[s:s + SLAB])
            return out
        node_slabs = []
        # node rows are sliced per level so that a level's slabs come together
        for L in range(1, self.nlev):
            a0, a1 = self.node_lev_off[L - 1], self.node_lev_off[L]
            node_slabs.append(slabs(a0, a1 - a0))
        out_slabs = slabs(len(nodes), len(outputs))
        self.node_slab_off = [0]
        all_slabs = []
        for sl in node_slabs:
            all_slabs += sl
            self.node_slab_off.append(len(all_slabs))
        self.out_slab_first = len(all_slabs)
        all_slabs += out_slabs
        self.nslab = len(all_slabs)
        self.slab_term_off = [0]
        self.slab_rows = []
        coef, mono = [], []
        for sl in all_slabs:
            k = max(len(rows[r]) for r in sl) if sl else 0
            items = [sorted(rows[r].items(), key=lambda mc: self.mono_index[mc[0]]) for r in sl]
            for t in range(k):
                for q in range(SLAB):
                    if q < len(sl) and t < len(items[q]):
                        m, c = items[q][t]
                        coef.append(c)
                        mono.append(self.mono_index[m])
                    else:
                        coef.append(0.0)
                        mono.append(0)
            self.slab_term_off.append(len(coef))
            self.slab_rows.append([self.row_dest[r] if r >= len(nodes) else r for r in sl] +
                                  [-1] * (SLAB - len(sl)))
        self.coef = np.array(coef, dtype=np.float64)
        self.mono = np.array(mono, dtype=np.int32)
        self.terms = int(sum(len(p) for p in rows))

    # NumPy restatement of the device evaluator (colloc_tables.h)
    def evaluate(self, base):
        """base: (G, nb) values of the base atoms for G points; returns
        (G, max dest + 1) output values by destination and the (G, natoms)
        atom values."""
        base = np.atleast_2d(np.asarray(base, dtype=np.float64))
        G = base.shape[0]
        ext = np.zeros((G, self.natoms))
        ext[:, :base.shape[1]] = base
        mv = np.ones((G, self.nmono))

        def monos(L):
            for k in range(self.mono_lev_off[L], self.mono_lev_off[L + 1]):
                v = np.ones(G)
                for q in range(self.mono_fac_off[k], self.mono_fac_off[k + 1]):
                    v = v * ext[:, self.mono_fac_atom[q]] ** float(self.mono_fac_pow[q])
                mv[:, k] = v

        def slab_rows(s):
            t0, t1 = self.slab_term_off[s], self.slab_term_off[s + 1]
            c = self.coef[t0:t1].reshape(-1, SLAB)
            m = self.mono[t0:t1].reshape(-1, SLAB)
            return (mv[:, m] * c[None]).sum(axis=1)   # (G, SLAB)

        nb = ext.shape[1] - self.nnode
        for L in range(self.nlev):
            monos(L)
            if L + 1 < self.nlev:
                for s in range(self.node_slab_off[L], self.node_slab_off[L + 1]):
                    v = slab_rows(s)
                    for q, r in enumerate(self.slab_rows[s]):
                        if r >= 0:
                            ext[:, nb + r] = node_value(self.node_names[r], v[:, q], self.node_r[r])
        nd = max(self.row_dest[self.nnode:], default=-1) + 1
        out = np.zeros((G, nd))
        for s in range(self.out_slab_first, self.nslab):
            v = slab_rows(s)
            for q, r in enumerate(self.slab_rows[s]):
                if r >= 0:
                    out[:, r] = v[:, q]
        return out, ext
