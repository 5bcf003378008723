"""OCP text format reader (the reference's problem-definition language).

The reference reads an OCP file by stripping all white space, splitting the
text into ``Keyword = value;`` statements, rewriting each statement into
Python and running the result with ``exec`` (OCP2ABVP.py:600-672, keyword map
:524-597, ``Variables`` macros :684-783).  This reader accepts the same
statements but never executes the text: expressions are tokenised and parsed
by a small recursive-descent parser straight into sympy trees.

Supported statements (same spelling as the reference):
  Constants, Variables (macro names, defined by ``name = expr;`` lines),
  StateVariables, ControlVariables, ParameterVariables, CostFunctional,
  TerminalPenalty, DifferentialEquations, EqualityConstraints,
  InequalityConstraints, StateVariableInequalityConstraints,
  InitialConstraints, TerminalConstraints, InitialTime, FinalTime, Nodes,
  Tolerance, InputFile, OutputFile, Display, MaximumMeshRefinements,
  MaximumNewtonIterations, MaximumNodes, StateEstimate, ControlEstimate,
  ParameterEstimate.
Extension (not in the reference): ``InstanceConstants = [name, ...];`` marks
constants whose values are supplied per instance at run time (batched solves).
"""
import re

import sympy as sp

_FUNCS = {
    "sin": sp.sin, "cos": sp.cos, "tan": sp.tan, "exp": sp.exp, "log": sp.log,
    "sqrt": sp.sqrt, "tanh": sp.tanh, "sinh": sp.sinh, "cosh": sp.cosh,
    "asin": sp.asin, "acos": sp.acos, "atan": sp.atan, "abs": sp.Abs, "fabs": sp.Abs,
    "atan2": sp.atan2, "pow": sp.Pow,
}


class OCPSyntaxError(ValueError):
    pass


class OCP:
    """Parsed optimal control problem; field names follow OCP2ABVP.OCP (:10-39)."""

    def __init__(self):
        self.x = []
        self.u = []
        self.p = []
        self.phi = sp.Integer(0)
        self.L = sp.Integer(0)
        self.f = []
        self.c = []
        self.d = []
        self.s = []
        self.Gamma = []
        self.Psi = []
        self.t_i = 0.0
        self.t_f = 1.0
        self.nodes = 100
        self.input_file = ""
        self.output_file = "ocp.data"
        self.tolerance = 1.0e-6
        self.maximum_nodes = 2000
        self.maximum_newton_iterations = 200
        self.maximum_mesh_refinements = 10
        self.display = 0
        self.constants = {}            # Symbol -> float (baked into the code)
        self.instance_constants = []   # Symbols read from a per-instance array
        self.state_estimate = None
        self.control_estimate = None
        self.parameter_estimate = None
        self.source = ""


# ---------------------------------------------------------------- tokenizer
_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)"
                    r"|([A-Za-z_][A-Za-z_0-9]*)|(\*\*|[-+*/^(),\[\]=])|(\"[^\"]*\"))")


def _tokenize(text):
    pos, out = 0, []
    text = text.strip()
    while pos < len(text):
        mt = _TOKEN.match(text, pos)
        if not mt or mt.end() == pos:
            raise OCPSyntaxError("unexpected character %r in %r" % (text[pos], text))
        num, name, op, string = mt.groups()
        if num is not None:
            out.append(("num", num))
        elif name is not None:
            out.append(("name", name))
        elif op is not None:
            out.append(("op", "^" if op == "**" else op))
        else:
            out.append(("str", string[1:-1]))
        pos = mt.end()
    return out


class _Parser:
    def __init__(self, tokens, resolve):
        self.toks = tokens
        self.i = 0
        self.resolve = resolve

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self, kind=None, val=None):
        tk = self.peek()
        if tk[0] is None or (kind and tk[0] != kind) or (val and tk[1] != val):
            raise OCPSyntaxError("expected %s %s, got %s" % (kind, val, tk))
        self.i += 1
        return tk

    def done(self):
        return self.i >= len(self.toks)

    def list_or_expr(self):
        if self.peek() == ("op", "["):
            self.take("op", "[")
            items = []
            if self.peek() != ("op", "]"):
                items.append(self.expr())
                while self.peek() == ("op", ","):
                    self.take()
                    items.append(self.expr())
            self.take("op", "]")
            return items
        return self.expr()

    def expr(self):
        v = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            rhs = self.term()
            v = v + rhs if op == "+" else v - rhs
        return v

    def term(self):
        v = self.unary()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            rhs = self.unary()
            v = v * rhs if op == "*" else v / rhs
        return v

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return -self.unary()
        if self.peek() == ("op", "+"):
            self.take()
            return self.unary()
        return self.power()

    def power(self):
        base = self.atom()
        if self.peek() == ("op", "^"):
            self.take()
            return base ** self.unary()          # right associative
        return base

    def atom(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            if re.fullmatch(r"\d+", val):
                return sp.Integer(int(val))
            return sp.Float(val, 17)
        if kind == "name":
            self.take()
            if self.peek() == ("op", "("):
                self.take()
                args = [self.expr()]
                while self.peek() == ("op", ","):
                    self.take()
                    args.append(self.expr())
                self.take("op", ")")
                fn = _FUNCS.get(val)
                if fn is None:
                    raise OCPSyntaxError("unknown function %s" % val)
                return fn(*args)
            return self.resolve(val)
        if (kind, val) == ("op", "("):
            self.take()
            v = self.expr()
            self.take("op", ")")
            return v
        raise OCPSyntaxError("unexpected token %s" % ((kind, val),))


# ------------------------------------------------------------------ reader
def _statements(text):
    """Split into ';'-terminated statements with comments removed (OCP2ABVP.py:617-646)."""
    body = []
    for line in text.splitlines():
        k = line.find("#")
        body.append(line if k < 0 else line[:k])
    joined = "\n".join(body)
    return [s.strip() for s in joined.split(";") if s.strip()]


def _names(rhs):
    rhs = re.sub(r"\s+", "", rhs)
    if not (rhs.startswith("[") and rhs.endswith("]")):
        raise OCPSyntaxError("expected a [list] of names, got %r" % rhs)
    inner = rhs[1:-1]
    return [n for n in inner.split(",") if n]


def parse_ocp(text):
    """Parse OCP text into an :class:`OCP` with sympy expressions."""
    ocp = OCP()
    ocp.source = text
    stmts = []
    for s in _statements(text):
        if "=" not in s:
            raise OCPSyntaxError("statement without '=': %r" % s)
        key, rhs = s.split("=", 1)
        stmts.append((re.sub(r"\s+", "", key), rhs.strip()))

    symbols = {"t": sp.Symbol("t")}
    macro_names = set()
    macro_text = {}
    macro_cache = {}
    for key, rhs in stmts:
        if key == "Variables":
            macro_names.update(_names(rhs))
    for key, rhs in stmts:
        if key in macro_names:
            macro_text[key] = rhs

    def resolve(name):
        if name in macro_names:
            if name not in macro_cache:
                if name not in macro_text:
                    raise OCPSyntaxError("macro %s has no definition" % name)
                macro_cache[name] = None
                macro_cache[name] = _parse_expr(macro_text[name])
            if macro_cache[name] is None:
                raise OCPSyntaxError("recursive macro %s" % name)
            return macro_cache[name]
        if name in ("YES",):
            return sp.Integer(1)
        if name in ("NO",):
            return sp.Integer(0)
        if name not in symbols:
            symbols[name] = sp.Symbol(name)
        return symbols[name]

    def _parse_expr(rhs):
        p = _Parser(_tokenize(rhs), resolve)
        v = p.list_or_expr()
        if not p.done():
            raise OCPSyntaxError("trailing tokens in %r" % rhs)
        return v

    def _number(rhs):
        v = _parse_expr(rhs)
        return float(v)

    def _string(rhs):
        mt = re.fullmatch(r'\s*"([^"]*)"\s*', rhs)
        if not mt:
            raise OCPSyntaxError("expected a quoted string, got %r" % rhs)
        return mt.group(1)

    def _decl(rhs):
        out = []
        for n in _names(rhs):
            symbols[n] = sp.Symbol(n)
            out.append(symbols[n])
        return out

    # Declarations first so that expressions refer to the same Symbol objects.
    for key, rhs in stmts:
        if key == "Constants":
            inner = re.sub(r"\s+", "", rhs)[1:-1]
            for item in [it for it in inner.split(",") if it]:
                name, val = item.split("=", 1)
                sym = sp.Symbol(name)
                symbols[name] = sym
                ocp.constants[sym] = float(_parse_expr(val))
        elif key == "StateVariables":
            ocp.x = _decl(rhs)
        elif key == "ControlVariables":
            ocp.u = _decl(rhs)
        elif key == "ParameterVariables":
            ocp.p = _decl(rhs)
    for key, rhs in stmts:
        if key == "InstanceConstants":
            for n in _names(rhs):
                sym = symbols.get(n)
                if sym is None or sym not in ocp.constants:
                    raise OCPSyntaxError("InstanceConstants names %s, which is not a constant" % n)
                ocp.instance_constants.append(sym)

    as_list = lambda v: v if isinstance(v, list) else [v]
    handlers = {
        "CostFunctional": lambda r: setattr(ocp, "L", _parse_expr(r)),
        "TerminalPenalty": lambda r: setattr(ocp, "phi", _parse_expr(r)),
        "DifferentialEquations": lambda r: setattr(ocp, "f", as_list(_parse_expr(r))),
        "EqualityConstraints": lambda r: setattr(ocp, "c", as_list(_parse_expr(r))),
        "InequalityConstraints": lambda r: setattr(ocp, "d", as_list(_parse_expr(r))),
        "StateVariableInequalityConstraints": lambda r: setattr(ocp, "s", as_list(_parse_expr(r))),
        "InitialConstraints": lambda r: setattr(ocp, "Gamma", as_list(_parse_expr(r))),
        "TerminalConstraints": lambda r: setattr(ocp, "Psi", as_list(_parse_expr(r))),
        "InitialTime": lambda r: setattr(ocp, "t_i", _number(r)),
        "FinalTime": lambda r: setattr(ocp, "t_f", _number(r)),
        "Nodes": lambda r: setattr(ocp, "nodes", int(_number(r))),
        "Tolerance": lambda r: setattr(ocp, "tolerance", _number(r)),
        "InputFile": lambda r: setattr(ocp, "input_file", _string(r)),
        "OutputFile": lambda r: setattr(ocp, "output_file", _string(r)),
        "Display": lambda r: setattr(ocp, "display", int(_number(r))),
        "MaximumMeshRefinements": lambda r: setattr(ocp, "maximum_mesh_refinements", int(_number(r))),
        "MaximumNewtonIterations": lambda r: setattr(ocp, "maximum_newton_iterations", int(_number(r))),
        "MaximumNodes": lambda r: setattr(ocp, "maximum_nodes", int(_number(r))),
        "StateEstimate": lambda r: setattr(ocp, "state_estimate", as_list(_parse_expr(r))),
        "ControlEstimate": lambda r: setattr(ocp, "control_estimate", as_list(_parse_expr(r))),
        "ParameterEstimate": lambda r: setattr(ocp, "parameter_estimate", as_list(_parse_expr(r))),
    }
    declared = {"Constants", "StateVariables", "ControlVariables", "ParameterVariables",
                "Variables", "InstanceConstants"}
    for key, rhs in stmts:
        if key in declared or key in macro_names:
            continue
        h = handlers.get(key)
        if h is None:
            raise OCPSyntaxError("unknown statement %s" % key)
        h(rhs)
    return ocp


def read_ocp(path):
    with open(path) as f:
        return parse_ocp(f.read())
