"""Synthetic OCP instances of the benchmark families, written as OCP text.

Family (one of):
  x' = A x + B u + s * tanh(C x),  cost int (q x'x + c_u u'u) dt,
  x(0) = x0, x(T) = 0,
with A, B, C drawn i.i.d. N(0, 1/n_x) (variance 1/n_x) from
numpy.random.default_rng(seed) in the order A, B, C, then x0 ~ U[-1, 1]^n_x.
The linear-quadratic family is s = 0 (no C drawn).  The reference has no
generator; instances are emitted in its OCP text grammar (OCP2ABVP.py:524-597)
so that they go through the same parser and emitter as any user problem.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class FamilyInstance:
    nx: int
    nu: int
    A: np.ndarray
    B: np.ndarray
    C: np.ndarray = None
    x0: np.ndarray = None
    s: float = 0.5
    q: float = 1.0
    c_u: float = 1.0
    T: float = 1.0
    nodes: int = 129
    tolerance: float = 1e-8
    instance_x0: bool = False
    homotopy: bool = False
    extra: dict = field(default_factory=dict)

    def instance_constants(self):
        """Names of the constants supplied per instance at run time, in the
        order of the InstanceConstants statement."""
        names = []
        if self.homotopy and self.C is not None and self.s != 0.0:
            names.append("S")
        if self.instance_x0:
            names += ["X0_%d" % (i + 1) for i in range(self.nx)]
        return names

    def constant_values(self, s=None):
        """Values of instance_constants(); s overrides the tanh scale."""
        vals = []
        for n in self.instance_constants():
            if n == "S":
                vals.append(self.s if s is None else s)
            else:
                vals.append(self.x0[int(n.split("_")[1]) - 1])
        return vals

    def homotopy_schedule(self):
        """Continuation path for the plain all-ones initial guess: solve with
        the tanh scale at 0 (the linear-quadratic problem, one Newton step),
        then at its value.  Returns the intermediate instance-constant vectors."""
        if "S" not in self.instance_constants():
            return []
        return [self.constant_values(s=0.0)]

    def ocp_text(self):
        nx, nu = self.nx, self.nu
        r = lambda v: repr(float(v))
        consts = []
        for i in range(nx):
            for j in range(nx):
                consts.append("A_%d_%d = %s" % (i + 1, j + 1, r(self.A[i, j])))
            for j in range(nu):
                consts.append("B_%d_%d = %s" % (i + 1, j + 1, r(self.B[i, j])))
            if self.C is not None:
                for j in range(nx):
                    consts.append("C_%d_%d = %s" % (i + 1, j + 1, r(self.C[i, j])))
        for i in range(nx):
            consts.append("X0_%d = %s" % (i + 1, r(self.x0[i])))
        tanh = self.C is not None and self.s != 0.0
        if tanh:
            consts.append("S = %s" % r(self.s))
        xs = ["x%d" % (i + 1) for i in range(nx)]
        us = ["u%d" % (j + 1) for j in range(nu)]
        f = []
        for i in range(nx):
            terms = ["A_%d_%d*x%d" % (i + 1, j + 1, j + 1) for j in range(nx)]
            terms += ["B_%d_%d*u%d" % (i + 1, j + 1, j + 1) for j in range(nu)]
            e = " + ".join(terms)
            if tanh:
                arg = " + ".join("C_%d_%d*x%d" % (i + 1, j + 1, j + 1) for j in range(nx))
                e += " + S*tanh(%s)" % arg
            f.append(e)
        cost = "%s*(%s) + %s*(%s)" % (r(self.q), " + ".join("%s*%s" % (x, x) for x in xs),
                                      r(self.c_u), " + ".join("%s*%s" % (u, u) for u in us))
        lines = [
            "# synthetic %s instance: nx=%d nu=%d" % ("tanh" if self.C is not None else "LQ", nx, nu),
            "Constants = [%s];" % ", ".join(consts),
            "StateVariables = [%s];" % ", ".join(xs),
            "ControlVariables = [%s];" % ", ".join(us),
            "InitialConstraints = [%s];" % ", ".join("x%d - X0_%d" % (i + 1, i + 1) for i in range(nx)),
            "TerminalConstraints = [%s];" % ", ".join(xs),
            "CostFunctional = %s;" % cost,
            "DifferentialEquations = [%s];" % ", ".join(f),
            "InitialTime = 0.0;",
            "FinalTime = %s;" % r(self.T),
            "Nodes = %d;" % self.nodes,
            "Tolerance = %s;" % r(self.tolerance),
            "MaximumNewtonIterations = 200;",
            "MaximumNodes = %d;" % max(2000, 4 * self.nodes),
        ]
        if self.instance_constants():
            lines.append("InstanceConstants = [%s];" % ", ".join(self.instance_constants()))
        return "\n".join(lines) + "\n"


def draw_family(nx, nu, seed, tanh=True):
    """A, B (, C) and x0 from default_rng(seed) in the order A, B, C, x0."""
    rng = np.random.default_rng(seed)
    sd = np.sqrt(1.0 / nx)
    A = rng.normal(0.0, sd, (nx, nx))
    B = rng.normal(0.0, sd, (nx, nu))
    C = rng.normal(0.0, sd, (nx, nx)) if tanh else None
    x0 = rng.uniform(-1.0, 1.0, nx)
    return A, B, C, x0


def config(index, nint=None, nx=None, nu=None):
    """Instances of the benchmark configurations (0-based, as listed in the
    measurement plan).  nint / nx / nu override sizes for reduced parity runs.

      0: LQ, n_x=4, n_u=2, seed 0, T=1, cost x'x + u'u, Nint=128, m=4
      1: tanh, n_x=20, n_u=10, seed 1, T=2, cost x'x + 0.1 u'u, Nint=4096, m=5
      2: tanh, n_x=64, n_u=32, seed 2, T=2, cost x'x + 0.1 u'u, Nint=4096, m=4
    Returns (FamilyInstance, m).
    """
    if index == 0:
        nx_, nu_, seed, T, n_, m, tanh, c_u = 4, 2, 0, 1.0, 128, 4, False, 1.0
    elif index == 1:
        nx_, nu_, seed, T, n_, m, tanh, c_u = 20, 10, 1, 2.0, 4096, 5, True, 0.1
    elif index == 2:
        nx_, nu_, seed, T, n_, m, tanh, c_u = 64, 32, 2, 2.0, 4096, 4, True, 0.1
    else:
        raise ValueError("configuration %d is not a single-instance problem" % index)
    nx_ = nx if nx is not None else nx_
    nu_ = nu if nu is not None else nu_
    n_ = nint if nint is not None else n_
    A, B, C, x0 = draw_family(nx_, nu_, seed, tanh)
    inst = FamilyInstance(nx=nx_, nu=nu_, A=A, B=B, C=C, x0=x0, s=0.5 if tanh else 0.0, q=1.0, c_u=c_u,
                          T=T, nodes=n_ + 1, tolerance=1e-8, homotopy=tanh)
    return inst, m


def reseeded(inst, seed):
    """The same family and sizes as inst with A, B, C, x0 drawn from another
    seed: a different generated program of identical sizes."""
    A, B, C, x0 = draw_family(inst.nx, inst.nu, seed, inst.C is not None)
    return FamilyInstance(nx=inst.nx, nu=inst.nu, A=A, B=B, C=C, x0=x0, s=inst.s, q=inst.q, c_u=inst.c_u, T=inst.T,
                          nodes=inst.nodes, tolerance=inst.tolerance, homotopy=inst.homotopy)


def batch_family(count, seed=1, nx=20, nu=10, nint_range=(256, 2048)):
    """Batched family: shared A, B, C from `seed`; per instance i,
    default_rng([seed, i]) draws x0 ~ U[-1,1]^nx, T ~ U[0.5, 2] and
    Nint ~ U{nint_range}."""
    A, B, C, _ = draw_family(nx, nu, seed, True)
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        x0 = rng.uniform(-1.0, 1.0, nx)
        T = rng.uniform(0.5, 2.0)
        nint = int(rng.integers(nint_range[0], nint_range[1] + 1))
        out.append(FamilyInstance(nx=nx, nu=nu, A=A, B=B, C=C, x0=x0, s=0.5, q=1.0, c_u=0.1, T=T,
                                  nodes=nint + 1, tolerance=1e-8, instance_x0=True, homotopy=True))
    return out


# A problem whose dynamics depend on a parameter (so its condensed system has
# an active parameter column): a forced oscillator whose frequency pp is a
# free parameter with a terminal penalty pulling it towards 1.2.
PARAM_OSC_OCP = """
StateVariables = [x1, x2];
ControlVariables = [u1];
ParameterVariables = [pp];
CostFunctional = x1*x1 + x2*x2 + u1*u1;
TerminalPenalty = 5*(pp - 1.2)^2;
DifferentialEquations = [pp*x2, -pp*x1 + u1];
InitialConstraints = [x1 - 1, x2];
TerminalConstraints = [x1];
InitialTime = 0;
FinalTime = 1;
Nodes = 65;
Tolerance = 1e-8;
"""


class TextProblem:
    """An OCP given as text (no generated family data)."""

    def __init__(self, text):
        self.text = text

    def ocp_text(self):
        return self.text


def standard_units():
    """Problems whose device code the build step compiles ahead of time:
    (label, FamilyInstance, m).  Mesh size does not change the generated
    code, so one entry serves every Nint of its family."""
    yield "lq4_m4", config(0)[0], 4
    small, _ = config(1, nx=4, nu=2, nint=32)
    yield "tanh4_m5", small, 5
    yield "tanh4_seed2_m5", reseeded(small, 2), 5
    yield "tanh20_m5", config(1)[0], 5
    yield "param_osc_m4", TextProblem(PARAM_OSC_OCP), 4
    # batched families: x0 (and the tanh scale) are instance constants, so
    # instance 0's text stands for every member
    yield "tanh4_batch_m5", batch_family(1, nx=4, nu=2, nint_range=(16, 64))[0], 5
    yield "tanh20_batch_m5", batch_family(1)[0], 5
