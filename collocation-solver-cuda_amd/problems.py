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


# Constrained problems of the reference's example set (Notebook/source), the
# same mathematics written here as OCP text: a trolley carrying a pendulum
# load with a bounded control (exTrolley), a rocket in free space with control
# and state-variable bounds and a free final time (exFreespaceRocket), a
# container crane with a free final time (exContainerTransfer) and an
# underwater vehicle with bounded controls (exUnderwaterVehicle).  Their
# InputFiles are not shipped with the reference, so they start from the
# all-ones guess.
TROLLEY_OCP = """
Constants = [pi = 3.14159265358979, m1 = 0.3, m2 = 0.5, L = 0.75, Iy = 0.002, g = 9.81];
Variables = [den, num3, num4];
StateVariables = [x1, x2, x3, x4];
ControlVariables = [u];
den = -m1*m2*L*L - m1*Iy - m2*m2*L*L - m2*Iy + m2*m2*L*L*cos(x2)*cos(x2);
num3 = m2*m2*L*L*L*sin(x2)*x4*x4 - m2*L*L*u + m2*Iy*L*x4*x4*sin(x2) - Iy*u + m2*m2*L*L*g*cos(x2)*sin(x2);
num4 = m2*L*(m2*L*cos(x2)*x4*x4*sin(x2) - cos(x2)*u + g*sin(x2)*(m1 + m2));
DifferentialEquations = [x3, x4, num3/den, num4/den];
CostFunctional = 0.5*(u*u + 5*x4*x4);
InitialConstraints = [x1, x2, x3, x4];
TerminalConstraints = [x1 - 1, x2, x3, x4];
InequalityConstraints = [u - 0.5, -0.5 - u];
InitialTime = 0.0;
FinalTime = 2.7;
Nodes = 271;
Tolerance = 1.0e-6;
MaximumNodes = 1000;
MaximumMeshRefinements = 20;
"""

ROCKET_OCP = """
StateVariables = [y1, y2, y3];
ControlVariables = [u];
ParameterVariables = [p];
DifferentialEquations = [p*y2, p*(u - 0.2*y2*y2)/y3, -0.01*p*u*u];
CostFunctional = p*(1 + 0.1*u*u);
InitialConstraints = [y1, y2, y3 - 1.0];
TerminalConstraints = [y1 - 10, y2];
InequalityConstraints = [u - 1.1, -1.1 - u];
StateVariableInequalityConstraints = [y2 - 1.7, -0.1 - y2];
InitialTime = 0.0;
FinalTime = 1.0;
Nodes = 101;
Tolerance = 1.0e-6;
"""

CONTAINER_OCP = """
Constants = [rho = 1.0];
StateVariables = [x1, x2, x3, x4, x5, x6];
ControlVariables = [u1, u2];
ParameterVariables = [p];
DifferentialEquations = [p*x4, p*x5, p*x6, p*(u1 + 17.2656*x3), p*u2, -p*(u1 + 27.0756*x3 + 2.0*x5*x6)/x2];
CostFunctional = p*(rho + 0.01*(u1*u1 + u2*u2));
InitialConstraints = [x1, x2 - 22.0, x3, x4, x5 + 1.0, x6];
TerminalConstraints = [x1 - 10.0, x2 - 14.0, x3, x4 - 2.5, x5, x6];
InequalityConstraints = [u1 - 2.83374, -2.83374 - u1, u2 - 0.71265, -0.80865 - u2];
ParameterEstimate = [9.0];
Tolerance = 1.0e-6;
MaximumNodes = 2000;
MaximumNewtonIterations = 500;
MaximumMeshRefinements = 25;
"""

UNDERWATER_OCP = """
Constants = [pi = 3.14159265358979, uxmax = 2.0, uzmax = 1.0, cx = 0.5, cz = 0.1, rx0 = 0.1];
Variables = [d1, d2, rx, rz];
StateVariables = [x1, x2, x3, x4, x5, x6, x7, x8, x9, x10];
ControlVariables = [u1, u2, u3, u4];
d1 = (x1 - cx)/rx0;
d2 = (x3 - cz)/cz;
rx = -uxmax*exp(-d1*d1)*(x1 - cx)*d2*d2;
rz = -uzmax*exp(-d1*d1)*d2*d2;
DifferentialEquations = [cos(x6)*cos(x5)*x7 + rx, sin(x6)*cos(x5)*x7, -sin(x5)*x7 + rz,
                         x8 + sin(x4)*tan(x5)*x9 + cos(x4)*tan(x5)*x10, cos(x4)*x9 - sin(x4)*x10,
                         (x9*sin(x4) + x10*cos(x4))/cos(x5), u1, u2, u3, u4];
CostFunctional = u1*u1 + u2*u2 + u3*u3 + u4*u4;
InitialConstraints = [x1, x2, x3 - 0.02, x4 - pi/2.0, x5 - 0.1, x6 + pi/4.0, x7 - 1.0, x8, x9 - 0.5, x10 - 0.1];
TerminalConstraints = [x1 - 1.0, x2 - 0.5, x3, x4 - pi/2.0, x5, x6, x7, x8, x9, x10];
InequalityConstraints = [u1 - 15.0, -15.0 - u1, u2 - 15.0, -15.0 - u2, u3 - 15.0, -15.0 - u3, u4 - 15.0, -15.0 - u4];
Tolerance = 1.0e-6;
MaximumNodes = 2000;
MaximumNewtonIterations = 500;
MaximumMeshRefinements = 20;
"""

CONSTRAINED_EXAMPLES = {"trolley": TROLLEY_OCP, "rocket": ROCKET_OCP, "container": CONTAINER_OCP,
                        "underwater": UNDERWATER_OCP}


class TextProblem:
    """An OCP given as text (no generated family data)."""

    def __init__(self, text):
        self.text = text

    def ocp_text(self):
        return self.text


def form_units():
    """Problems whose device code the tests build in an explicit callback
    form ("code": straight-line, "tables": polynomial tables), to compare
    the two: (label, OCP text, m, form)."""
    small, _ = config(1, nx=4, nu=2, nint=32)
    items = [("tanh4", small.ocp_text(), 5),
             ("tanh4_batch", batch_family(1, nx=4, nu=2, nint_range=(16, 64))[0].ocp_text(), 5),
             ("trolley", CONSTRAINED_EXAMPLES["trolley"], 4), ("param_osc", PARAM_OSC_OCP, 4)]
    for label, text, m in items:
        for form in ("code", "tables"):
            yield "%s_%s_m%d" % (label, form, m), text, m, form


def standard_units():
    """Problems whose device code the build step compiles ahead of time:
    (label, FamilyInstance, m).  Mesh size does not change the generated
    code, so one entry serves every Nint of its family."""
    yield "lq4_m4", config(0)[0], 4
    small, _ = config(1, nx=4, nu=2, nint=32)
    yield "tanh4_m5", small, 5
    yield "tanh4_seed2_m5", reseeded(small, 2), 5
    yield "tanh20_m5", config(1)[0], 5
    yield "tanh64_m4", config(2)[0], 4
    yield "param_osc_m4", TextProblem(PARAM_OSC_OCP), 4
    for name, text in CONSTRAINED_EXAMPLES.items():
        yield "%s_m4" % name, TextProblem(text), 4
    # batched families: x0 (and the tanh scale) are instance constants, so
    # instance 0's text stands for every member
    yield "tanh4_batch_m5", batch_family(1, nx=4, nu=2, nint_range=(16, 64))[0], 5
    yield "tanh20_batch_m5", batch_family(1)[0], 5
