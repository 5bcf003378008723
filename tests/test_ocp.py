"""OCP text reader: the reference's statement grammar (OCP2ABVP.py:524-783)
parsed without executing the text."""
import os

import pytest
import sympy as sp

from colloc_amd import ocp as ocpmod

BASIC = """
# a comment line
Constants = [a = 2.0, b = 3*0.5];
Variables = [v, w];
v = x1 + a*x2;     # macro used below
w = v*v;
StateVariables = [x1, x2];
ControlVariables = [u];
ParameterVariables = [q];
CostFunctional = w + u^2;
TerminalPenalty = q**2;
DifferentialEquations = [x2, -b*x1 + u*q];
InequalityConstraints = [u - 1, -u - 1];
StateVariableInequalityConstraints = [x1 - 5];
InitialConstraints = [x1 - 1, x2];
TerminalConstraints = [x1];
InitialTime = 0;
FinalTime = 2.5;
Nodes = 17;
Tolerance = 1e-7;
InputFile = "guess.data";
OutputFile = "out.data";
MaximumNewtonIterations = 50;
MaximumNodes = 999;
MaximumMeshRefinements = 4;
Display = 1;
"""


def test_statements_and_settings():
    o = ocpmod.parse_ocp(BASIC)
    x1, x2, u, q = sp.symbols("x1 x2 u q")
    a, b = sp.Symbol("a"), sp.Symbol("b")
    assert [s.name for s in o.x] == ["x1", "x2"]
    assert [s.name for s in o.u] == ["u"] and [s.name for s in o.p] == ["q"]
    assert o.constants == {a: 2.0, b: 1.5}
    # macros expand, constants stay symbolic
    assert sp.simplify(o.L - ((x1 + a * x2) ** 2 + u ** 2)) == 0
    assert sp.simplify(o.phi - q ** 2) == 0
    assert [sp.simplify(e1 - e2) for e1, e2 in zip(o.f, [x2, -b * x1 + u * q])] == [0, 0]
    assert len(o.d) == 2 and len(o.s) == 1 and len(o.Gamma) == 2 and len(o.Psi) == 1
    assert (o.t_i, o.t_f, o.nodes, o.tolerance) == (0.0, 2.5, 17, 1e-7)
    assert (o.input_file, o.output_file) == ("guess.data", "out.data")
    assert (o.maximum_newton_iterations, o.maximum_nodes, o.maximum_mesh_refinements, o.display) == (50, 999, 4, 1)


def test_sizes_of_the_bvp_dae():
    from colloc_amd import emitter
    sym = emitter.BvpDaeSymbolic(ocpmod.parse_ocp(BASIC))
    # y = [x, lambda, gamma]; z = [u, theta, mu, xi, nu, sigma]; p = [p, kappa_i, kappa_f]
    assert sym.size_y == 2 + 2 + 1
    assert sym.size_z == 1 + 0 + 2 + 1 + 2 + 1
    assert sym.size_p == 1 + 2 + 1


@pytest.mark.parametrize("text, msg", [
    ("StateVariables = [x];\nFoo = 3;", "unknown statement"),
    ("Variables = [v];\nv = v + 1;\nStateVariables = [x];\nCostFunctional = v;", "recursive macro"),
    ("Constants = [a = 1];\nInstanceConstants = [b];", "not a constant"),
    ("StateVariables = [x];\nCostFunctional = x +* 2;", None),
    ("StateVariables = [x]\nCostFunctional = x;", None),
])
def test_syntax_errors(text, msg):
    with pytest.raises(ocpmod.OCPSyntaxError, match=msg):
        ocpmod.parse_ocp(text)


def test_text_is_never_executed(tmp_path):
    marker = tmp_path / "executed"
    text = ("StateVariables = [x];\nCostFunctional = __import__('os').system('touch %s');\n" % marker)
    with pytest.raises(ocpmod.OCPSyntaxError):
        ocpmod.parse_ocp(text)
    assert not os.path.exists(marker)


def test_instance_constants_and_functions():
    text = ("Constants = [k = 0.5, S = 1];\nInstanceConstants = [S];\nStateVariables = [x];\n"
            "ControlVariables = [u];\nCostFunctional = exp(x) + sqrt(1 + u*u) + abs(x);\n"
            "DifferentialEquations = [S*tanh(k*x) + sin(u) + pow(x, 3)];\n"
            "InitialConstraints = [x - 1];\nTerminalConstraints = [x];")
    o = ocpmod.parse_ocp(text)
    assert [s.name for s in o.instance_constants] == ["S"]
    x, u, k, S = sp.symbols("x u k S")
    assert sp.simplify(o.f[0] - (S * sp.tanh(k * x) + sp.sin(u) + x ** 3)) == 0
