"""Extract the reference's printed coefficient tables into a JSON fixture.

The reference hard-codes Lobatto IIIA tableaux (collocation_coefficients.py:24-635)
and Gauss-Legendre nodes/weights (gauss_coefficients.py:12-131) as decimal
literals.  This script reads those two files as TEXT (nothing is imported or
executed), collects the literals and writes ``reference_tables.json`` next to
itself.  The fixture is data: the numbers the reference's solver runs with.

Usage: python tests/golden/make_reference_tables.py [/path/to/reference]
"""
import ast
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _lobatto(text):
    out = {}
    m = None
    for line in text.splitlines():
        hdr = re.match(r"\s*(?:el)?if m == (\d+):", line)
        if hdr:
            m = int(hdr.group(1))
            continue
        if line.startswith("@cuda.jit") or line.startswith("def "):
            m = None
        if m is None:
            continue
        entry = out.setdefault(m, {"A": [[None] * m for _ in range(m)], "b": [None] * m,
                                   "c": [None] * m, "A_assignments": []})
        a = re.match(r"\s*self\.A\[(\d+), (\d+)\] = (\S+)", line)
        if a:
            i, j, v = int(a.group(1)), int(a.group(2)), float(a.group(3))
            entry["A"][i][j] = v
            entry["A_assignments"].append([i, j])
            continue
        for key in ("b", "c"):
            v = re.match(r"\s*self\.%s\[(\d+)\] = (\S+)" % key, line)
            if v:
                entry[key][int(v.group(1))] = float(v.group(2))
    return out


def _eval_num(node, env):
    """Evaluate a literal arithmetic expression tree (numbers, names, + - * /)."""
    if isinstance(node, ast.Constant):
        return float(node.value)
    if isinstance(node, ast.Name):
        return env[node.id]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_eval_num(node.operand, env)
    if isinstance(node, ast.BinOp):
        a, b = _eval_num(node.left, env), _eval_num(node.right, env)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):
            return a / b
    raise ValueError("unsupported literal expression")


def _gauss(text):
    out = {}
    n = None
    env = {}
    for line in text.splitlines():
        hdr = re.match(r"\s*(?:el)?if n == (\d+):", line)
        if hdr:
            n, env = int(hdr.group(1)), {}
            continue
        if n is None:
            continue
        s = line.strip()
        v = re.match(r"([xw]\d+) = ([-0-9.eE+]+)$", s)
        if v:
            env[v.group(1)] = float(v.group(2))
            continue
        lst = re.match(r"([tw]) = (\[.*\])$", s)
        if lst:
            tree = ast.parse(lst.group(2), mode="eval").body
            out.setdefault(n, {})[lst.group(1)] = [_eval_num(e, env) for e in tree.elts]
    return out


def main():
    with open(os.path.join(REF, "collocation_coefficients.py")) as f:
        lob = _lobatto(f.read())
    with open(os.path.join(REF, "gauss_coefficients.py")) as f:
        gau = _gauss(f.read())
    data = {"source": {"lobatto": "collocation_coefficients.py:24-635",
                       "gauss": "gauss_coefficients.py:12-131"},
            "lobatto": {str(k): v for k, v in sorted(lob.items())},
            "gauss": {str(k): v for k, v in sorted(gau.items())}}
    with open(os.path.join(HERE, "reference_tables.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("lobatto m:", sorted(lob), "gauss n:", sorted(gau))


if __name__ == "__main__":
    main()
