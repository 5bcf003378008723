"""Per-problem device code: emit, compile with hiprtc, cache the code object.

The cache key is the OCP text, the stage count and the text of every source
that shapes the generated unit (emitter, tables, kernel headers), so a cached
code object is reused only for exactly the same generated program.  Cached
objects live in-tree under _jit/ (built ahead of time by the build step for
the standard problems, so that no compiler runs on the solve path).
"""
import hashlib
import json
import os

from . import build, emitter, native, polyform, tables
from .ocp import _statements, parse_ocp

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_DIR = os.path.join(HERE, "_jit")
_SOURCES = ["emitter.py", "polyform.py", "tables.py", "ocp.py", "jit.py",
            os.path.join("csrc", "kernels", "colloc_common.h"),
            os.path.join("csrc", "kernels", "colloc_gj.h"),
            os.path.join("csrc", "kernels", "colloc_tables.h"),
            os.path.join("csrc", "kernels", "colloc_large.h"),
            os.path.join("csrc", "kernels", "colloc_kernels.h")]


def _sources_digest():
    h = hashlib.sha256()
    for rel in _SOURCES:
        with open(os.path.join(HERE, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def sizes_of(ocp):
    """(ny, nz, np, nc) of the BVP-DAE of an OCP without differentiating it."""
    nx, npar = len(ocp.x), len(ocp.p)
    ny = 2 * nx + npar
    nz = len(ocp.u) + len(ocp.c) + 2 * len(ocp.d) + 2 * len(ocp.s)
    np_ = npar + len(ocp.Gamma) + len(ocp.Psi)
    return ny, nz, np_, len(ocp.instance_constants)


# statements that only set the mesh, tolerances, files or the initial guess;
# they do not change the generated device code
_MESH_ONLY = {"InitialTime", "FinalTime", "Nodes", "Tolerance", "InputFile", "OutputFile", "Display",
              "MaximumMeshRefinements", "MaximumNewtonIterations", "MaximumNodes", "StateEstimate",
              "ControlEstimate", "ParameterEstimate"}


def code_text(ocp_text):
    """The statements of an OCP text that determine the generated code."""
    keep = []
    for s in _statements(ocp_text):
        key = s.split("=", 1)[0].strip()
        if key not in _MESH_ONLY:
            keep.append(" ".join(s.split()))
    return ";\n".join(keep)


def cache_key(ocp_text, m, form="auto"):
    h = hashlib.sha256()
    h.update(code_text(ocp_text).encode())
    h.update(b"\0m=%d\0" % m)
    if form != "auto":
        h.update(b"\0form=%s\0" % form.encode())
    h.update(_sources_digest().encode())
    return h.hexdigest()[:32]


def choose_form(ocp, form="auto"):
    """"tables" (the polynomial table form, polyform.py) or "code" (straight-
    line callbacks from sympy) for a problem: tables when ny exceeds what the
    register-resident kernels hold or the callbacks would exceed
    emitter.TABLE_TERMS terms, unless form forces one.  Returns (form,
    BvpDaePoly or None)."""
    if form == "code":
        return "code", None
    ny = sizes_of(ocp)[0]
    try:
        P = polyform.BvpDaePoly(ocp)
    except polyform.Unsupported:
        if form == "tables" or ny > emitter.LARGE_NY:
            raise
        return "code", None
    if form == "tables" or ny > emitter.LARGE_NY or P.term_count() > emitter.TABLE_TERMS:
        return "tables", P
    return "code", None


def code_object(ocp_text, m, verbose=False, form="auto"):
    """gfx950 code object (bytes) of the problem's kernels, compiling on a
    cache miss.  form: "auto", "code" or "tables" (see choose_form)."""
    key = cache_key(ocp_text, m, form)
    path = os.path.join(CACHE_DIR, key + ".hsaco")
    if os.path.exists(path):
        with open(path, "rb") as f:
            return f.read()
    ocp = parse_ocp(ocp_text)
    kind, P = choose_form(ocp, form)
    if kind == "tables":
        unit = emitter.build_unit_tables(P, m, tables.collocation_tables(m))
        active = P.active_p_columns()
    else:
        sym = emitter.BvpDaeSymbolic(ocp)
        try:   # the polynomial form's Jacobian program for the matrix-core float32 Jacobian
            poly = P or polyform.BvpDaePoly(ocp)
        except polyform.Unsupported:
            poly = None
        unit = emitter.build_unit(sym, m, tables.collocation_tables(m), poly=poly)
        active = emitter.active_p_columns(sym)
    code, log = native.jit_compile(unit, "colloc_%s.hip" % key, build.KERNEL_DIR)
    if verbose and log.strip():
        print(log)
    os.makedirs(CACHE_DIR, exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "wb") as f:
        f.write(code)
    os.replace(tmp, path)
    with open(os.path.join(CACHE_DIR, key + ".json"), "w") as f:
        json.dump({"m": m, "sizes": sizes_of(ocp), "bytes": len(code), "form": kind,
                   "active_params": active}, f)
    return code


def unit_info(ocp_text, m, form="auto"):
    """Metadata of the problem's code object: sizes, form and the parameters
    whose condensed columns are kept (compiles the code object on a cache
    miss)."""
    path = os.path.join(CACHE_DIR, cache_key(ocp_text, m, form) + ".json")
    if not os.path.exists(path):
        code_object(ocp_text, m, form=form)
    with open(path) as f:
        return json.load(f)
