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

from . import build, emitter, native, tables
from .ocp import _statements, parse_ocp

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_DIR = os.path.join(HERE, "_jit")
_SOURCES = ["emitter.py", "tables.py", "ocp.py", "jit.py",
            os.path.join("csrc", "kernels", "colloc_common.h"),
            os.path.join("csrc", "kernels", "colloc_gj.h"),
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


def cache_key(ocp_text, m):
    h = hashlib.sha256()
    h.update(code_text(ocp_text).encode())
    h.update(b"\0m=%d\0" % m)
    h.update(_sources_digest().encode())
    return h.hexdigest()[:32]


def code_object(ocp_text, m, verbose=False):
    """gfx950 code object (bytes) of the problem's kernels, compiling on a cache miss."""
    key = cache_key(ocp_text, m)
    path = os.path.join(CACHE_DIR, key + ".hsaco")
    if os.path.exists(path):
        with open(path, "rb") as f:
            return f.read()
    ocp = parse_ocp(ocp_text)
    sym = emitter.BvpDaeSymbolic(ocp)
    unit = emitter.build_unit(sym, m, tables.collocation_tables(m))
    code, log = native.jit_compile(unit, "colloc_%s.hip" % key, build.KERNEL_DIR)
    if verbose and log.strip():
        print(log)
    os.makedirs(CACHE_DIR, exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "wb") as f:
        f.write(code)
    os.replace(tmp, path)
    with open(os.path.join(CACHE_DIR, key + ".json"), "w") as f:
        json.dump({"m": m, "sizes": sizes_of(ocp), "bytes": len(code),
                   "active_params": emitter.active_p_columns(sym)}, f)
    return code


def unit_info(ocp_text, m):
    """Metadata of the problem's code object: sizes and the parameters whose
    condensed columns are kept (compiles the code object on a cache miss)."""
    path = os.path.join(CACHE_DIR, cache_key(ocp_text, m) + ".json")
    if not os.path.exists(path):
        code_object(ocp_text, m)
    with open(path) as f:
        return json.load(f)
