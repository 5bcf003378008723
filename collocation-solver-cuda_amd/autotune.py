"""Autotuner of the run-time launch shapes and the BABD pairing rule.

The kernels are specialised to the problem's sizes at compile time; what
remains open at run time, and depends on the batch shape and the device, is
chosen here by measurement on the handle's current mesh and state:

* the series condensation's launch: one workgroup per interval, or a
  persistent grid of k workgroups per compute unit taking intervals from a
  work queue (cs_set_launch CS_LAUNCH_CONDENSE_GRID; not for ny > 64, whose
  large-block series has no such launch);
* the residual kernel's workgroup size (CS_LAUNCH_RESIDUAL_BLOCK; the
  straight-line callbacks, one thread per stage; the table form's
  evaluator has a fixed workgroup);
* the cyclic reduction's pairing rule (cs_set_babd_substitution: the
  substitution levels' span limit, the "QR tile shape" of the block
  reduction), among the candidates whose Newton direction agrees with the
  Householder-only one (the reference's orthogonal elimination) to
  `direction_tol` relative.  The directions are compared with float64
  rows and records (a property of the elimination order, not of the
  storage); the candidates are timed with the handle's own BABD precision.

Each candidate is timed with the library's HIP-event profile
(cs_profile_get) over `reps` launches of its stage; the launch shapes do
not change any result (the GPU tests check this), the pairing rule changes
the direction only at rounding level (checked here).  Choices are cached
per (problem code, m, direction precision, BABD precision, batch size,
total nodes, device compute units) in _tune/ beside this file, so a later run with the same
shape applies them without measuring.
"""
import hashlib
import json
import os

import numpy as np

from . import emitter, jit, native

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_DIR = os.path.join(HERE, "_tune")
KERNEL = {"residual": 0, "jacobian": 1, "condense": 2, "babd": 3, "recover": 4}


def cache_key(S):
    cus = S.get_launch(native.CS_LAUNCH_COMPUTE_UNITS)
    h = hashlib.sha256()
    h.update(jit.cache_key(S.bvp.ocp_text, S.m, S.form).encode())
    h.update(b"|%d|%d|%d|%d|%d|%d" % (S.direction_precision, S.babd_precision, S.B, int(S.N), S.m, cus))
    return h.hexdigest()[:24]


def _timed(S, kernel, fn, reps):
    """Average device time (ms) of `kernel`'s profile group over reps calls of fn."""
    S.profile(False)
    fn()   # once untimed (first-touch effects)
    S.profile(True)
    for _ in range(reps):
        fn()
    ms, n = S.profile_get(kernel)
    S.profile(False)
    return ms / max(reps, 1)


def _direction(S, alpha):
    S.compute_f_q(alpha, outputs=False)
    S.construct_jacobian(alpha, outputs=False)
    S.reduce_jacobian(outputs=False)
    S.solve_babd()
    S.recover_delta_k()
    return S.get_direction()


def _rel(a, b):
    d = max(np.max(np.abs(x - y)) for x, y in zip(a, b))
    s = max(np.max(np.abs(y)) for y in b)
    return d / max(s, 1e-300)


def tune(S, alpha=0.1, reps=3, grid_factors=(2, 3, 4, 8), residual_blocks=(64, 128, 256),
         spans=(0, 2, 4, 8, 16, 32, 64), direction_tol=1e-7, use_cache=True, log=None):
    """Measure the candidates on S's current mesh and state, apply the
    fastest of each and return {"choices": ..., "timings_ms": ...}.  The
    state is left as it was (only the direction buffers change)."""
    key = cache_key(S)
    path = os.path.join(CACHE_DIR, key + ".json")
    if use_cache and os.path.exists(path):
        with open(path) as f:
            rec = json.load(f)
        apply(S, rec["choices"])
        rec["cached"] = True
        return rec
    say = log or (lambda s: None)
    cus = S.get_launch(native.CS_LAUNCH_COMPUTE_UNITS)
    timings = {}
    choices = {}
    S.compute_f_q(alpha, outputs=False)
    S.construct_jacobian(alpha, outputs=False)
    if S.ny > emitter.LARGE_NY:
        grid_factors = ()
    if jit.unit_info(S.bvp.ocp_text, S.m, S.form)["form"] == "tables":
        residual_blocks = ()
    # condensation launch (results identical for every grid)
    cand = [0] + [k * cus for k in grid_factors]
    tc = {}
    for g in cand:
        S.set_launch(native.CS_LAUNCH_CONDENSE_GRID, g)
        tc[g] = _timed(S, "condense", lambda: S.reduce_jacobian(outputs=False), reps)
        say("condensation grid %d: %.3f ms" % (g, tc[g]))
    best = min(tc, key=tc.get)
    S.set_launch(native.CS_LAUNCH_CONDENSE_GRID, best)
    choices["condense_grid"] = best
    timings["condense_grid"] = {str(k): v for k, v in tc.items()}
    # residual workgroup (straight-line callbacks only: the table form fixes its own)
    tr = {}
    for b in residual_blocks:
        S.set_launch(native.CS_LAUNCH_RESIDUAL_BLOCK, b)
        tr[b] = _timed(S, "residual", lambda: S.compute_f_q(alpha, outputs=False), reps)
        say("residual block %d: %.3f ms" % (b, tr[b]))
    if tr:
        best = min(tr, key=tr.get)
        S.set_launch(native.CS_LAUNCH_RESIDUAL_BLOCK, best)
        choices["residual_block"] = best
        timings["residual_block"] = {str(k): v for k, v in tr.items()}
    # pairing rule, among the directions that agree with Householder-only
    if spans:
        bits = S.babd_precision
        S.set_babd_precision(64)
        S.set_babd_substitution(0)
        ref = _direction(S, alpha)
        err = {}
        for sp in spans:
            S.set_babd_substitution(sp)
            err[sp] = float(_rel(_direction(S, alpha), ref))
        S.set_babd_precision(bits)   # level 0 reads the condensation's rows, float64 for either
        tb = {}
        for sp in spans:
            if err[sp] <= direction_tol:
                S.set_babd_substitution(sp)
                tb[sp] = _timed(S, "babd", lambda: S.solve_babd(), reps)
                say("pairing 1/%d: %.3f ms, direction %.2e" % (sp, tb[sp], err[sp]))
            else:
                say("pairing 1/%d: direction %.2e beyond %.1e, not a candidate" % (sp, err[sp], direction_tol))
        best = min(tb, key=tb.get)
        S.set_babd_substitution(best)
        choices["babd_spans"] = best
        timings["babd_spans"] = {str(k): v for k, v in tb.items()}
        timings["babd_direction_rel_diff"] = {str(k): v for k, v in err.items()}
    rec = {"choices": choices, "timings_ms": timings, "key": key, "compute_units": cus, "cached": False}
    os.makedirs(CACHE_DIR, exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as f:
        json.dump(rec, f, indent=1)
    os.replace(tmp, path)
    return rec


def apply(S, choices):
    if "condense_grid" in choices:
        S.set_launch(native.CS_LAUNCH_CONDENSE_GRID, choices["condense_grid"])
    if "residual_block" in choices:
        S.set_launch(native.CS_LAUNCH_RESIDUAL_BLOCK, choices["residual_block"])
    if "babd_spans" in choices:
        S.set_babd_substitution(choices["babd_spans"])
