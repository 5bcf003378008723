"""Host side of the solver: the reference's problem object, stage functions
and driver, backed by libcolloc.so.

Mirrors, with the same names and argument meaning:
  * bvp_problem.BvpDae             -- the generated problem class (OCP2ABVP.py:329-447)
  * the stage functions of collocation_solver_parallel_2d_shared.py
    (form_initial_input_parallel :287, compute_f_q_parallel :399,
    construct_jacobian_parallel :611, reduce_jacobian_parallel :840,
    recover_delta_k_parallel :1580, compute_segment_residual_parallel :1770,
    recover_solution_parallel :2034) and the BABD solve of
    solve_babd_system.py (:93-572), as methods of a device-resident
    CollocationSolver;
  * collocation_solver_parallel (:36-251), the continuation / Newton /
    remesh driver.
The arithmetic runs in the HIP kernels; this module moves arrays across the
C ABI and runs the once-per-mesh host logic (remesh, file output).
"""
import numpy as np
import sympy as sp

from . import jit, native
from .bvpdae_io import bvpdae_read_data, bvpdae_write_data
from .mesh import mesh_sanity_check, remesh
from .native import CSDims, check, lib, ptr
from .ocp import parse_ocp

MAX_LINESEARCH = 20   # solver:64


class BvpDae:
    """Problem sizes, settings and initial guess (the generated BvpDae class).

    The initial guess is all ones on linspace(t_i, t_f, Nodes) unless an
    InputFile can be read or estimates are given, as in OCP2ABVP.py:355-447.
    """

    def __init__(self, ocp_text, instance_constants=None):
        self.ocp_text = ocp_text
        o = parse_ocp(ocp_text)
        self.ocp = o
        self.size_y, self.size_z, self.size_p, self.size_c = jit.sizes_of(o)
        self.size_inequality = len(o.d)
        self.size_sv_inequality = len(o.s)
        self.output_file = o.output_file
        self.tolerance = o.tolerance
        self.maximum_nodes = o.maximum_nodes
        self.maximum_newton_iterations = o.maximum_newton_iterations
        self.maximum_mesh_refinements = o.maximum_mesh_refinements
        if self.size_c:
            if instance_constants is None:
                instance_constants = [o.constants[s] for s in o.instance_constants]
            self.constants = np.asarray(instance_constants, dtype=np.float64)
        else:
            self.constants = np.zeros(0)
        read = 1
        if o.input_file:
            read, t0, y0, z0, p0 = bvpdae_read_data(o.input_file)
        if read != 0:
            self.N = o.nodes
            self.t_initial, self.t_final = o.t_i, o.t_f
            self.T0 = np.linspace(self.t_initial, self.t_final, self.N)
            self.Y0 = np.ones((self.N, self.size_y))
            self.Z0 = np.ones((self.N, self.size_z))
            self.P0 = np.ones(self.size_p)
        else:
            self.N = t0.shape[0]
            self.T0 = t0
            self.Y0 = np.ones((self.N, self.size_y))
            self.Z0 = np.ones((self.N, self.size_z))
            self.P0 = np.ones(self.size_p)
            ny = min(y0.shape[1], self.size_y)
            self.Y0[:, :ny] = y0[:, :ny]
            nu = min(z0.shape[1], len(o.u))
            self.Z0[:, :nu] = z0[:, :nu]
            npp = min(p0.shape[0], self.size_p)
            self.P0[:npp] = p0[:npp]
        if o.state_estimate or o.control_estimate or o.parameter_estimate:
            self._solution_estimate()

    def _solution_estimate(self):
        o = self.ocp
        subs = {s: v for s, v in o.constants.items()}
        tsym = sp.Symbol("t")

        def ev(e, tv):
            return float(sp.sympify(e).xreplace(subs).subs(tsym, tv))
        for i, tv in enumerate(self.T0):
            for j, e in enumerate(o.state_estimate or []):
                self.Y0[i, j] = ev(e, tv)
            for j, e in enumerate(o.control_estimate or []):
                self.Z0[i, j] = ev(e, tv)
        for j, e in enumerate(o.parameter_estimate or []):
            self.P0[j] = ev(e, 0.0)


class CollocationSolver:
    """Device-resident solver state for one BVP-DAE and stage count m.

    The handle holds a batch of up to `batch` instances of the problem (same
    generated code, each with its own mesh, constants, state and parameters);
    set_mesh makes a batch of one.  For a batch of one every method returns
    what the reference's stage function returns; for B > 1 the per-solve
    scalars (norms, iteration counts, maximum residuals) are arrays of B
    values and the arrays are the instances' arrays concatenated.
    """

    def __init__(self, bvp, m, n_max=None, batch=1, form="auto"):
        self.bvp = bvp
        self.m = m
        self.form = form   # device code of the callbacks: "auto", "code" or "tables" (jit.choose_form)
        self.direction_precision = 64
        self.babd_precision = 64
        self.ny, self.nz, self.np = bvp.size_y, bvp.size_z, bvp.size_p
        self.n_max = int(n_max or max(bvp.maximum_nodes, bvp.N))
        self.batch = int(batch)
        L = lib()
        dims = CSDims(self.ny, self.nz, self.np, m, self.n_max, bvp.size_c, self.batch)
        h = native.c_void_p()
        check(L.cs_create(native.ctypes.byref(dims), native.ctypes.byref(h)))
        self.h = h
        self.load_problem(bvp)
        self.N = 0
        self.B = 1
        self.nodes = np.zeros(1, dtype=np.int32)

    def load_problem(self, bvp):
        """Load the device code of a problem of this handle's sizes (the
        generated callbacks and the kernels specialised to them); the mesh and
        state are kept.  The role of importing a generated bvp_problem module
        (solver:10-17)."""
        code = jit.code_object(bvp.ocp_text, self.m, form=self.form)
        buf = native.ctypes.create_string_buffer(code, len(code))
        check(lib().cs_load_module(self.h, buf, len(code)))
        self.bvp = bvp
        if bvp.size_c:
            self.set_constants(bvp.constants)

    def set_constants(self, values):
        """Per-instance constants of the problem (InstanceConstants order),
        given to every instance of the batch."""
        c = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().cs_set_constants(self.h, ptr(c), c.shape[0]))

    def set_batch_constants(self, values):
        """Constants of each instance: values[B][nc]."""
        c = np.ascontiguousarray(values, dtype=np.float64).reshape(self.B, -1)
        check(lib().cs_set_batch_constants(self.h, ptr(c), c.shape[0], c.shape[1]))

    def close(self):
        if getattr(self, "h", None):
            lib().cs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------
    def set_mesh(self, t_span):
        t = np.ascontiguousarray(t_span, dtype=np.float64)
        check(lib().cs_set_mesh(self.h, ptr(t), t.shape[0]))
        self.N = t.shape[0]
        self.B = 1
        self.nodes = np.array([self.N], dtype=np.int32)
        self.t_span = t

    def set_batch_mesh(self, t_spans):
        """Meshes of a batch: a list of B node arrays."""
        nodes = np.ascontiguousarray([len(t) for t in t_spans], dtype=np.int32)
        t = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.float64) for t in t_spans]))
        check(lib().cs_set_batch_mesh(self.h, len(t_spans), native.iptr(nodes), ptr(t)))
        self.N = int(nodes.sum())
        self.B = len(t_spans)
        self.nodes = nodes
        self.t_span = t

    def node_offsets(self):
        return np.concatenate([[0], np.cumsum(self.nodes)])

    def interval_offsets(self):
        return np.concatenate([[0], np.cumsum(self.nodes - 1)])

    def _per_solve(self, v):
        return v[0] if self.B == 1 else v

    def _shapes(self):
        nst = (self.N - self.B) * self.m
        ps = (self.B, self.np) if self.B > 1 else (self.np,)
        return (self.N, self.ny), (nst, self.ny), (nst, self.nz), ps

    def set_state(self, y, y_dot, z_tilde, p):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (y, y_dot, z_tilde, p)]
        for a, s in zip(arrs, self._shapes()):
            if a.size != int(np.prod(s)):
                raise ValueError("state array of size %d, expected %s" % (a.size, s))
        check(lib().cs_set_state(self.h, *[ptr(a) for a in arrs]))

    def get_state(self):
        out = [np.empty(s) for s in self._shapes()]
        check(lib().cs_get_state(self.h, *[ptr(a) for a in out]))
        return tuple(out)

    def get_state_parts(self, y=False, y_dot=False, z_tilde=False, p=False):
        """get_state copying only the requested arrays (the others are None)."""
        out = [np.empty(s) if want else None for s, want in zip(self._shapes(), (y, y_dot, z_tilde, p))]
        check(lib().cs_get_state(self.h, *[ptr(a) for a in out]))
        return tuple(out)

    # -- stages --------------------------------------------------------------
    def form_initial_input(self, y0, z0, p0):
        """form_initial_input_parallel (solver:287): the handle's state becomes
        y0, the initial y_dot and z_tilde, and p0 (read it with get_state)."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        z0 = np.ascontiguousarray(z0, dtype=np.float64).reshape(self.N, self.nz)
        p0 = np.ascontiguousarray(p0, dtype=np.float64)
        check(lib().cs_form_initial_input(self.h, ptr(y0), ptr(z0), ptr(p0)))

    def restart(self):
        """Re-form the state from the last initial input (device-resident)."""
        check(lib().cs_restart(self.h))

    def compute_f_q(self, alpha, outputs=True):
        """compute_f_q_parallel (solver:399): (norm, y_tilde, f_a, f_b, r_bc)."""
        nrm = np.zeros(self.B)
        nint = self.N - self.B
        nst = nint * self.m
        if outputs:
            yt = np.empty((nst, self.ny)); fa = np.empty((nint, self.m * (self.ny + self.nz)))
            fb = np.empty((nint, self.ny))
            rbc = np.empty(self.ny + self.np) if self.B == 1 else np.empty((self.B, self.ny + self.np))
        else:
            yt = fa = fb = rbc = None
        check(lib().cs_compute_f_q(self.h, alpha, ptr(nrm), ptr(yt), ptr(fa), ptr(fb), ptr(rbc)))
        return self._per_solve(nrm), yt, fa, fb, rbc

    def construct_jacobian(self, alpha, outputs=True):
        """construct_jacobian_parallel (solver:611): per-stage Df, Dg and Dr."""
        nst = (self.N - self.B) * self.m
        nj = self.ny + self.nz + self.np
        if outputs:
            Df = np.empty((nst, self.ny, nj)); Dg = np.empty((nst, self.nz, nj))
            shp = (self.ny + self.np, 2 * self.ny + self.np)
            Dr = np.empty(shp if self.B == 1 else (self.B,) + shp)
        else:
            Df = Dg = Dr = None
        check(lib().cs_construct_jacobian(self.h, alpha, ptr(Df), ptr(Dg), ptr(Dr)))
        return Df, Dg, Dr

    def reduce_jacobian(self, outputs=True):
        """reduce_jacobian_parallel (solver:840): A, C (= I), H, b."""
        nint = self.N - self.B
        if outputs:
            A = np.empty((nint, self.ny, self.ny)); H = np.empty((nint, self.ny, self.np)); b = np.empty((nint, self.ny))
        else:
            A = H = b = None
        check(lib().cs_reduce_jacobian(self.h, ptr(A), ptr(H), ptr(b)))
        C = None if A is None else np.broadcast_to(np.eye(self.ny), A.shape)
        return A, C, H, b

    def solve_babd(self):
        """BABD solve (solve_babd_system.py:93-572): delta_y (N, ny), delta_p."""
        dy = np.empty((self.N, self.ny))
        dp = np.empty(self.np) if self.B == 1 else np.empty((self.B, self.np))
        check(lib().cs_solve_babd(self.h, ptr(dy), ptr(dp)))
        return dy, dp

    def set_babd_substitution(self, min_spans):
        """Cyclic-reduction pairing rule: substitution while a pair spans at
        most 1/min_spans of an instance's intervals, Householder QR above
        (0: Householder everywhere, as the reference)."""
        check(lib().cs_set_babd_substitution(self.h, int(min_spans)))

    def set_launch(self, which, value):
        """A launch shape that does not change results (native.CS_LAUNCH_*):
        the series condensation's persistent grid (0: one workgroup per
        interval) or the residual kernel's workgroup size."""
        check(lib().cs_set_launch(self.h, int(which), int(value)))

    def get_launch(self, which):
        v = np.zeros(1, dtype=np.int32)
        check(lib().cs_get_launch(self.h, int(which), native.iptr(v)))
        return int(v[0])

    def set_direction_precision(self, bits):
        """Condensation arithmetic of the Newton direction: 64 (default) or
        32 (an inexact-Newton direction; residuals, BABD solve and the
        convergence test stay float64)."""
        check(lib().cs_set_direction_precision(self.h, int(bits)))
        self.direction_precision = int(bits)

    def set_newton_graph(self, on):
        """Run the Newton loops device resident, one HIP-graph replay per
        iteration (the same results as the host-driven loop)."""
        check(lib().cs_set_newton_graph(self.h, 1 if on else 0))

    def set_babd_precision(self, bits):
        """Stored type of the cyclic reduction's rows and records: 64
        (default) or 32 (float32 levels, float64 boundary solve and back
        substitution)."""
        check(lib().cs_set_babd_precision(self.h, int(bits)))
        self.babd_precision = int(bits)

    def set_refinement(self, steps):
        """Iterative refinement steps per Newton direction (0: none): the
        float64 linear residual rho = F + J delta and a correction solve with
        the same direction kernels, delta <- delta + e."""
        check(lib().cs_set_refinement(self.h, int(steps)))

    def linear_residual(self):
        """max |F + J delta| per solve for the current direction (the last
        solve_babd / recover_delta_k) at the last compute_f_q and
        construct_jacobian."""
        out = np.zeros(self.B)
        check(lib().cs_linear_residual(self.h, ptr(out)))
        return self._per_solve(out)

    def refine_direction(self, steps=1):
        """Refinement steps on the current stage-level direction; f_a, f_b,
        r_bc then hold the last linear residual (recompute with compute_f_q)."""
        check(lib().cs_refine_direction(self.h, int(steps)))

    def get_direction(self):
        """The current Newton direction (delta_y, delta_y_dot, delta_z_tilde,
        delta_p) as it stands, nothing recomputed."""
        nst = (self.N - self.B) * self.m
        dy = np.empty((self.N, self.ny)); dyd = np.empty((nst, self.ny)); dz = np.empty((nst, self.nz))
        dp = np.empty(self.np) if self.B == 1 else np.empty((self.B, self.np))
        check(lib().cs_get_direction(self.h, ptr(dy), ptr(dyd), ptr(dz), ptr(dp)))
        return dy, dyd, dz, dp

    def set_condense_kernel(self, kind):
        """Series condensation kernel of the FP32 direction: 0 matrix cores
        (default, where the stage blocks fit), 1 vector ALU."""
        check(lib().cs_set_condense_kernel(self.h, int(kind)))

    def recover_delta_k(self):
        """recover_delta_k_parallel (solver:1580): delta_y_dot, delta_z_tilde."""
        nst = (self.N - self.B) * self.m
        dyd = np.empty((nst, self.ny)); dz = np.empty((nst, self.nz))
        check(lib().cs_recover_delta_k(self.h, ptr(dyd), ptr(dz)))
        return dyd, dz

    def newton(self, alpha, tol, max_iter, patience=0, min_ratio=0.0, window=1):
        """Damped Newton loop (solver:109-174): (iter_time, norm, norm history).
        For a batch: arrays of B iteration counts and norms, and a list of B
        histories.  patience > 0 adds the progress test of cs_newton_progress
        (window 1, for warm-started solves) or, for window > 1, the stall test
        of cs_newton_stall from iteration `patience` on (for fallback
        continuation stages); the reference has neither."""
        it = np.zeros(self.B, dtype=np.int32)
        nrm = np.zeros(self.B)
        hist = np.zeros((self.B, max_iter))
        if patience > 0 and window > 1:
            check(lib().cs_newton_stall(self.h, alpha, tol, max_iter, int(patience), int(window), float(min_ratio),
                                        native.iptr(it), ptr(nrm), ptr(hist), max_iter))
        elif patience > 0:
            check(lib().cs_newton_progress(self.h, alpha, tol, max_iter, int(patience), float(min_ratio),
                                           native.iptr(it), ptr(nrm), ptr(hist), max_iter))
        else:
            check(lib().cs_newton(self.h, alpha, tol, max_iter, native.iptr(it), ptr(nrm), ptr(hist), max_iter))
        hists = [hist[b, :min(int(it[b]) + 1, max_iter)].copy() for b in range(self.B)]
        if self.B == 1:
            return int(it[0]), float(nrm[0]), hists[0]
        return it.astype(int), nrm, hists

    def segment_residual(self, alpha, tol, residual_type=1, scale_by_time=True, residual2=False):
        """compute_segment_residual_parallel (solver:1770), or with residual2
        compute_segment_residual_parallel2 (solver:1895): (residual, max_residual).
        residual_type / scale_by_time are the reference's module flags (solver:30-31)."""
        res = np.empty(self.N)
        mx = np.zeros(self.B)
        check(lib().cs_segment_residual_ex(self.h, alpha, tol, int(residual_type), int(bool(scale_by_time)),
                                           int(bool(residual2)), ptr(res), ptr(mx)))
        return res, self._per_solve(mx)

    def recover_solution(self):
        """recover_solution_parallel (solver:2034): z at the nodes."""
        z = np.empty((self.N, self.nz))
        check(lib().cs_recover_solution(self.h, ptr(z)))
        return z

    def condense_stats(self):
        """(intervals condensed by Gauss-Jordan because the series solve did
        not converge for them, series terms summed over intervals, intervals
        condensed), since the solver was created."""
        n = native.ctypes.c_longlong()
        k = native.ctypes.c_longlong()
        c = native.ctypes.c_longlong()
        check(lib().cs_condense_stats(self.h, native.ctypes.byref(n), native.ctypes.byref(k), native.ctypes.byref(c)))
        return n.value, k.value, c.value

    def condense_fallbacks(self):
        return self.condense_stats()[0]

    def status(self):
        s = native.c_int()
        check(lib().cs_status(self.h, native.ctypes.byref(s)))
        return s.value

    def profile(self, enable):
        check(lib().cs_profile(self.h, 1 if enable else 0))

    def profile_get(self, kernel):
        ms = native.c_double()
        n = native.c_int()
        check(lib().cs_profile_get(self.h, native.KERNELS[kernel], native.ctypes.byref(ms), native.ctypes.byref(n)))
        return ms.value, n.value


def collocation_solver_parallel(bvp_dae, m=4, solver=None, write_output=False, log=None, homotopy=None,
                                residual_type=1, scale_by_time=True, residual2=False):
    """Continuation + Newton + mesh loop (solver:36-237).

    Same control flow as the reference: alpha starts at 0.1 (alpha_m = 1e-6
    with inequalities, else 0.1, beta = 0.9); each continuation step runs the
    damped Newton loop, checks the limits, computes the segment residual and
    either remeshes (residual > 1) or accepts.  Returns a dict with the
    solution on the final mesh and the per-step history.

    residual_type, scale_by_time, residual2: the reference's module flags
    (solver:30-33) selecting the mesh-acceptance residual.

    homotopy (an extension, not in the reference): a list of instance-constant
    vectors to pass through before the problem's own constants.  Newton is run
    to the tolerance at each of them in turn, each warm-started from the
    previous solution; the reference algorithm then starts from the result.
    """
    ny, nz, np_ = bvp_dae.size_y, bvp_dae.size_z, bvp_dae.size_p
    tol = bvp_dae.tolerance
    max_iter = bvp_dae.maximum_newton_iterations
    max_mesh = bvp_dae.maximum_mesh_refinements
    max_nodes = bvp_dae.maximum_nodes
    min_nodes = 3
    alpha = 0.1
    alpha_m = 1e-6 if (bvp_dae.size_inequality > 0 or bvp_dae.size_sv_inequality > 0) else 0.1
    beta = 0.9
    own = solver is None
    if own:
        solver = CollocationSolver(bvp_dae, m, n_max=max(max_nodes, bvp_dae.N) + 4 * bvp_dae.N)
    t_span = np.array(bvp_dae.T0, dtype=np.float64)
    N = t_span.shape[0]
    para = np.array(bvp_dae.P0, dtype=np.float64)
    solver.set_mesh(t_span)
    solver.form_initial_input(bvp_dae.Y0, bvp_dae.Z0, para)
    success = 1
    max_residual = 1.0 / tol
    history = []
    for cvals in (homotopy or []):
        solver.set_constants(cvals)
        iter_time, nrm, norms = solver.newton(alpha, tol, max_iter)
        history.append({"alpha": alpha, "N": N, "iterations": iter_time, "norms": norms,
                        "constants": list(map(float, cvals))})
        if log:
            log("homotopy %s: newton iterations=%d norm=%.3e" % (list(cvals), iter_time, nrm))
    if homotopy:
        solver.set_constants(bvp_dae.constants)
    mesh_it = 0
    for alphacal in range(max_iter):
        mesh_it = 0
        iter_time, nrm, norms = solver.newton(alpha, tol, max_iter)
        history.append({"alpha": alpha, "N": N, "iterations": iter_time, "norms": norms})
        if log:
            log("alpha=%g N=%d newton iterations=%d norm=%.3e" % (alpha, N, iter_time, nrm))
        if alphacal >= max_iter - 1 or iter_time >= max_iter - 1 or N > max_nodes or mesh_it > max_mesh \
                or N < min_nodes:
            success = 0
            break
        residual, max_residual = solver.segment_residual(alpha, tol, residual_type, scale_by_time, residual2)
        history[-1]["max_residual"] = max_residual
        if max_residual > 1:
            y, _, _, para = solver.get_state()
            z = solver.recover_solution()
            N, t_span, y, z = remesh(ny, nz, N, t_span, y, z, residual)
            mesh_it += 1
            if N > solver.n_max:
                success = 0
                break
            solver.set_mesh(t_span)
            solver.form_initial_input(y, z, para)
            if mesh_sanity_check(N, mesh_it, max_mesh, max_nodes, min_nodes):
                success = 0
                break
        else:
            if alpha <= alpha_m:
                break
            alpha *= beta
    y, y_dot, z_tilde, para = solver.get_state()
    z = solver.recover_solution()
    if write_output and alpha <= alpha_m:
        bvpdae_write_data(bvp_dae.output_file, N, ny, nz, np_, t_span, y, z, para)
    result = {"t": t_span, "y": y, "z": z, "p": para, "y_dot": y_dot, "z_tilde": z_tilde, "success": success,
              "alpha": alpha, "max_residual": max_residual, "history": history, "N": N}
    if own:
        solver.close()
    return result


def solve_batch(solver, t_spans, Y0s, Z0s, P0s, constants=None, homotopy=(), alpha=0.1, tol=1e-8, max_iter=200,
                residual_type=1, scale_by_time=True, residual2=False):
    """One damped-Newton solve (solver:109-174) of every instance of a batch
    in one device-resident sequence of launches, followed by the acceptance
    residual (solver:1770-1892) of each.

    t_spans, Y0s, Z0s, P0s: per instance mesh and initial guess (the
    reference's T0, Y0, Z0, P0).  constants: [B][nc] instance constants of
    the final solve; homotopy: a list of [B][nc] constant arrays solved first,
    each warm-starting the next (as the homotopy of
    collocation_solver_parallel).  Returns a dict with per-instance arrays
    "iterations" ([stages][B]), "norm", "max_residual" and "converged"
    (norm < tol, max residual <= 1 and fewer than max_iter - 1 iterations).
    """
    B = len(t_spans)
    solver.set_batch_mesh(t_spans)
    solver.form_initial_input(np.concatenate(Y0s), np.concatenate(Z0s),
                              np.stack([np.asarray(p, dtype=np.float64) for p in P0s]) if B > 1 else P0s[0])
    return newton_stages(solver, constants, homotopy, alpha, tol, max_iter, residual_type, scale_by_time,
                          residual2)


def newton_stages(solver, constants, homotopy, alpha, tol, max_iter, residual_type=1, scale_by_time=True,
                   residual2=False, patience=0, min_ratio=0.0, stall=None):
    """Newton at each homotopy constant vector in turn, then at `constants`,
    then the acceptance residual (see solve_batch).  patience / min_ratio:
    the progress test of cs_newton_progress for the final solve.  stall =
    (start, window, min_ratio): the stall test of cs_newton_stall in every
    stage (an instance stopped by it goes on to the next stage from where it
    stopped)."""
    B = solver.B
    its = []
    nrm = None

    def run(final=False):
        if final and patience > 0:
            it, n, _ = solver.newton(alpha, tol, max_iter, patience, min_ratio)
            return np.atleast_1d(it), np.atleast_1d(n)
        if stall is not None:
            it, n, _ = solver.newton(alpha, tol, max_iter, stall[0], stall[2], window=stall[1])
            return np.atleast_1d(it), np.atleast_1d(n)
        it, n, _ = solver.newton(alpha, tol, max_iter)
        return np.atleast_1d(it), np.atleast_1d(n)
    for cst in homotopy:
        solver.set_batch_constants(cst)
        it, nrm = run()
        its.append(it)
    if constants is not None:
        solver.set_batch_constants(constants)
    it, nrm = run(final=True)
    its.append(it)
    _, mx = solver.segment_residual(alpha, tol, residual_type, scale_by_time, residual2)
    mx = np.atleast_1d(mx)
    ok = (nrm < tol) & (mx <= 1.0) & (it < max_iter - 1)
    return {"iterations": np.array(its), "norm": nrm, "max_residual": mx, "converged": ok, "B": B}


def batch_solutions(solver):
    """Per-instance solutions of the solver's current batch: a list of dicts
    with t, y, z (at the nodes, recover_solution_parallel) and p."""
    y, _, _, p = solver.get_state()
    z = solver.recover_solution()
    p = np.atleast_2d(p)
    off = solver.node_offsets()
    return [{"t": solver.t_span[off[b]:off[b + 1]], "y": y[off[b]:off[b + 1]], "z": z[off[b]:off[b + 1]],
             "p": p[b]} for b in range(solver.B)]


def plan_chunks(nodes, max_nodes, max_batch):
    """Split instances (node counts, in order) into consecutive chunks of at
    most max_batch instances and max_nodes nodes."""
    chunks, cur, tot = [], [], 0
    for i, n in enumerate(nodes):
        if n > max_nodes:
            raise ValueError("instance %d has %d nodes, more than a chunk holds (%d)" % (i, n, max_nodes))
        if cur and (tot + n > max_nodes or len(cur) == max_batch):
            chunks.append(cur)
            cur, tot = [], 0
        cur.append(i)
        tot += n
    if cur:
        chunks.append(cur)
    return chunks


def collocation_solver_batch(bvp_dae, m, t_spans, Y0s, Z0s, P0s, constants=None, homotopy=(), solver=None,
                             max_nodes=None, max_batch=None, solutions=True, log=None):
    """Solve many instances of one problem with batched launches.

    The instances are split into chunks that fit the solver's handle; each
    chunk is solved by solve_batch.  constants / homotopy stages are [B][nc]
    arrays over all instances.  Returns (stats, solutions) with stats the
    per-instance arrays of solve_batch concatenated over chunks and
    solutions a list of per-instance dicts (None unless solutions=True).
    """
    B = len(t_spans)
    nodes = [len(t) for t in t_spans]
    if solver is None:
        max_nodes = max_nodes or sum(nodes)
        max_batch = max_batch or B
        solver = CollocationSolver(bvp_dae, m, n_max=max_nodes, batch=max_batch)
    max_nodes = min(max_nodes or solver.n_max, solver.n_max)
    max_batch = min(max_batch or solver.batch, solver.batch)
    chunks = plan_chunks(nodes, max_nodes, max_batch)
    tol = bvp_dae.tolerance
    max_iter = bvp_dae.maximum_newton_iterations
    nc = bvp_dae.size_c
    cons = None if constants is None else np.asarray(constants, dtype=np.float64).reshape(B, nc)
    homs = [np.asarray(h, dtype=np.float64).reshape(B, nc) for h in homotopy]
    stats = {"iterations": np.zeros((len(homs) + 1, B), dtype=int), "norm": np.zeros(B),
             "max_residual": np.zeros(B), "converged": np.zeros(B, dtype=bool)}
    sols = [None] * B if solutions else None
    for ch in chunks:
        r = solve_batch(solver, [t_spans[i] for i in ch], [Y0s[i] for i in ch], [Z0s[i] for i in ch],
                        [P0s[i] for i in ch], None if cons is None else cons[ch], [h[ch] for h in homs],
                        tol=tol, max_iter=max_iter)
        stats["iterations"][:, ch] = r["iterations"]
        for k in ("norm", "max_residual", "converged"):
            stats[k][ch] = r[k]
        if solutions:
            for i, s in zip(ch, batch_solutions(solver)):
                sols[i] = s
        if log:
            log("chunk of %d instances (%d nodes): %d converged" % (len(ch), sum(nodes[i] for i in ch),
                                                                     int(r["converged"].sum())))
    return stats, sols
