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
    """Device-resident solver state for one BVP-DAE and stage count m."""

    def __init__(self, bvp, m, n_max=None):
        self.bvp = bvp
        self.m = m
        self.ny, self.nz, self.np = bvp.size_y, bvp.size_z, bvp.size_p
        self.n_max = int(n_max or max(bvp.maximum_nodes, bvp.N))
        code = jit.code_object(bvp.ocp_text, m)
        L = lib()
        dims = CSDims(self.ny, self.nz, self.np, m, self.n_max, bvp.size_c)
        h = native.c_void_p()
        check(L.cs_create(native.ctypes.byref(dims), native.ctypes.byref(h)))
        self.h = h
        buf = native.ctypes.create_string_buffer(code, len(code))
        check(L.cs_load_module(self.h, buf, len(code)))
        if bvp.size_c:
            self.set_constants(bvp.constants)
        self.N = 0

    def set_constants(self, values):
        """Per-instance constants of the problem (InstanceConstants order)."""
        c = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().cs_set_constants(self.h, ptr(c), c.shape[0]))

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
        self.t_span = t

    def _shapes(self):
        nst = (self.N - 1) * self.m
        return (self.N, self.ny), (nst, self.ny), (nst, self.nz), (self.np,)

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

    # -- stages --------------------------------------------------------------
    def form_initial_input(self, y0, z0, p0):
        """form_initial_input_parallel (solver:287): returns y, y_dot, z_tilde."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        z0 = np.ascontiguousarray(z0, dtype=np.float64).reshape(self.N, self.nz)
        p0 = np.ascontiguousarray(p0, dtype=np.float64)
        check(lib().cs_form_initial_input(self.h, ptr(y0), ptr(z0), ptr(p0)))

    def restart(self):
        """Re-form the state from the last initial input (device-resident)."""
        check(lib().cs_restart(self.h))

    def compute_f_q(self, alpha, outputs=True):
        """compute_f_q_parallel (solver:399): (norm, y_tilde, f_a, f_b, r_bc)."""
        nrm = native.c_double()
        nst = (self.N - 1) * self.m
        if outputs:
            yt = np.empty((nst, self.ny)); fa = np.empty((self.N - 1, self.m * (self.ny + self.nz)))
            fb = np.empty((self.N - 1, self.ny)); rbc = np.empty(self.ny + self.np)
        else:
            yt = fa = fb = rbc = None
        check(lib().cs_compute_f_q(self.h, alpha, native.ctypes.byref(nrm), ptr(yt), ptr(fa), ptr(fb), ptr(rbc)))
        return nrm.value, yt, fa, fb, rbc

    def construct_jacobian(self, alpha, outputs=True):
        """construct_jacobian_parallel (solver:611): per-stage Df, Dg and Dr."""
        nst = (self.N - 1) * self.m
        nj = self.ny + self.nz + self.np
        if outputs:
            Df = np.empty((nst, self.ny, nj)); Dg = np.empty((nst, self.nz, nj))
            Dr = np.empty((self.ny + self.np, 2 * self.ny + self.np))
        else:
            Df = Dg = Dr = None
        check(lib().cs_construct_jacobian(self.h, alpha, ptr(Df), ptr(Dg), ptr(Dr)))
        return Df, Dg, Dr

    def reduce_jacobian(self, outputs=True):
        """reduce_jacobian_parallel (solver:840): A, C (= I), H, b."""
        nint = self.N - 1
        if outputs:
            A = np.empty((nint, self.ny, self.ny)); H = np.empty((nint, self.ny, self.np)); b = np.empty((nint, self.ny))
        else:
            A = H = b = None
        check(lib().cs_reduce_jacobian(self.h, ptr(A), ptr(H), ptr(b)))
        C = None if A is None else np.broadcast_to(np.eye(self.ny), A.shape)
        return A, C, H, b

    def solve_babd(self):
        """BABD solve (solve_babd_system.py:93-572): delta_y (N, ny), delta_p."""
        dy = np.empty((self.N, self.ny)); dp = np.empty(self.np)
        check(lib().cs_solve_babd(self.h, ptr(dy), ptr(dp)))
        return dy, dp

    def recover_delta_k(self):
        """recover_delta_k_parallel (solver:1580): delta_y_dot, delta_z_tilde."""
        nst = (self.N - 1) * self.m
        dyd = np.empty((nst, self.ny)); dz = np.empty((nst, self.nz))
        check(lib().cs_recover_delta_k(self.h, ptr(dyd), ptr(dz)))
        return dyd, dz

    def newton(self, alpha, tol, max_iter):
        """Damped Newton loop (solver:109-174): (iter_time, norm, norm history)."""
        it = native.c_int()
        nrm = native.c_double()
        hist = np.zeros(max_iter)
        check(lib().cs_newton(self.h, alpha, tol, max_iter, native.ctypes.byref(it), native.ctypes.byref(nrm),
                              ptr(hist), max_iter))
        return it.value, nrm.value, hist[:min(it.value + 1, max_iter)].copy()

    def segment_residual(self, alpha, tol, residual_type=1, scale_by_time=True, residual2=False):
        """compute_segment_residual_parallel (solver:1770), or with residual2
        compute_segment_residual_parallel2 (solver:1895): (residual, max_residual).
        residual_type / scale_by_time are the reference's module flags (solver:30-31)."""
        res = np.empty(self.N)
        mx = native.c_double()
        check(lib().cs_segment_residual_ex(self.h, alpha, tol, int(residual_type), int(bool(scale_by_time)),
                                           int(bool(residual2)), ptr(res), native.ctypes.byref(mx)))
        return res, mx.value

    def recover_solution(self):
        """recover_solution_parallel (solver:2034): z at the nodes."""
        z = np.empty((self.N, self.nz))
        check(lib().cs_recover_solution(self.h, ptr(z)))
        return z

    def condense_stats(self):
        """(intervals condensed by Gauss-Jordan because the series solve did
        not converge for them, series terms summed over intervals), since the
        solver was created."""
        n = native.ctypes.c_longlong()
        k = native.ctypes.c_longlong()
        check(lib().cs_condense_stats(self.h, native.ctypes.byref(n), native.ctypes.byref(k)))
        return n.value, k.value

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
