"""CPU restatement of the reference's collocation solver driver and stages.

Follows collocation_solver_parallel_2d_shared.py ("solver" below).  The
problem object supplies size_y, size_z, size_p and vectorised callbacks
f(Y,Z,P), g(Y,Z,P,alpha), Df, Dg (stage axis first) and r(y0,y1,p), Dr --
the roles of the generated _abvp_* callbacks (OCP2ABVP.py:203-327).

Array layouts are the reference's: y (N, ny); y_dot (Nint*m, ny) and
z_tilde (Nint*m, nz) with row i*m + j for stage j of interval i
(solver:278-283); f_a (Nint, m*(ny+nz)); f_b (Nint, ny).
"""
import numpy as np

from . import babd
from .tables import compute_I, compute_L, gauss, lobatto

MAX_LINESEARCH = 20          # solver:64
M_PARTITIONS = 8             # solver:78


def _dims(prob):
    return prob.size_y, prob.size_z, prob.size_p


def form_initial_input(prob, m, N, y0, z0, p0, rk):
    """Initial y_dot, z_tilde from node values (solver:287-356)."""
    ny, nz, _ = _dims(prob)
    c = rk.c
    ytmp = (1.0 - c)[None, :, None] * y0[:-1, None, :] + c[None, :, None] * y0[1:, None, :]
    ztil = (1.0 - c)[None, :, None] * z0[:-1, None, :] + c[None, :, None] * z0[1:, None, :]
    ytmp = ytmp.reshape(-1, ny)
    ztil = ztil.reshape(-1, nz)
    y_dot = prob.f(ytmp, ztil, p0)
    return y0, y_dot, ztil


def collocation_update(m, N, rk, t_span, y, y_dot):
    """Stage values y_tilde_ij = y_i + dt_i sum_k a_jk ydot_ik (solver:520-559)."""
    ny = y.shape[1]
    dt = np.diff(t_span)
    Yd = y_dot.reshape(N - 1, m, ny)
    s = np.einsum("jk,ikl->ijl", rk.A, Yd)
    return (y[:-1, None, :] + dt[:, None, None] * s).reshape(-1, ny)


def compute_f_q(prob, m, N, rk, t_span, y, y_dot, z_tilde, p, alpha):
    """Collocation residual (solver:399-466, kernels :478-512).

    Returns f_q (N*ny + Nint*m*(ny+nz) + np,), y_tilde, f_a, f_b, r_bc.
    """
    ny, nz, np_ = _dims(prob)
    nint = N - 1
    y_tilde = collocation_update(m, N, rk, t_span, y, y_dot)
    rh = prob.f(y_tilde, z_tilde, p) - y_dot
    rg = prob.g(y_tilde, z_tilde, p, alpha)
    f_a = np.concatenate([rh.reshape(nint, m, ny), rg.reshape(nint, m, nz)], axis=2).reshape(nint, -1)
    dt = np.diff(t_span)
    sum_i = np.einsum("j,ijl->il", rk.b, y_dot.reshape(nint, m, ny))
    f_b = y[1:] - y[:-1] - dt[:, None] * sum_i
    r_bc = prob.r(y[0], y[N - 1], p)
    f_q = np.concatenate([np.concatenate([f_a, f_b], axis=1).ravel(), r_bc])
    return f_q, y_tilde, f_a, f_b, r_bc


def construct_jacobian(prob, m, N, rk, t_span, y, y_tilde, z_tilde, p, alpha):
    """J, V, D, W per interval plus B_0, B_n, V_n (solver:611-797).

    Shapes: J (Nint, K, ny), V (Nint, K, np), D (Nint, ny, K), W (Nint, K, K),
    K = m (ny + nz).  W block (j, k) = [dt a_jk h_y - d_jk I, d_jk h_z;
    dt a_jk g_y, d_jk g_z] with the stage-j derivatives.
    """
    ny, nz, np_ = _dims(prob)
    nint = N - 1
    n = ny + nz
    K = m * n
    dt = np.diff(t_span)
    Df = prob.Df(y_tilde, z_tilde, p).reshape(nint, m, ny, ny + nz + np_)
    Dg = prob.Dg(y_tilde, z_tilde, p, alpha).reshape(nint, m, nz, ny + nz + np_)
    J = np.zeros((nint, m, n, ny)); V = np.zeros((nint, m, n, np_))
    J[:, :, :ny] = Df[..., :ny]; J[:, :, ny:] = Dg[..., :ny]
    V[:, :, :ny] = Df[..., ny + nz:]; V[:, :, ny:] = Dg[..., ny + nz:]
    D = np.zeros((nint, ny, m, n))
    for j in range(m):
        D[:, :, j, :ny] = (dt * rk.b[j])[:, None, None] * np.eye(ny)[None]
    W = np.zeros((nint, m, n, m, n))
    for j in range(m):
        for k in range(m):
            fac = (dt * rk.A[j, k])[:, None, None]
            W[:, j, :ny, k, :ny] = fac * Df[:, j, :, :ny]
            W[:, j, ny:, k, :ny] = fac * Dg[:, j, :, :ny]
        W[:, j, :ny, j, :ny] -= np.eye(ny)[None]
        W[:, j, :ny, j, ny:] = Df[:, j, :, ny:ny + nz]
        W[:, j, ny:, j, ny:] = Dg[:, j, :, ny:ny + nz]
    d_r = prob.Dr(y[0], y[N - 1], p)
    B_0 = d_r[:, 0:ny]; B_n = d_r[:, ny:2 * ny]; V_n = d_r[:, 2 * ny:2 * ny + np_]
    return (J.reshape(nint, K, ny), V.reshape(nint, K, np_), D.reshape(nint, ny, K),
            W.reshape(nint, K, K), B_0, B_n, V_n)


def reduce_jacobian(prob, m, N, W, D, J, V, f_a, f_b):
    """Condensation (solver:840-1001): with P W = L U per interval,
    A = -I + D W^-1 J,  C = I,  H = D W^-1 V,  b = -f_b - D W^-1 f_a."""
    ny, nz, np_ = _dims(prob)
    rhs = np.concatenate([J, V, f_a[:, :, None]], axis=2)
    X = np.linalg.solve(W, rhs)                # LAPACK getrf/getrs, first-max partial pivoting
    DX = D @ X
    A = -np.eye(ny)[None] + DX[:, :, :ny]
    H = DX[:, :, ny:ny + np_]
    b = -f_b - DX[:, :, ny + np_]
    C = np.broadcast_to(np.eye(ny), A.shape).copy()
    return A, C, H, b


def recover_delta_k(prob, m, N, delta_y, delta_p, f_a, J, V, W):
    """delta_k_i = W_i^-1 (-f_a - J_i dy_i - V_i dp) (solver:1580-1650)."""
    ny, nz, _ = _dims(prob)
    vec = -f_a - np.einsum("ikl,il->ik", J, delta_y[:-1]) - V @ delta_p
    delta_k = np.linalg.solve(W, vec[:, :, None])[:, :, 0]
    dk = delta_k.reshape(N - 1, m, ny + nz)
    return delta_k, dk[:, :, :ny].reshape(-1, ny).copy(), dk[:, :, ny:].reshape(-1, nz).copy()


def newton_direction(prob, m, N, rk, t_span, y, y_tilde, z_tilde, p, alpha, f_a, f_b, r_bc, M=M_PARTITIONS):
    """One Newton direction exactly as solver:123-151 sequences it."""
    ny, nz, np_ = _dims(prob)
    J, V, D, W, B_0, B_n, V_n = construct_jacobian(prob, m, N, rk, t_span, y, y_tilde, z_tilde, p, alpha)
    A, C, H, b = reduce_jacobian(prob, m, N, W, D, J, V, f_a, f_b)
    delta_y, delta_p = babd.solve_babd(ny, np_, N, A, C, H, b, B_0, B_n, V_n, r_bc, M=M)
    _, d_ydot, d_ztil = recover_delta_k(prob, m, N, delta_y, delta_p, f_a, J, V, W)
    return delta_y, d_ydot, d_ztil, delta_p


def newton_solve(prob, m, rk, t_span, y, y_dot, z_tilde, p, alpha, tol, max_iter, M=M_PARTITIONS, log=None):
    """Damped Newton loop with the reference's line search (solver:109-174).

    Stops when ||f_q||_inf < tol before taking a step; the step length tries
    1, 1/2, ..., 2^-19 and takes the first with strict decrease; a failed line
    search leaves the iterate unchanged.  Returns the new state, the loop
    counter iter_time at exit and the norm history.
    """
    N = t_span.shape[0]
    norms = []
    iter_time = 0
    for iter_time in range(max_iter):
        f_q, y_tilde, f_a, f_b, r_bc = compute_f_q(prob, m, N, rk, t_span, y, y_dot, z_tilde, p, alpha)
        norm_f_q = np.max(np.abs(f_q))
        norms.append(norm_f_q)
        if log is not None:
            log(iter_time, norm_f_q)
        if norm_f_q < tol:
            break
        dy, dyd, dz, dp = newton_direction(prob, m, N, rk, t_span, y, y_tilde, z_tilde, p, alpha,
                                           f_a, f_b, r_bc, M=M)

        def trial_norm(a0):
            f_q_new = compute_f_q(prob, m, N, rk, t_span, y + a0 * dy, y_dot + a0 * dyd, z_tilde + a0 * dz,
                                  p + a0 * dp, alpha)[0]
            return np.max(np.abs(f_q_new))
        a0 = line_search(trial_norm, norm_f_q)
        if a0 is not None:
            y, y_dot, z_tilde, p = y + a0 * dy, y_dot + a0 * dyd, z_tilde + a0 * dz, p + a0 * dp
    return y, y_dot, z_tilde, p, iter_time, norms


def line_search(trial_norm, norm0, max_trials=MAX_LINESEARCH):
    """The reference's step rule (solver:155-174): alpha0 = 1, 1/2, ...,
    2^-(max_trials-1); the first step whose trial norm is strictly below
    norm0 is taken.  Returns that step, or None when no trial decreased the
    norm (the reference then leaves the iterate unchanged)."""
    a0 = 1.0
    for _ in range(max_trials):
        if trial_norm(a0) < norm0:
            return a0
        a0 /= 2.0
    return None


def _gauss_point_values(prob, m, N, t_span, y, y_dot, z_tilde, i_gp, gc):
    """ydot, y and z at Gauss point i_gp of every interval (get_y_dot/get_y/get_z, solver:1671-1742)."""
    ny, nz, _ = _dims(prob)
    nint = N - 1
    dt = np.diff(t_span)
    Yd = y_dot.reshape(nint, m, ny)
    Zt = z_tilde.reshape(nint, m, nz)
    L = compute_L(m, gc.t[i_gp])
    I = compute_I(m, gc.t[i_gp])
    yd_tau = np.einsum("k,ikl->il", L, Yd)
    y_tau = y[:-1].copy()
    for k in range(m):                       # get_y accumulates dt * I_k * ydot_k term by term
        y_tau += dt[:, None] * I[k] * Yd[:, k, :]
    z_tau = np.einsum("k,ikl->il", L, Zt)
    return yd_tau, y_tau, z_tau


def compute_segment_residual(prob, m, N, t_span, y, y_dot, z_tilde, p, alpha, tol, residual_type=1,
                             scale_by_time=True):
    """Per-interval acceptance residual (solver:1770-1892).  residual_type 0
    (max), 1 (sum of |.|), 2 (sum of squares, sqrt at the end).  Quirks kept:
    the time scaling rho *= dt sits inside the component loop
    (solver:1868-1869, 1883-1884); for residual_type 0 the g part takes |h|
    of the same component index (solver:1880; components beyond ny count 0).
    The last entry is the boundary norm ||r||_inf / tol (solver:1811-1817)."""
    ny, nz, np_ = _dims(prob)
    nint = N - 1
    gc = gauss(m + 1)
    dt = np.diff(t_span)
    rho_h = np.zeros(nint)
    rho_g = np.zeros(nint)
    for i in range(m + 1):
        yd_tau, y_tau, z_tau = _gauss_point_values(prob, m, N, t_span, y, y_dot, z_tilde, i, gc)
        h_res = prob.f(y_tau, z_tau, p) - yd_tau
        for k in range(ny):
            if residual_type == 2:
                rho_h += gc.w[i] * h_res[:, k] * h_res[:, k]
            elif residual_type == 1:
                rho_h += gc.w[i] * np.abs(h_res[:, k])
            else:
                rho_h = np.maximum(rho_h, gc.w[i] * np.abs(h_res[:, k]))
            if scale_by_time:
                rho_h *= dt
        if nz > 0:
            g_res = prob.g(y_tau, z_tau, p, alpha)
            for k in range(nz):
                if residual_type == 2:
                    rho_g += gc.w[i] * g_res[:, k] * g_res[:, k]
                elif residual_type == 1:
                    rho_g += gc.w[i] * np.abs(g_res[:, k])
                else:
                    rho_g = np.maximum(rho_g, gc.w[i] * np.abs(h_res[:, k]) if k < ny else 0.0)
                if scale_by_time:
                    rho_g *= dt
    residual = np.zeros(N)
    if residual_type == 2:
        residual[:nint] = np.sqrt(rho_h + rho_g) / tol
    elif residual_type == 1:
        residual[:nint] = (np.abs(rho_h) + np.abs(rho_g)) / tol
    else:
        residual[:nint] = np.maximum(rho_h, rho_g) / tol
    if ny + np_ > 0:
        residual[N - 1] = np.max(np.abs(prob.r(y[0], y[N - 1], p))) / tol
    return residual, np.max(residual)


def compute_segment_residual2(prob, m, N, t_span, y, y_dot, z_tilde, p, alpha, tol):
    """compute_segment_residual_parallel2 (solver:1895-2013): per interval the
    largest |h_k| / (1 + max over Gauss points |ydot_k|) and
    |g_k| / (1 + max |z_k|), over / tol; last entry ||r||_inf / tol."""
    ny, nz, np_ = _dims(prob)
    nint = N - 1
    gc = gauss(m + 1)
    vals = [_gauss_point_values(prob, m, N, t_span, y, y_dot, z_tilde, i, gc) for i in range(m + 1)]
    ydm = np.max([np.abs(v[0]) for v in vals], axis=0)
    zm = np.max([np.abs(v[2]) for v in vals], axis=0) if nz > 0 else None
    rho_h = np.zeros(nint)
    rho_g = np.zeros(nint)
    for yd_tau, y_tau, z_tau in vals:
        h_res = prob.f(y_tau, z_tau, p) - yd_tau
        rho_h = np.maximum(rho_h, np.max(np.abs(h_res) / (1.0 + ydm), axis=1))
        if nz > 0:
            g_res = prob.g(y_tau, z_tau, p, alpha)
            rho_g = np.maximum(rho_g, np.max(np.abs(g_res) / (1.0 + zm), axis=1))
    residual = np.zeros(N)
    residual[:nint] = np.maximum(rho_h, rho_g) / tol
    if ny + np_ > 0:
        residual[N - 1] = np.max(np.abs(prob.r(y[0], y[N - 1], p))) / tol
    return residual, np.max(residual)


def recover_solution(m, N, z_tilde):
    """z at the nodes (solver:2034-2075): first stage of each interval, last
    stage of the final interval (Lobatto basis at tau = 0 and tau = 1)."""
    nz = z_tilde.shape[1]
    Zt = z_tilde.reshape(N - 1, m, nz)
    z = np.zeros((N, nz))
    z[:N - 1] = np.einsum("j,ijk->ik", compute_L(m, 0.0), Zt)
    z[N - 1] = compute_L(m, 1.0) @ Zt[N - 2]
    return z


def remesh(size_y, size_z, N, tspan, y0, z0, residual):
    """Mesh deletion/insertion (solver:2106-2311)."""
    t_tmp, y_tmp, z_tmp, r_tmp = [], [], [], []
    i = 0
    thr = 1e-2
    while i < N - 4:
        if residual[i] <= thr and all(residual[i + q] <= thr for q in range(1, 5)):
            for q in (0, 2, 4):
                t_tmp.append(tspan[i + q]); y_tmp.append(y0[i + q]); z_tmp.append(z0[i + q])
                r_tmp.append(residual[i + q])
            i += 5
        else:
            t_tmp.append(tspan[i]); y_tmp.append(y0[i]); z_tmp.append(z0[i]); r_tmp.append(residual[i])
            i += 1
    while i < N:
        t_tmp.append(tspan[i]); y_tmp.append(y0[i]); z_tmp.append(z0[i]); r_tmp.append(residual[i])
        i += 1
    n_tmp = len(t_tmp)
    t_new, y_new, z_new = [], [], []
    for i in range(n_tmp - 1):
        res = r_tmp[i]
        if res > 1:
            parts = 4 if res > 10 else 2
            dtq = (t_tmp[i + 1] - t_tmp[i]) / parts
            dyq = (y_tmp[i + 1] - y_tmp[i]) / parts
            dzq = (z_tmp[i + 1] - z_tmp[i]) / parts
            for q in range(parts):
                t_new.append(t_tmp[i] + q * dtq)
                y_new.append(y_tmp[i] + q * dyq)
                z_new.append(z_tmp[i] + q * dzq)
        else:
            t_new.append(t_tmp[i]); y_new.append(y_tmp[i]); z_new.append(z_tmp[i])
    t_new.append(t_tmp[-1]); y_new.append(y_tmp[-1]); z_new.append(z_tmp[-1])
    return len(t_new), np.array(t_new), np.array(y_new), np.array(z_new)


def collocation_solver(prob, m, t_span0, y0, z0, p0, tol, max_iter=200, max_mesh=10, max_nodes=2000,
                       inequalities=False, M=M_PARTITIONS, homotopy=()):
    """Continuation + Newton + mesh loop (solver:36-237).

    homotopy: problems solved in turn (Newton to tol, warm-started) before
    `prob`; not part of the reference, mirrors the product's option."""
    rk = lobatto(m)
    alpha = 0.1
    alpha_m = 1e-6 if inequalities else 0.1
    beta = 0.9
    min_nodes = 3
    t_span = np.array(t_span0, dtype=np.float64)
    N = t_span.shape[0]
    p = np.array(p0, dtype=np.float64)
    y, y_dot, z_tilde = form_initial_input(prob, m, N, np.asarray(y0, np.float64), np.asarray(z0, np.float64), p, rk)
    success = 1
    history = []
    max_residual = 1.0 / tol
    for hp in homotopy:
        y, y_dot, z_tilde, p, iter_time, norms = newton_solve(
            hp, m, rk, t_span, y, y_dot, z_tilde, p, alpha, tol, max_iter, M=M)
        history.append({"alpha": alpha, "N": N, "iterations": iter_time, "norms": norms})
    mesh_it = 0
    iter_time = 0
    for alphacal in range(max_iter):
        mesh_it = 0
        y, y_dot, z_tilde, p, iter_time, norms = newton_solve(
            prob, m, rk, t_span, y, y_dot, z_tilde, p, alpha, tol, max_iter, M=M)
        history.append({"alpha": alpha, "N": N, "iterations": iter_time, "norms": norms})
        if alphacal >= max_iter - 1 or iter_time >= max_iter - 1 or N > max_nodes \
                or mesh_it > max_mesh or N < min_nodes:
            success = 0
            break
        residual, max_residual = compute_segment_residual(prob, m, N, t_span, y, y_dot, z_tilde, p, alpha, tol)
        history[-1]["max_residual"] = max_residual
        if max_residual > 1:
            z = recover_solution(m, N, z_tilde)
            N, t_span, y, z = remesh(prob.size_y, prob.size_z, N, t_span, y, z, residual)
            mesh_it += 1
            y, y_dot, z_tilde = form_initial_input(prob, m, N, y, z, p, rk)
            if mesh_it > max_mesh or N > max_nodes or N < min_nodes:
                success = 0
                break
        else:
            if alpha <= alpha_m:
                break
            alpha *= beta
    z = recover_solution(m, N, z_tilde)
    return {"t": t_span, "y": y, "z": z, "p": p, "y_dot": y_dot, "z_tilde": z_tilde,
            "success": success, "alpha": alpha, "max_residual": max_residual, "history": history}
