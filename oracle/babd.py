"""CPU restatement of the reference's BABD solver (solve_babd_system.py).

The condensed system is, for intervals i = 0 .. N-2,
    A_i dy_i + C_i dy_{i+1} + H_i dp = b_i,
closed by the boundary rows  B_0 dy_0 + B_n dy_{N-1} + V_n dp = -r_bc.
The reference splits the interval chain into M partitions, eliminates the
interior nodes of each partition by a chain of 2ny x ny Householder QRs
(partition_factorization_kernel, solve_babd_system.py:161-258), solves the
reduced (M+1)-block system on the host with LAPACK QR (qr_decomposition /
backward_substitution, :338-423) and back-substitutes the interiors
(partition_backward_substitution_kernel, :537-572).
"""
import numpy as np


def householder_qr(a, eps=np.finfo(np.float64).eps):
    """Householder QR with explicit Q (matrix_factorization_cuda.py:5-82).

    Follows the reference's conventions: reflector sign g = -sign(v_k)|v|
    (sign(0) = +1), an ABSOLUTE stop when |v|^2 < eps, R taken as the first
    m rows of the reflected copy (sub-diagonal round-off left in place) and Q
    accumulated backwards from the identity.
    """
    n, m = a.shape
    if m > n:
        raise TypeError("qr requires a matrix with columns <= rows")
    cpy = np.array(a, dtype=np.float64, copy=True)
    q = np.eye(n)
    vs_list = []
    col = m
    for k in range(m):
        vv = cpy[k:, k].copy()
        s = 1.0 if vv[0] >= 0 else -1.0
        g = -s * np.sqrt(np.dot(vv, vv))
        vv[0] -= g
        vs = np.dot(vv, vv)
        if vs < eps:
            col = k
            break
        beta = -2.0 / vs
        w = vv @ cpy[k:, k:m]
        cpy[k:, k:m] += beta * np.outer(vv, w)
        vs_list.append((beta, vv))
    r = np.zeros((n, m))
    r[:m, :m] = cpy[:m, :m]
    for k in range(col - 1, -1, -1):
        beta, vv = vs_list[k]
        u = vv @ q[k:, k:]
        q[k:, k:] += beta * np.outer(vv, u)
    return q, r


def partition_indices(M, N):
    """Partition start nodes (solve_babd_system.py:99-106): [0, q, ..., (M-1)q, N-1]."""
    q = (N - 1) // M
    return np.array([i * q for i in range(M)] + [N - 1], dtype=np.int64)


def partition_factorization(size_y, size_p, M, N, A, C, H, b):
    """Per-partition QR elimination (solve_babd_system.py:93-258).

    A, C: (N-1, ny, ny); H: (N-1, ny, np); b: (N-1, ny).
    Returns index, R, E, J, G, d (per interval) and the partition-end blocks
    A_tilde, C_tilde, H_tilde (M, ny, .) and b_tilde (M, ny).
    """
    ny = size_y
    index = partition_indices(M, N)
    nint = N - 1
    R = np.zeros((nint, ny, ny)); E = np.zeros((nint, ny, ny)); G = np.zeros((nint, ny, ny))
    Jr = np.zeros((nint, ny, size_p)); d = np.zeros((nint, ny))
    Ct = np.zeros((nint, ny, ny)); Gt = np.zeros((nint, ny, ny))
    Ht = np.zeros((nint, ny, size_p)); bt = np.zeros((nint, ny))
    A_e = np.zeros((M, ny, ny)); C_e = np.zeros((M, ny, ny))
    H_e = np.zeros((M, ny, size_p)); b_e = np.zeros((M, ny))
    zero = np.zeros((ny, ny))
    for i in range(M):
        rs, re = index[i], index[i + 1] - 1
        Ct[rs], Gt[rs], Ht[rs], bt[rs] = C[rs], A[rs], H[rs], b[rs]
        for j in range(rs, re):
            q, r = householder_qr(np.vstack([Ct[j], A[j + 1]]))
            qt = q.T
            EC = qt @ np.vstack([zero, C[j + 1]])
            E[j], Ct[j + 1] = EC[:ny], EC[ny:]
            GG = qt @ np.vstack([Gt[j], zero])
            G[j], Gt[j + 1] = GG[:ny], GG[ny:]
            JH = qt @ np.vstack([Ht[j], H[j + 1]])
            Jr[j], Ht[j + 1] = JH[:ny], JH[ny:]
            db = qt @ np.concatenate([bt[j], b[j + 1]])
            d[j], bt[j + 1] = db[:ny], db[ny:]
            R[j] = r[:ny]
        A_e[i], C_e[i], H_e[i], b_e[i] = Gt[re], Ct[re], Ht[re], bt[re]
    return index, R, E, Jr, G, d, A_e, C_e, H_e, b_e


def solve_reduced(size_s, size_p, M, A_t, C_t, H_t, b_t, B_0, B_n, V_n, b_n):
    """Reduced (M+1)-block system: construct_babd_mshoot + qr_decomposition +
    backward_substitution + recover_babd_solution (solve_babd_system.py:311-440).

    Unknowns are delta_s_0 .. delta_s_M (the partition boundary nodes) and dp.
    Returns (delta_s (M+1, ny), delta_p (np,)).
    """
    n = M + 1
    s = size_s
    # construct_babd_mshoot keeps H columns 0:ny (solve_babd_system.py:317), so the
    # reference needs np <= ny; here the full H block is kept.
    Ct = [None] * n; Gt = [None] * n; Ht = [None] * n; bt = [None] * n
    Rr = [None] * n; Er = [None] * n; Gr = [None] * n; Kr = [None] * n; dr = [None] * n
    Ct[0], Gt[0], Ht[0], bt[0] = C_t[0], A_t[0], H_t[0], b_t[0]
    for i in range(n - 2):
        Q, Rq = np.linalg.qr(np.vstack([Ct[i], A_t[i + 1]]), mode="complete")
        Rr[i] = Rq[:s, :]
        EC = Q.T @ np.vstack([np.zeros((s, s)), C_t[i + 1]])
        Er[i], Ct[i + 1] = EC[:s], EC[s:]
        GG = Q.T @ np.vstack([Gt[i], np.zeros((s, s))])
        Gr[i], Gt[i + 1] = GG[:s], GG[s:]
        JH = Q.T @ np.vstack([Ht[i], H_t[i + 1]])
        Kr[i], Ht[i + 1] = JH[:s, :size_p], JH[s:, :size_p]
        db = Q.T @ np.concatenate([bt[i], b_t[i + 1]])
        dr[i], bt[i + 1] = db[:s], db[s:]
    up = np.hstack([Ct[n - 2], Gt[n - 2], Ht[n - 2]])
    down = np.hstack([B_n, B_0, V_n])
    Q, Rq = np.linalg.qr(np.vstack([up, down]), mode="complete")
    R_last = Rq[:s, :s]; G_last = Rq[:s, s:2 * s]; K_last = Rq[:s, 2 * s:2 * s + size_p]
    R_n = Rq[s:2 * s, s:2 * s]; K_n = Rq[s:2 * s, 2 * s:2 * s + size_p]
    Rp = Rq[2 * s:2 * s + size_p, 2 * s:2 * s + size_p]
    dd = Q.T @ np.concatenate([bt[n - 2], b_n])
    d_last, d_n, dp_n = dd[:s], dd[s:2 * s], dd[2 * s:2 * s + size_p]

    def _solve(Mat, rhs):
        try:
            return np.linalg.solve(Mat, rhs)
        except np.linalg.LinAlgError:       # solve_babd_system.py:393-423 substitutes zeros
            return np.zeros(rhs.shape)

    delta_s = np.zeros((n, s))
    delta_p = _solve(Rp, dp_n) if size_p > 0 else np.zeros(0)
    ds1 = _solve(R_n, d_n - K_n @ delta_p)
    delta_s[0] = ds1
    delta_s[n - 1] = _solve(R_last, d_last - G_last @ ds1 - K_last @ delta_p)
    for i in range(n - 2, 0, -1):
        rhs = dr[i - 1] - Er[i - 1] @ delta_s[i + 1] - Gr[i - 1] @ ds1 - Kr[i - 1] @ delta_p
        delta_s[i] = _solve(Rr[i - 1], rhs)
    return delta_s, delta_p


def partition_backward_substitution(size_y, size_p, M, N, index, delta_s_r, delta_p, R, G, E, J, d):
    """Interior nodes from the partition factors (solve_babd_system.py:495-572)."""
    delta_s = np.zeros((N, size_y))
    for i in range(M):
        rs, re = index[i], index[i + 1] - 1
        delta_s[rs] = delta_s_r[i]
        delta_s[re + 1] = delta_s_r[i + 1]
        for j in range(re, rs, -1):
            Rj = np.triu(R[j - 1])                      # :551-553 zeroes the strict lower part
            vec = d[j - 1] - G[j - 1] @ delta_s[rs] - E[j - 1] @ delta_s[j + 1] - J[j - 1] @ delta_p
            delta_s[j] = np.linalg.solve(Rj, vec)
    return delta_s


def solve_babd(size_y, size_p, N, A, C, H, b, B_0, B_n, V_n, r_bc, M=8):
    """Whole BABD solve exactly as the reference's driver sequences it (solver:134-147)."""
    index, R, E, Jr, G, d, A_t, C_t, H_t, b_t = partition_factorization(size_y, size_p, M, N, A, C, H, b)
    ds_r, dp = solve_reduced(size_y, size_p, M, A_t, C_t, H_t, b_t, B_0, B_n, V_n, -r_bc)
    dy = partition_backward_substitution(size_y, size_p, M, N, index, ds_r, dp, R, G, E, Jr, d)
    return dy, dp
