"""Solution files in the reference's text format (BVPDAEReadWriteData.py:14-96).

Layout, one item per line:
  nt:  / N  / t:  / t_0 t_1 ... (space separated, trailing space)
  ny:  / ny / y:  / one line per node with ny values
  nz:  / nz / z:  / one line per node with nz values
  np:  / np / p:  / np values on one line
Reading returns (error, t, y, z, p) with error 0 on success and 1 on any
failure, as bvpdae_read_data does.
"""
import numpy as np


def bvpdae_read_data(fname):
    try:
        with open(fname) as f:
            lines = f.read().split("\n")
        it = iter(lines)

        def nxt():
            return next(it)

        nxt()
        N = int(nxt())
        nxt()
        t0 = np.array([float(v) for v in nxt().split()[:N]], dtype=np.float64)
        if t0.shape[0] != N:
            raise ValueError("short time line")
        nxt()
        n_y = int(nxt())
        nxt()
        y0 = np.array([[float(v) for v in nxt().split()[:n_y]] for _ in range(N)], dtype=np.float64).reshape(N, n_y)
        nxt()
        n_z = int(nxt())
        nxt()
        z0 = np.array([[float(v) for v in nxt().split()[:n_z]] for _ in range(N)], dtype=np.float64).reshape(N, n_z)
        nxt()
        n_p = int(nxt())
        nxt()
        p0 = np.array([float(v) for v in nxt().split()[:n_p]], dtype=np.float64)
        if p0.shape[0] != n_p:
            raise ValueError("short parameter line")
        return 0, t0, y0, z0, p0
    except Exception:
        return 1, [], [], [], []


def bvpdae_write_data(fname, N, n_y, n_z, n_p, t_span, y0, z0, p0):
    try:
        parts = ["nt:\n%d\nt:\n" % N, "".join("%r " % float(v) for v in t_span[:N]), "\n"]
        parts.append("ny:\n%d\ny:\n" % n_y)
        for i in range(N):
            parts.append("".join("%r " % float(v) for v in y0[i, :n_y]) + "\n")
        parts.append("nz:\n%d\nz:\n" % n_z)
        for i in range(N):
            parts.append("".join("%r " % float(v) for v in z0[i, :n_z]) + "\n")
        parts.append("np:\n%d\np:\n" % n_p)
        parts.append("".join("%r " % float(v) for v in p0[:n_p]))
        with open(fname, "w") as f:
            f.write("".join(parts))
        return 0
    except Exception:
        return 1
