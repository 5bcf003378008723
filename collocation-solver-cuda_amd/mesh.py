"""Mesh adaptation between continuation steps (host side, once per mesh).

Restates remesh (collocation_solver_parallel_2d_shared.py:2106-2311):
 1. deletion: wherever five consecutive nodes i..i+4 all have residual
    <= 1e-2, keep nodes i, i+2, i+4 and continue after them;
 2. insertion: for each remaining interval with residual r_i > 10 insert
    three equispaced nodes, for 1 < r_i <= 10 insert one midpoint;
    y and z are interpolated linearly (as are the residuals, which only
    steer later decisions of the same pass).
"""
import numpy as np

DELETE_THRESHOLD = 1e-2


def remesh(size_y, size_z, N, tspan, y0, z0, residual):
    keep = []
    i = 0
    while i < N - 4:
        if residual[i] <= DELETE_THRESHOLD and all(residual[i + k] <= DELETE_THRESHOLD for k in range(1, 5)):
            keep.extend((i, i + 2, i + 4))
            i += 5
        else:
            keep.append(i)
            i += 1
    keep.extend(range(i, N))
    idx = np.array(keep, dtype=np.int64)
    t_t, y_t, z_t, r_t = tspan[idx], y0[idx], z0[idx], np.asarray(residual)[idx]
    t_out, y_out, z_out = [], [], []
    for k in range(len(idx) - 1):
        parts = 4 if r_t[k] > 10 else (2 if r_t[k] > 1 else 1)
        dt = (t_t[k + 1] - t_t[k]) / parts
        dy = (y_t[k + 1] - y_t[k]) / parts
        dz = (z_t[k + 1] - z_t[k]) / parts
        for q in range(parts):
            t_out.append(t_t[k] + q * dt)
            y_out.append(y_t[k] + q * dy)
            z_out.append(z_t[k] + q * dz)
    t_out.append(t_t[-1])
    y_out.append(y_t[-1])
    z_out.append(z_t[-1])
    return len(t_out), np.array(t_out), np.array(y_out).reshape(-1, size_y), np.array(z_out).reshape(-1, size_z)


def mesh_sanity_check(N, mesh_it, max_mesh, max_nodes, min_nodes):
    """solver:2367-2377: True when the mesh leaves the allowed range."""
    return mesh_it > max_mesh or N > max_nodes or N < min_nodes
