"""CPU oracle for the collocation BVP-DAE solve -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy, the algorithm of the reference solver
(UW-OCP/Collocation-Solver-CUDA) so that the HIP product path can be checked
against it.  Every function cites the reference file:line it follows.

Who may use it: ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- and there only as the checker or as the
timed CPU baseline.  The product package never imports anything from here; the
product path fails loudly when its HIP library is missing instead of falling
back to this code.

Pinning (what checks each part independently of the formulas restated here):

* ``tables``  -- closed forms (Lobatto nodes are 0, 1 and the roots of
  P'_{m-1}; A_jk = int_0^{c_j} l_k; Gauss-Legendre nodes/weights) and the
  reference's own printed coefficient tables (fixture
  ``tests/golden/reference_tables.json``, extracted from the reference's text).
* ``families`` -- hand-derived closed forms of f, g, r and their Jacobians for
  the linear-quadratic and tanh OCP families, checked by central finite
  differences and the Hamiltonian identities in the tests.
* ``solver``/``babd`` -- the condensed BABD direction equals the solution of the
  fully assembled Newton matrix (np.linalg.solve) for small meshes; the Newton
  solve of the LQ problem equals the matrix-exponential closed form; the
  partition QR satisfies Q^T Q = I and the BABD residual invariant.
* ``solver.line_search`` / ``newton_solve`` / ``collocation_solver`` (the
  driver) -- hand-made merit sequences (strict decrease, 20 trials, first
  decrease taken); a scalar BVP whose Newton iterates and step halvings follow
  a scalar closed form; an ascent direction that fails all 20 trials and
  leaves the state unchanged; the alpha continuation 0.1 * 0.9^k down to
  1e-6 on a one-inequality OCP whose constrained optimum is known in closed
  form; remeshing of a stiff LQ problem until the segment residual is <= 1.
* ``solver.compute_segment_residual`` -- hand-computed tiny cases, including
  the reference's in-loop time-scaling quirk.
* ``solver.remesh`` -- hand-built residual patterns.
* ``mlp`` (the warm-start network, no reference counterpart) -- central finite
  differences of the loss, GELU's closed forms, torch's bf16 rounding.
"""
