"""Oracle: the conjugate-gradient loop of ``jax.scipy.sparse.linalg.cg``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference calls JAX's ``cg`` at ``src/ddg/dg2d/sipg.py:604-609, 910-915``.
JAX is a third-party dependency with no pinned version
(``pyproject.toml:30-33`` lists ``"jax"``) and is not installed here, so this
restates its published algorithm (``jax/_src/scipy/sparse/linalg.py``,
``_cg_solve``):

    r0 = b - A x0;  p0 = z0 = M r0;  gamma0 = <r0, z0>
    atol2 = max(tol^2 <b,b>, atol^2)
    while (rs > atol2) and (k < maxiter):      rs = gamma if M is None else <r,r>
        Ap = A p;  alpha = gamma / <p, Ap>
        x += alpha p;  r -= alpha Ap;  z = M r
        gamma' = <r, z>;  p = z + (gamma'/gamma) p;  k += 1

JAX returns ``(x, None)``; this returns ``(x, info)`` with the iteration
count and final squared residual so the tests can compare stopping points.
Parity unpinned: exact iteration counts under XLA's reduction order (only a
JAX run could pin them); solutions are pinned through CG-vs-LU tolerances.
"""

from __future__ import annotations

import numpy as np


def cg(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    if maxiter is None:
        maxiter = 10 * b.size
    ident = M is None
    Mf = (lambda v: v) if ident else M
    bs = float(np.dot(b, b))
    atol2 = max(tol * tol * bs, atol * atol)
    r = b - A(x)
    z = Mf(r)
    p = z
    gamma = float(np.dot(r, z))
    k = 0
    rs = gamma if ident else float(np.dot(r, r))
    while rs > atol2 and k < maxiter:
        Ap = A(p)
        alpha = gamma / float(np.dot(p, Ap))
        x = x + alpha * p
        r = r - alpha * Ap
        z = Mf(r)
        gamma_new = float(np.dot(r, z))
        p = z + (gamma_new / gamma) * p
        gamma = gamma_new
        k += 1
        rs = gamma if ident else float(np.dot(r, r))
    return x, {"iterations": k, "residual2": rs, "converged": rs <= atol2}
