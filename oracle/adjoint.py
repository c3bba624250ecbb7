"""Oracle: parameter derivatives of the SIPG form and the misfit loss.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference differentiates solves with ``jax.grad`` through
``lax.custom_linear_solve`` (src/ddg/dg2d/sipg.py:604-609, 910-915) and
the apply through the VJP of the form (304-337).  JAX is absent here, so:

* ``form_param_grad`` takes d a(u, v; u_dir)/d kappa_k and d a/d tau_k
  from the literal form ``sipg_bilinear_form`` by unit perturbations.  The
  form is affine in each kappa_k and each tau_k (every term carries one
  kappa factor; the penalty carries tau_k kappa_k), so a unit difference is
  the exact derivative, not an approximation; no closed form is retyped.
* ``misfit_loss`` solves with dense LU, for centred finite differences of
  the loss (``tests/test_oracle_adjoint.py`` pins the adjoint identity
  dL/dtheta = -d a(u, lam)/d theta with it).
"""

from __future__ import annotations

import numpy as np

from . import sipg as S


def form_param_grad(mesh, u, v, *, bc_types, kappa, tau_scale, u_dir=None):
    """(da/dkappa (K,), da/dtau (K,)) of a(u, v; u_dir) at (kappa, tau)."""
    K = mesh.K
    kap = np.asarray(kappa, dtype=np.float64).copy()
    tau = np.broadcast_to(np.asarray(tau_scale, dtype=np.float64), (K,)).copy()

    def a(kp, tp):
        return S.sipg_bilinear_form(mesh, u, v, bc_types=bc_types, u_dir=u_dir,
                                    tau_scale=tp[None, :], kappa=kp)

    a0 = a(kap, tau)
    gk, gt = np.empty(K), np.empty(K)
    for k in range(K):
        e = np.zeros(K)
        e[k] = 1.0
        gk[k] = a(kap + e, tau) - a0
        gt[k] = a(kap, tau + e) - a0
    return gk, gt


def solve_dense(mesh, f, *, bc_types, kappa, tau_scale, u_dir=None, q_neu=None):
    """Dense LU solve of A(kappa, tau) u = J M f - L_D(u_dir) + L_N(q_neu);
    returns u as (Np, K)."""
    K, Np = mesh.K, mesh.Np
    tau = np.broadcast_to(np.asarray(tau_scale, dtype=np.float64), (K,))[None, :]
    n = Np * K
    A = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = S.sipg_apply_2d(mesh, e.reshape(K, Np).T, bc_types=bc_types, tau_scale=tau,
                                  kappa=kappa).T.reshape(-1)
        e[j] = 0.0
    b = S.mass_apply(mesh, f)
    if u_dir is not None:
        b = b - S.sipg_apply_2d(mesh, np.zeros((Np, K)), bc_types=bc_types, u_dir=u_dir,
                                tau_scale=tau, kappa=kappa)
    if q_neu is not None:
        b = b + S.sipg_neumann_load_2d(mesh, q_neu, bc_types=bc_types)
    return np.linalg.solve(A, b.T.reshape(-1)).reshape(K, Np).T, A


def misfit_loss(mesh, f, target, *, bc_types, kappa, tau_scale, u_dir=None, q_neu=None):
    """0.5 (u - target)^T (J M) (u - target) and u."""
    u, _ = solve_dense(mesh, f, bc_types=bc_types, kappa=kappa, tau_scale=tau_scale,
                       u_dir=u_dir, q_neu=q_neu)
    e = u - target
    return 0.5 * float(np.sum(e * S.mass_apply(mesh, e))), u
