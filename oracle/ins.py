"""Oracle: the incompressible Navier--Stokes solver's two SIPG solves.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates the pressure and viscous solves of
``src/ddg/dg2d/incompressible_ns.py`` as the reference writes them, on top of
the oracle's operator, loads and JAX-semantics CG:

* ``pressure_solve`` -- ``_pressure_solve`` (669-746): right-hand side
  ``J M f + L_N - L_D``; for a domain without Dirichlet faces, a
  ``1e-6 * |A p| / |p|`` shift from the mean-zero ``sin(idx)`` probe and a
  mean-zero right-hand side (718-736); CG from ``x0`` (or zero);
* ``viscous_solve`` -- ``_viscous_solve`` (749-807): right-hand side
  ``J M f - L_D + L_N`` and CG on ``alpha M + A`` from ``x0``.

With ``iterative=None`` both solve the same systems by LU (the reference's
``A_factored`` path).  The operator is pinned in ``tests/test_oracle_*.py``;
what these add is the call sequence, which ``tests/test_oracle_ins.py``
checks against closed-form solutions.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from . import sipg as S
from .cg import cg as jax_cg


def _flat(u):
    return np.asarray(u, dtype=np.float64).T.reshape(-1)


def pressure_solve(mesh, source, *, bc_types, q_neu, u_dir, tau_scale=200.0,
                   iterative=None, x0=None, preconditioner=None, return_info=False):
    """incompressible_ns.py:669-746 (affine branch)."""
    Np, K = mesh.Np, mesh.K
    rhs = S.mass_apply(mesh, source) + S.sipg_neumann_load_2d(mesh, q_neu, bc_types=bc_types)
    if u_dir is not None:
        rhs = rhs - S.sipg_dirichlet_load_2d(mesh, u_dir, bc_types=bc_types,
                                             tau_scale=tau_scale)
    b = _flat(rhs)
    raw = S.flat_apply(mesh, bc_types, tau_scale)
    singular = not np.any(np.asarray(bc_types) == S.BC_DIRICHLET)
    if singular:
        probe = np.sin(np.arange(Np * K, dtype=np.float64))
        probe = probe - probe.mean()
        shift = 1.0e-6 * np.linalg.norm(raw(probe)) / np.linalg.norm(probe)
        matvec = lambda v: raw(v) + shift * v
        b = b - b.mean()
    else:
        matvec = raw
    if iterative is None:
        A = S.sipg_matrix_2d(mesh, bc_types=bc_types, tau_scale=tau_scale)
        if singular:
            A = A + shift * np.eye(A.shape[0])
        x, info = sla.lu_solve(sla.lu_factor(A), b), None
    else:
        xi = np.zeros(Np * K) if x0 is None else _flat(x0)
        x, info = jax_cg(matvec, b, xi, tol=iterative.tol, atol=iterative.atol,
                         maxiter=iterative.maxiter, M=preconditioner)
    p = x.reshape(K, Np).T
    return (p, info) if return_info else p


def viscous_solve(mesh, source, *, bc_types, u_dir, q_neu, alpha, nu, tau_scale=200.0,
                  iterative=None, x0=None, preconditioner=None, return_info=False):
    """incompressible_ns.py:749-807 (affine branch); ``nu`` unused there."""
    Np, K = mesh.Np, mesh.K
    rhs = S.mass_apply(mesh, source) - S.sipg_dirichlet_load_2d(
        mesh, u_dir, bc_types=bc_types, tau_scale=tau_scale)
    if q_neu is not None:
        rhs = rhs + S.sipg_neumann_load_2d(mesh, q_neu, bc_types=bc_types)
    b = _flat(rhs)
    if iterative is None:
        A = S.sipg_matrix_2d(mesh, bc_types=bc_types, tau_scale=tau_scale, alpha=alpha)
        x, info = sla.lu_solve(sla.lu_factor(A), b), None
    else:
        matvec = S.flat_apply(mesh, bc_types, tau_scale, alpha)
        xi = np.zeros(Np * K) if x0 is None else _flat(x0)
        x, info = jax_cg(matvec, b, xi, tol=iterative.tol, atol=iterative.atol,
                         maxiter=iterative.maxiter, M=preconditioner)
    u = x.reshape(K, Np).T
    return (u, info) if return_info else u
