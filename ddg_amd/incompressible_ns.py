"""Drop-in for the two SIPG solves of ``ddg.dg2d.incompressible_ns``.

The reference's incompressible Navier--Stokes time steppers call two
private solves per step (``src/ddg/dg2d/incompressible_ns.py``):

* ``_pressure_solve`` (669-746): the pressure Poisson problem, right-hand
  side ``J M f + L_N - L_D``; on a domain without Dirichlet faces the
  singular operator is shifted by ``1e-6 |A p| / |p|`` (mean-zero
  ``sin(idx)`` probe) and the right-hand side made mean-zero (718-736);
* ``_viscous_solve`` (749-807): the viscous Helmholtz problem
  ``(alpha M + A) u = J M f - L_D + L_N``.

Both take the same keywords as the reference (``A_factored`` for the LU
path, ``iterative`` for CG from ``x0`` with an optional ``preconditioner``)
and run on the device: the right-hand side is one ``ddgx_sipg_rhs`` launch,
the CG is ``ddgx_pcg_solve`` with the operator's stopping rule of
``jax.scipy.sparse.linalg.cg``.  The reference builds an apply closure and
hands it to JAX's ``cg``; here the closure is not needed because the solve
owns the operator.  The time steppers themselves (advection, BDF, velocity
reconstruction) are outside this package.

``return_info=True`` (an extension) also returns the CG information.
"""

from __future__ import annotations

from .sipg import solve_sipg_helmholtz_2d, solve_sipg_poisson_2d


def _pressure_solve(mesh, source, *, bc_types, q_neu, u_dir, A_factored, tau_scale=200.0,
                    iterative=None, x0=None, preconditioner=None, cub=None, return_info=False):
    """Pressure Poisson solve (incompressible_ns.py:669-746)."""
    return solve_sipg_poisson_2d(mesh, source, bc_types=bc_types, u_dir=u_dir, q_neu=q_neu,
                                 tau_scale=tau_scale, A_factored=A_factored,
                                 iterative=iterative, x0=x0, preconditioner=preconditioner,
                                 cub=cub, return_info=return_info)


def _viscous_solve(mesh, source, *, bc_types, u_dir, q_neu, alpha, nu, A_factored,
                   tau_scale=200.0, iterative=None, x0=None, preconditioner=None, cub=None,
                   return_info=False):
    """Viscous Helmholtz solve (incompressible_ns.py:749-807); ``nu`` is
    unused, as in the reference."""
    return solve_sipg_helmholtz_2d(mesh, source, alpha=alpha, nu=nu, bc_types=bc_types,
                                   u_dir=u_dir, q_neu=q_neu, tau_scale=tau_scale,
                                   A_factored=A_factored, iterative=iterative, x0=x0,
                                   preconditioner=preconditioner, cub=cub,
                                   return_info=return_info)
