"""Pins of the oracle's INS solve sequence (oracle/ins.py, restating
incompressible_ns.py:669-807) by closed-form solutions.

SIPG is consistent, so a polynomial of degree <= N with its exact Dirichlet
values and Neumann fluxes is reproduced to round-off; a load with the wrong
sign or a missing mass factor is not.  In the pure-Neumann pressure case the
reference shifts the singular operator by eps = 1e-6 |A s| / |s| (s the
mean-zero sin(idx) probe), so the solution x of the shifted system obeys
A (x - p) = -eps x exactly, with p the mean-zero exact field.
"""
import numpy as np
import pytest

from conftest import unit_square
from ddg_amd import mesh as pm
from oracle import ins
from oracle import sipg as S


def _poly(m):
    x, y = m.x, m.y
    u = x * x - 0.5 * x * y + 0.3 * y ** 3
    ux, uy = 2 * x - 0.5 * y, -0.5 * x + 0.9 * y * y
    lap = 2.0 + 1.8 * y
    return u, ux, uy, lap


def _flux(m, ux, uy):
    """du/dn at the face nodes, (3 Nfp, K)."""
    g = lambda a: a.reshape(-1, order="F")[m.vmapM]
    nx, ny = m.nx.reshape(-1, order="F"), m.ny.reshape(-1, order="F")
    return (nx * g(ux) + ny * g(uy)).reshape((3 * m.Nfp, m.K), order="F")


def _walls(m, neumann_right):
    bdy = pm.boundary_faces(m)
    cx, _ = pm.face_centres(m)
    neu = bdy & (cx > 1 - 1e-9) if neumann_right else np.zeros_like(bdy)
    return np.where(neu, 2, np.where(bdy, 1, 0)).astype(np.int32)


@pytest.mark.parametrize("iterative", [None, S.IterativeConfig(tol=1e-13, maxiter=5000)])
def test_pressure_with_dirichlet_reproduces_polynomial(iterative):
    _, m = unit_square(4, 4, 3, jitter=0.1, seed=3)
    p, px, py, lap = _poly(m)
    bc = _walls(m, neumann_right=True)
    x = ins.pressure_solve(m, -lap, bc_types=bc, q_neu=_flux(m, px, py), u_dir=p,
                           iterative=iterative)
    assert np.abs(x - p).max() < 1e-9 * np.abs(p).max()


def test_pure_neumann_pressure_solves_the_shifted_system():
    _, m = unit_square(4, 4, 3, jitter=0.1, seed=4)
    p, px, py, lap = _poly(m)
    bc = np.where(pm.boundary_faces(m), 2, 0).astype(np.int32)
    x = ins.pressure_solve(m, -lap, bc_types=bc, q_neu=_flux(m, px, py), u_dir=None)
    A = S.sipg_matrix_2d(m, bc_types=bc)
    xf = x.T.reshape(-1)
    pf = p.T.reshape(-1)
    pf = pf - pf.mean()
    res = A @ (xf - pf)
    eps = -float(res @ xf) / float(xf @ xf)
    assert np.abs(res + eps * xf).max() < 1e-9 * np.abs(A @ pf).max()
    s = np.sin(np.arange(pf.size, dtype=np.float64))
    s -= s.mean()
    assert eps == pytest.approx(1e-6 * np.linalg.norm(A @ s) / np.linalg.norm(s), rel=1e-6)


@pytest.mark.parametrize("alpha", [0.5, 150.0])
def test_viscous_reproduces_polynomial_and_warm_start(alpha):
    _, m = unit_square(4, 4, 3, jitter=0.1, seed=5)
    u, ux, uy, lap = _poly(m)
    bc = _walls(m, neumann_right=True)
    kw = dict(bc_types=bc, u_dir=u, q_neu=_flux(m, ux, uy), alpha=alpha, nu=0.01)
    lu = ins.viscous_solve(m, alpha * u - lap, **kw)
    assert np.abs(lu - u).max() < 1e-9 * np.abs(u).max()
    it = S.IterativeConfig(tol=1e-12, maxiter=5000)
    cold, ic = ins.viscous_solve(m, alpha * u - lap, iterative=it, return_info=True, **kw)
    warm, iw = ins.viscous_solve(m, alpha * u - lap, iterative=it, x0=lu, return_info=True,
                                 **kw)
    assert np.abs(cold - u).max() < 1e-8 * np.abs(u).max()
    assert np.abs(warm - u).max() < 1e-8 * np.abs(u).max()
    assert iw["iterations"] < ic["iterations"]
