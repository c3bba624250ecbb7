"""Pins for the multigrid oracle (oracle/multigrid.py).

Mathematical pins: transfers reproduce polynomials, restriction is the
transpose of prolongation, forest children sit where the child maps put
them, the V-cycle is symmetric positive definite.  Reference pins: the
assertions of test/test_multigrid_2d.py:77-202 (PCG-MG vs a direct solve,
iteration-count inequalities), at the reference's sizes.
"""

import numpy as np
import pytest

from oracle import mesh as om
from oracle import multigrid as omg
from oracle import reference_element as re
from oracle import sipg as osp


def _mesh(p, K=8):
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, K, K)
    return om.build_mesh_2d(VX, VY, EToV, p)


def _all_dirichlet(mesh):
    E = np.asarray(mesh.EToE)
    return np.where(E == np.arange(mesh.K)[:, None], osp.BC_DIRICHLET,
                    osp.BC_INTERIOR).astype(np.int32)


def _poly(r, s, i, j):
    return r ** i * s ** j


@pytest.mark.parametrize("lo,hi", [(1, 2), (2, 4), (3, 4), (1, 6)])
def test_p_interp_reproduces_polynomials(lo, hi):
    """Reproducing every monomial of degree <= lo fixes I uniquely."""
    m_lo, m_hi = _mesh(lo, 1), _mesh(hi, 1)
    I = omg.p_interp_matrix(m_lo, m_hi)
    assert I.shape == (m_hi.Np, m_lo.Np)
    for i in range(lo + 1):
        for j in range(lo + 1 - i):
            got = I @ _poly(m_lo.r, m_lo.s, i, j)
            np.testing.assert_allclose(got, _poly(m_hi.r, m_hi.s, i, j), atol=1e-13)


@pytest.mark.parametrize("N", [1, 3])
def test_h_child_interp_reproduces_polynomials(N):
    """Reproducing every monomial of degree <= N fixes each child matrix."""
    m = _mesh(N, 1)
    Ic = omg.h_child_interp_matrices(m)
    corners = {"A": (-1, -1), "B": (1, -1), "C": (-1, 1)}
    corners["mAB"], corners["mBC"], corners["mCA"] = (0, -1), (0, 0), (-1, 0)
    for c, names in enumerate(omg.RED_CHILDREN):
        P = np.array([corners[n] for n in names], dtype=float)
        # child node (r,s) -> parent coordinates, by the affine map of the child
        pr = 0.5 * (-(m.r + m.s) * P[0, 0] + (1 + m.r) * P[1, 0] + (1 + m.s) * P[2, 0])
        ps = 0.5 * (-(m.r + m.s) * P[0, 1] + (1 + m.r) * P[1, 1] + (1 + m.s) * P[2, 1])
        for i in range(N + 1):
            for j in range(N + 1 - i):
                got = Ic[c] @ _poly(m.r, m.s, i, j)
                np.testing.assert_allclose(got, _poly(pr, ps, i, j), atol=1e-13)


def test_restriction_is_prolongation_transpose():
    rng = np.random.default_rng(0)
    m4, m2 = _mesh(4, 2), _mesh(2, 2)
    I = omg.p_interp_matrix(m2, m4)
    P, R = omg.p_prolong(I), omg.p_restrict(I)
    u, v = rng.standard_normal(m2.Np * m2.K), rng.standard_normal(m4.Np * m4.K)
    assert np.isclose(v @ P(u), u @ R(v), rtol=1e-13)
    Ic = omg.h_child_interp_matrices(m2)
    Kc = 5
    Ph, Rh = omg.h_prolong(Ic), omg.h_restrict(Ic)
    u, v = rng.standard_normal(Kc * m2.Np), rng.standard_normal(4 * Kc * m2.Np)
    assert np.isclose(v @ Ph(u), u @ Rh(v), rtol=1e-13)


def test_forest_children_geometry():
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 2, 2)
    hier = omg.forest_hierarchy(VX, VY, EToV, 2)
    assert [h[2].shape[0] for h in hier] == [8, 32, 128]
    for (cx, cy, cE), (fx, fy, fE) in zip(hier[:-1], hier[1:]):
        # shared vertices keep their ids and coordinates
        np.testing.assert_array_equal(fx[:cx.size], cx)
        np.testing.assert_array_equal(fy[:cy.size], cy)
        for k in range(cE.shape[0]):
            A, B, C = (np.array([cx[v], cy[v]]) for v in cE[k])
            pos = {"A": A, "B": B, "C": C, "mAB": (A + B) / 2, "mBC": (B + C) / 2,
                   "mCA": (C + A) / 2}
            for c, names in enumerate(omg.RED_CHILDREN):
                got = np.array([[fx[v], fy[v]] for v in fE[4 * k + c]])
                np.testing.assert_allclose(got, np.array([pos[n] for n in names]), atol=1e-15)


def test_vcycle_is_spd():
    mesh = _mesh(3, 4)
    bc = _all_dirichlet(mesh)
    mg = omg.make_sipg_mg_preconditioner(mesh, bc_types=bc)
    rng = np.random.default_rng(3)
    n = mesh.Np * mesh.K
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    assert abs(v @ mg(u) - u @ mg(v)) < 1e-10 * abs(v @ mg(u)) + 1e-12
    assert u @ mg(u) > 0 and v @ mg(v) > 0
    # linear
    np.testing.assert_allclose(mg(2.0 * u + v), 2.0 * mg(u) + mg(v), rtol=1e-10, atol=1e-12)


def test_vcycle_single_level_is_direct_solve():
    mesh = _mesh(2, 2)
    bc = _all_dirichlet(mesh)
    mg = omg.make_sipg_mg_preconditioner(mesh, bc_types=bc, p_coarse=2)
    A = osp.sipg_matrix_2d(mesh, bc_types=bc)
    b = np.random.default_rng(1).standard_normal(mesh.Np * mesh.K)
    np.testing.assert_allclose(A @ mg(b), b, rtol=1e-10, atol=1e-10)


def test_pmg_preconditioned_solve_is_correct():
    """test_multigrid_2d.py:77-96."""
    mesh = _mesh(p=4, K=8)
    bc = _all_dirichlet(mesh)
    A = osp.flat_apply(mesh, bc, 200.0, 0.0)
    b = np.random.default_rng(0).standard_normal(mesh.Np * mesh.K)
    x_direct = np.linalg.solve(osp.sipg_matrix_2d(mesh, bc_types=bc), b)
    mg = omg.make_sipg_mg_preconditioner(mesh, bc_types=bc, alpha=0.0)
    x_mg, iters = omg.pcg_count(A, b, mg, tol=1e-10)
    err = np.linalg.norm(x_mg - x_direct) / np.linalg.norm(x_direct)
    assert err < 1e-7


def test_pmg_beats_block_jacobi_and_is_degree_robust():
    """test_multigrid_2d.py:101-125."""
    pmg, bj = {}, {}
    for p in (2, 5):
        mesh = _mesh(p=p, K=8)
        bc = _all_dirichlet(mesh)
        A = osp.flat_apply(mesh, bc, 200.0, 0.0)
        b = np.random.default_rng(1).standard_normal(mesh.Np * mesh.K)
        bjp = osp.make_block_jacobi_preconditioner(mesh, bc_types=bc)
        mg = omg.make_sipg_mg_preconditioner(mesh, bc_types=bc)
        _, bj[p] = omg.pcg_count(A, b, bjp)
        _, pmg[p] = omg.pcg_count(A, b, mg)
    assert pmg[5] < 0.5 * bj[5]
    assert pmg[5] < 2.0 * pmg[2]
    assert bj[5] > 1.5 * bj[2]


def test_pmg_handles_the_helmholtz_operator():
    """test_multigrid_2d.py:130-150."""
    mesh = _mesh(p=4, K=8)
    bc = _all_dirichlet(mesh)
    alpha = 50.0
    A = osp.flat_apply(mesh, bc, 200.0, alpha)
    b = np.random.default_rng(2).standard_normal(mesh.Np * mesh.K)
    x_direct = np.linalg.solve(osp.sipg_matrix_2d(mesh, bc_types=bc, alpha=alpha), b)
    mg = omg.make_sipg_mg_preconditioner(mesh, bc_types=bc, alpha=alpha)
    bjp = osp.make_block_jacobi_preconditioner(mesh, bc_types=bc)
    x_mg, mg_it = omg.pcg_count(A, b, mg, tol=1e-10)
    _, bj_it = omg.pcg_count(A, b, bjp, tol=1e-10)
    assert np.linalg.norm(x_mg - x_direct) / np.linalg.norm(x_direct) < 1e-7
    assert mg_it <= bj_it


def test_hpmg_preconditioned_solve_is_correct():
    """test_multigrid_2d.py:155-173."""
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 2, 2)
    mg, finest = omg.make_sipg_hpmg_preconditioner(VX, VY, EToV, 3, 2, bc_fn=_all_dirichlet)
    assert finest.K == 128
    bc = _all_dirichlet(finest)
    A = osp.flat_apply(finest, bc, 200.0, 0.0)
    b = np.random.default_rng(4).standard_normal(finest.Np * finest.K)
    x_direct = np.linalg.solve(osp.sipg_matrix_2d(finest, bc_types=bc), b)
    x_mg, it = omg.pcg_count(A, b, mg, tol=1e-10)
    assert np.linalg.norm(x_mg - x_direct) / np.linalg.norm(x_direct) < 1e-7


def test_hpmg_is_mesh_independent():
    """test_multigrid_2d.py:178-202."""
    hp, bj = {}, {}
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 2, 2)
    for n_h in (1, 3):
        mg, finest = omg.make_sipg_hpmg_preconditioner(VX, VY, EToV, 3, n_h,
                                                       bc_fn=_all_dirichlet)
        bc = _all_dirichlet(finest)
        A = osp.flat_apply(finest, bc, 200.0, 0.0)
        b = np.random.default_rng(5).standard_normal(finest.Np * finest.K)
        bjp = osp.make_block_jacobi_preconditioner(finest, bc_types=bc)
        _, hp[n_h] = omg.pcg_count(A, b, mg)
        _, bj[n_h] = omg.pcg_count(A, b, bjp)
    assert hp[3] < 1.6 * hp[1]
    assert bj[3] > 1.8 * bj[1]
    assert hp[3] < 0.3 * bj[3]
