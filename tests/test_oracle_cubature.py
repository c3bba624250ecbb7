"""Pins of the oracle's cubature, curved-mesh and cubature-SIPG restatements
(oracle/cubature.py, oracle/curved.py, the ``cub`` branch of
oracle/sipg.py) by what the mathematics fixes:

* the conical rule integrates every monomial r^a s^b with a + b <= degree
  exactly: the integrals over the reference triangle are evaluated in exact
  rational arithmetic (simplex moments i! j! / (i+j+2)!);
* on an affine mesh the cubature data reduce to the nodal ones: cubature
  mass = J M_ref, face cubature mass = sJ m1d, cubature apply = nodal apply;
* on a curved mesh, sum_q w_q J_q is each element's area, which the
  boundary line integral (1/2) int (x dy - y dx) along the element's
  polynomial edges gives independently; sJ at the face points is |dx/dt|
  of the 1D interpolant of the face nodes;
* the hand-written gradient of the cubature form equals the literal form
  on unit vectors; the cubature matrix is symmetric positive definite;
* the reference's curved-mesh checks (test/test_curved_boundary_2d.py:62-95)
  and a manufactured solution on the curved annulus.
"""
from fractions import Fraction
from math import comb, factorial

import numpy as np
import pytest

from oracle import cubature as oc
from oracle import curved as ocv
from oracle import mesh as om
from oracle import reference_element as re
from oracle import sipg as S


def _exact_triangle_moment(a, b):
    """int over {r, s >= -1, r + s <= 0} of r^a s^b, exactly."""
    tot = Fraction(0)
    for i in range(a + 1):
        for j in range(b + 1):
            c = comb(a, i) * comb(b, j) * 2 ** (i + j) * (-1) ** (a - i + b - j)
            tot += c * Fraction(factorial(i) * factorial(j), factorial(i + j + 2))
    return 4 * tot


@pytest.mark.parametrize("degree", [1, 2, 5, 8, 14, 22])
def test_triangle_cubature_is_exact_to_its_degree(degree):
    r, s, w = oc.triangle_cubature(degree)
    assert abs(w.sum() - 2.0) < 1e-14
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            want = float(_exact_triangle_moment(a, b))
            assert abs(np.sum(w * r ** a * s ** b) - want) < 1e-13, (a, b)


def test_gauss_jacobi_10_rule():
    for n in (1, 3, 7):
        x, w = oc._gauss_jacobi_10(n)
        for k in range(2 * n):
            m = lambda p: 0.0 if p % 2 else 2.0 / (p + 1)
            assert abs(np.sum(w * x ** k) - (m(k) - m(k + 1))) < 1e-14


def _annulus(N, n=8, curved=True):
    VX, VY, EToV = om.rectangle_with_square_hole_mesh_2d(-2, 2, -2, 2, n, n, -0.5, 0.5, -0.5, 0.5)
    m = om.build_mesh_2d(VX, VY, EToV, N)
    return ocv.apply_circular_hole_2d(m, cx=0.0, cy=0.0, radius=0.5, detection_tol=0.5) \
        if curved else m


def _bc_walls(m):
    bdy = np.asarray(m.EToE) == np.arange(m.K)[:, None]
    return np.where(bdy, 1, 0).astype(np.int32)


@pytest.mark.parametrize("N", [1, 3, 4])
def test_affine_mesh_cubature_equals_nodal(N):
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0, 1, 0, 1, 3, 3)
    rng = np.random.default_rng(N)
    inner = (VX > 0) & (VX < 1) & (VY > 0) & (VY < 1)
    VX = VX + np.where(inner, rng.uniform(-0.05, 0.05, VX.size), 0)
    m = om.build_mesh_2d(VX, VY, EToV, N)
    cub = oc.attach_cubature(m)
    M = S.local_mass(m)
    Mk = oc.per_element_mass(cub)
    assert np.abs(Mk - m.J[0][:, None, None] * M[None]).max() < 1e-13 * np.abs(Mk).max()
    np.testing.assert_allclose(cub.J, np.broadcast_to(m.J[0], cub.J.shape), rtol=1e-13)
    sJ = m.sJ.reshape((m.Nfp, 3, m.K), order="F")[0]
    mf = S.face_cub_mass(cub)
    want = sJ[:, None, None, :] * S.face_1d_mass_matrices(m)[:, :, :, None]
    assert np.abs(mf - want).max() < 1e-13 * np.abs(want).max()
    bc = np.where(np.asarray(m.EToE) == np.arange(m.K)[:, None],
                  np.where(np.abs(m.nx.reshape((m.Nfp, 3, m.K), order="F")[0].T) > 0.5, 1, 2),
                  0).astype(np.int32)
    u = rng.standard_normal((m.Np, m.K))
    ud = rng.standard_normal((m.Np, m.K))
    y0 = S.sipg_apply_2d(m, u, bc_types=bc, u_dir=ud)
    y1 = S.sipg_apply_2d(m, u, bc_types=bc, u_dir=ud, cub=cub)
    assert np.abs(y1 - y0).max() < 1e-12 * np.abs(y0).max()
    np.testing.assert_allclose(S.mass_apply(m, u, cub=cub), S.mass_apply(m, u), rtol=0,
                               atol=1e-13 * np.abs(S.mass_apply(m, u)).max())


def _edge_poly(t_nodes, vals, N):
    return np.polynomial.polynomial.Polynomial.fit(t_nodes, vals, N, domain=[-1, 1], window=[-1, 1])


@pytest.mark.parametrize("N", [2, 4])
def test_curved_element_area_and_face_jacobian(N):
    m = _annulus(N)
    cub = oc.attach_cubature(m)
    area_cub = np.sum(cub.w[:, None] * cub.J, axis=0)
    Fm = np.asarray(m.Fmask)
    t_nodes = (m.r[Fm[:, 0]], -m.r[Fm[:, 1]], -m.s[Fm[:, 2]])
    tg, wg = np.polynomial.legendre.leggauss(2 * N + 2)
    moved = np.abs(m.J - m.J[0]).max(axis=0) > 1e-9
    assert moved.sum() > 0
    for k in range(m.K):
        area = 0.0
        for f in range(3):
            px = _edge_poly(t_nodes[f], m.x[Fm[:, f], k], N)
            py = _edge_poly(t_nodes[f], m.y[Fm[:, f], k], N)
            area += 0.5 * np.sum(wg * (px(tg) * py.deriv()(tg) - py(tg) * px.deriv()(tg)))
            tq = np.polynomial.legendre.leggauss(len(cub.face_w))[0]
            sj = np.hypot(px.deriv()(tq), py.deriv()(tq))
            assert np.abs(cub.face_sJ[f, :, k] - sj).max() < 1e-12 * sj.max()
        assert abs(area_cub[k] - area) < 1e-12 * abs(area)


def test_reference_curved_mesh_checks():
    """test_curved_boundary_2d.py:62-95 at N=3 on a 16x16 grid."""
    aff, cur = _annulus(3, 16, curved=False), _annulus(3, 16)
    assert np.abs(cur.x - aff.x).max() > 1e-3
    xb = cur.x.T.reshape(-1)[cur.vmapB]
    yb = cur.y.T.reshape(-1)[cur.vmapB]
    rb = np.hypot(xb, yb)
    np.testing.assert_allclose(rb[rb < 1.0], 0.5, atol=1e-10)
    assert cur.J.min() > 0


@pytest.mark.parametrize("use_cub", [False, True])
def test_curved_apply_is_gradient_of_literal_form(use_cub):
    m = _annulus(2, 8)
    cub = oc.attach_cubature(m) if use_cub else None
    bc = _bc_walls(m)
    rng = np.random.default_rng(3)
    u, ud = rng.standard_normal((m.Np, m.K)), rng.standard_normal((m.Np, m.K))
    y = S.sipg_apply_2d(m, u, bc_types=bc, u_dir=ud, cub=cub)
    yy = np.empty_like(y)
    e = np.zeros_like(u)
    for i in range(m.Np):
        for k in range(m.K):
            e[i, k] = 1.0
            yy[i, k] = S.sipg_bilinear_form(m, u, e, bc_types=bc, u_dir=ud, cub=cub)
            e[i, k] = 0.0
    assert np.abs(y - yy).max() <= 1e-12 * np.abs(y).max()


def test_curved_cubature_matrix_symmetric_positive_definite():
    m = _annulus(3, 8)
    cub = oc.attach_cubature(m)
    A = S.sipg_matrix_2d(m, bc_types=_bc_walls(m), cub=cub)
    assert np.abs(A - A.T).max() < 1e-10 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


@pytest.mark.parametrize("use_cub", [False, True])
def test_curved_manufactured_solution(use_cub):
    """test_curved_boundary_2d.py:98-116 (there with 24x24 cells and cub=None)."""
    m = _annulus(3, 16)
    cub = oc.attach_cubature(m) if use_cub else None
    ue = np.sin(np.pi * m.x / 4) * np.sin(np.pi * m.y / 4)
    src = 2 * (np.pi / 4) ** 2 * ue
    u = S.solve_sipg_poisson_2d(m, src, bc_types=_bc_walls(m), u_dir=ue, cub=cub)
    err = np.sqrt(np.sum(m.J * (u - ue) ** 2) / np.sum(m.J * ue ** 2))
    assert err < 2e-2
