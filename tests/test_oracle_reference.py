"""Oracle reference element pinned by the reference's known answers
(test/test_reference.py:18-51, test/test_reference_2d.py:20-59) and by the
product's independent implementation."""
import numpy as np
import pytest

from oracle import reference_element as re


def test_lgl_n4_closed_form():
    r = re.jacobi_gauss_lobatto_nodes(0.0, 0.0, 4)
    np.testing.assert_allclose(r, [-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0], atol=1e-14)


@pytest.mark.parametrize("N", [1, 4, 9, 16])
def test_gauss_weights_sum_to_two(N):
    _, w = re.jacobi_gauss_quadrature(0.0, 0.0, N)
    assert abs(w.sum() - 2.0) < 1e-13


def test_gauss_rule_integrates_monomials():
    # the (N+1)-point rule is exact to degree 2N+1: int x^k = 2/(k+1) for even k
    x, w = re.jacobi_gauss_quadrature(0.0, 0.0, 6)
    for k in range(14):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(w @ x ** k - exact) < 1e-13


def test_legendre_orthonormal():
    N = 6
    x, w = re.jacobi_gauss_quadrature(0.0, 0.0, 2 * N + 1)
    P = np.stack([re.jacobi_p(x, 0.0, 0.0, j) for j in range(N + 1)], 1)
    np.testing.assert_allclose(P.T @ np.diag(w) @ P, np.eye(N + 1), atol=1e-13)


@pytest.mark.parametrize("N", [3, 5, 7, 9])
def test_dr_1d_exact(N):
    r = re.jacobi_gauss_lobatto_nodes(0.0, 0.0, N)
    Dr = re.d_matrix_1d(N, r, re.vandermonde_1d(N, r))
    for p in range(N + 1):
        exp = np.zeros_like(r) if p == 0 else p * r ** (p - 1)
        np.testing.assert_allclose(Dr @ r ** p, exp, atol=1e-11)


def _ref2d(N):
    x, y = re.equilateral_nodes_2d(N)
    r, s = re.xy_to_rs(x, y)
    V = re.vandermonde_2d(N, r, s)
    return r, s, V


def test_triangle_area_via_mass():
    r, s, V = _ref2d(5)
    M = np.linalg.inv(V @ V.T)
    assert abs(np.ones(len(r)) @ M @ np.ones(len(r)) - 2.0) < 1e-10


@pytest.mark.parametrize("N", [2, 3, 4, 6])
def test_d_matrices_exact(N):
    r, s, V = _ref2d(N)
    Dr, Ds = re.d_matrices_2d(N, r, s, V)
    for p in range(N + 1):
        for q in range(N - p + 1):
            f = r ** p * s ** q
            dr = p * r ** max(p - 1, 0) * s ** q if p else 0 * r
            ds = q * r ** p * s ** max(q - 1, 0) if q else 0 * r
            assert np.abs(Dr @ f - dr).max() < 1e-10
            assert np.abs(Ds @ f - ds).max() < 1e-10


def test_face_masks_on_edges():
    N = 5
    r, s, _ = _ref2d(N)
    Fm = re.face_masks_2d(N, r, s)
    assert Fm.shape == (N + 1, 3)
    assert np.abs(s[Fm[:, 0]] + 1).max() < 1e-12
    assert np.abs(r[Fm[:, 1]] + s[Fm[:, 1]]).max() < 1e-12
    assert np.abs(r[Fm[:, 2]] + 1).max() < 1e-12


def test_face_nodes_are_lgl():
    # warp-and-blend places the face nodes at the 1D LGL points
    for N in range(1, 7):
        r, s, _ = _ref2d(N)
        Fm = re.face_masks_2d(N, r, s)
        np.testing.assert_allclose(r[Fm[:, 0]], re.jacobi_gauss_lobatto_nodes(0, 0, N), atol=1e-13)


def test_face_mass_is_1d_mass():
    # inv(V V^T) on LGL nodes integrates products of degree-N interpolants:
    # 1^T m 1 = 2 and x^T m x = 2/3
    for N in range(1, 7):
        x = re.jacobi_gauss_lobatto_nodes(0, 0, N)
        m = re.face_mass_1d(N, x)
        assert abs(np.ones(N + 1) @ m @ np.ones(N + 1) - 2.0) < 1e-12
        assert abs(x @ m @ x - 2.0 / 3.0) < 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_product_reference_element_matches_oracle(N):
    from ddg_amd.reference import ref_element, fmask_canonical
    ref = ref_element(N)
    r, s, V = _ref2d(N)
    Dr, Ds = re.d_matrices_2d(N, r, s, V)
    np.testing.assert_allclose(ref.r, r, atol=1e-14)
    np.testing.assert_allclose(ref.s, s, atol=1e-14)
    np.testing.assert_allclose(ref.V, V, atol=1e-12)
    np.testing.assert_allclose(ref.Dr, Dr, atol=1e-11)
    np.testing.assert_allclose(ref.Ds, Ds, atol=1e-11)
    assert np.array_equal(fmask_canonical(N), re.face_masks_2d(N, r, s))
