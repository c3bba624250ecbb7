"""Host-side curved-mesh and cubature setup of the product (ddg_amd.mesh,
ddg_amd.curved, ddg_amd.cubature) against the oracle's restatements of the
reference (mesh.py:159-206, curved.py, cubature.py)."""
import numpy as np
import pytest

from ddg_amd import cubature as pc
from ddg_amd import curved as pcv
from ddg_amd import mesh as pm
from oracle import cubature as oc
from oracle import curved as ocv
from oracle import mesh as om
from oracle import sipg as S


def _pair(N, n=8):
    args = (-2.0, 2.0, -2.0, 2.0, n, n, -0.5, 0.5, -0.5, 0.5)
    VX, VY, E = pm.rectangle_with_square_hole_mesh_2d(*args)
    oVX, oVY, oE = om.rectangle_with_square_hole_mesh_2d(*args)
    np.testing.assert_array_equal(VX, oVX)
    np.testing.assert_array_equal(VY, oVY)
    np.testing.assert_array_equal(E, oE)
    m = pcv.apply_circular_hole_2d(pm.build_mesh_2d(VX, VY, E, N), cx=0.0, cy=0.0, radius=0.5,
                                   detection_tol=0.5)
    mo = ocv.apply_circular_hole_2d(om.build_mesh_2d(VX, VY, E, N), cx=0.0, cy=0.0, radius=0.5,
                                    detection_tol=0.5)
    return m, mo


@pytest.mark.parametrize("N", [1, 3, 6])
def test_curved_mesh_matches_oracle(N):
    m, mo = _pair(N)
    for f in ("x", "y", "rx", "sx", "ry", "sy", "J", "nx", "ny", "sJ", "Fscale"):
        a, b = np.asarray(getattr(m, f)), np.asarray(getattr(mo, f))
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), f
    np.testing.assert_array_equal(m.vmapP, mo.vmapP)


@pytest.mark.parametrize("degree", [1, 6, 14, 22])
def test_triangle_cubature_matches_oracle(degree):
    for a, b in zip(pc.triangle_cubature(degree), oc.triangle_cubature(degree)):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


@pytest.mark.parametrize("N", [1, 4, 6])
def test_attach_cubature_matches_oracle(N):
    m, mo = _pair(N)
    cub, cubo = pc.attach_cubature(m), oc.attach_cubature(mo)
    for f in ("r", "s", "w", "Lq", "Drq", "Dsq", "J", "rx", "sx", "ry", "sy", "face_w",
              "face_Lq", "face_L1d", "face_sJ", "face_nx", "face_ny"):
        a, b = np.asarray(getattr(cub, f)), np.asarray(getattr(cubo, f))
        assert a.shape == b.shape, f
        assert np.abs(a - b).max() <= 1e-11 * max(1.0, np.abs(b).max()), f
    vw = pc.volume_weights(cub)
    wJ = cubo.w[:, None] * cubo.J
    np.testing.assert_allclose(vw[:, 3, :], wJ.T, rtol=1e-12)
    np.testing.assert_allclose(vw[:, 1, :], (wJ * (cubo.rx * cubo.sx + cubo.ry * cubo.sy)).T,
                               rtol=1e-11, atol=1e-14)
    fm = pc.face_mass(cub)
    np.testing.assert_allclose(fm, np.transpose(S.face_cub_mass(cubo), (3, 0, 1, 2)),
                               rtol=1e-12, atol=1e-15)
    f = np.random.default_rng(N).standard_normal((m.Np, m.K))
    np.testing.assert_allclose(pc.cubature_mass_apply(cub, f), oc.cubature_mass_apply(cubo, f),
                               rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(pc.per_element_mass(cub), oc.per_element_mass(cubo),
                               rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("N", [2, 5])
def test_element_matrices_match_oracle_cubature(N):
    """Packed per-element cubature stiffness and mass (what the device
    operator reads) against the oracle's volume cubature term and mass."""
    from ddg_amd.operator import element_matrices
    m, mo = _pair(N)
    cub, cubo = pc.attach_cubature(m), oc.attach_cubature(mo)
    em = element_matrices(cub)
    iu = np.triu_indices(m.Np)
    S = _fill(em["estiff"], m.K, m.Np, iu)
    Q = _fill(em["emass"], m.K, m.Np, iu)
    u = np.random.default_rng(N).standard_normal((m.Np, m.K))
    urq, usq = cubo.Drq @ u, cubo.Dsq @ u
    cx = cubo.w[:, None] * cubo.J * (cubo.rx * urq + cubo.sx * usq)
    cy = cubo.w[:, None] * cubo.J * (cubo.ry * urq + cubo.sy * usq)
    vol = cubo.Drq.T @ (cubo.rx * cx + cubo.ry * cy) + cubo.Dsq.T @ (cubo.sx * cx + cubo.sy * cy)
    got = np.einsum("kij,jk->ik", S, u)
    assert np.abs(got - vol).max() <= 1e-11 * np.abs(vol).max()
    np.testing.assert_allclose(Q, oc.per_element_mass(cubo), rtol=1e-11, atol=1e-15)


def _fill(packed, K, Np, iu):
    F = np.zeros((K, Np, Np))
    F[:, iu[0], iu[1]] = packed
    F[:, iu[1], iu[0]] = packed
    return F
