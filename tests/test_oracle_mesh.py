"""Oracle and product meshes: the reference's invariants
(test/test_mesh_2d.py:36-109) and field-by-field agreement."""
import numpy as np
import pytest

from conftest import unit_square


def test_jacobian_closed_form():
    pm, om = unit_square(4, 4, 3)
    for m in (pm, om):
        np.testing.assert_allclose(m.J, 0.25 * 0.25 / 4, atol=1e-12)


def test_face_node_matching():
    for m in unit_square(3, 3, 4, jitter=0.1, seed=2):
        xf, yf = m.x.flatten(order="F"), m.y.flatten(order="F")
        d = np.hypot(xf[m.vmapM] - xf[m.vmapP], yf[m.vmapM] - yf[m.vmapP])
        assert d.max() < 1e-13


def test_unit_normals_and_connectivity():
    for m in unit_square(4, 3, 4):
        assert np.abs(np.hypot(m.nx, m.ny) - 1).max() < 1e-13
        for k in range(m.K):
            for f in range(3):
                kp, fp = m.EToE[k, f], m.EToF[k, f]
                if kp == k and fp == f:
                    continue
                assert m.EToE[kp, fp] == k and m.EToF[kp, fp] == f


def test_boundary_node_count():
    for m in unit_square(4, 4, 3):
        assert len(m.mapB) == 4 * (4 + 4 + 4 + 4)


@pytest.mark.parametrize("N", [1, 3, 4, 6])
def test_product_mesh_matches_oracle(N):
    pm, om = unit_square(5, 4, N, jitter=0.1, seed=7)
    for f in ("EToE", "EToF", "vmapM", "vmapP", "mapB", "Fmask"):
        assert np.array_equal(np.asarray(getattr(pm, f)), np.asarray(getattr(om, f))), f
    for f in ("x", "y", "rx", "sx", "ry", "sy", "J", "nx", "ny", "sJ", "Fscale"):
        a, b = getattr(pm, f), getattr(om, f)
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), f
