"""Oracle SIPG operator pinned against the literal bilinear form, the
reference's invariants and manufactured solutions
(test/test_sipg_poisson_2d.py:112-247, test/test_block_jacobi_preconditioner.py,
test/test_matfree_preconditioner.py)."""
import numpy as np
import pytest

from conftest import unit_square
from oracle import sipg as S
from ddg_amd import mesh as pmesh


def _bcs(m):
    return {"dir": pmesh.all_dirichlet(m), "mixed": pmesh.dirichlet_x_neumann_y(m)}


def _top_neumann(m):
    """Dirichlet on x walls and y=0, Neumann on y=1 (test_sipg_poisson_2d.py:83-102)."""
    bdy = pmesh.boundary_faces(m)
    nx, ny = pmesh.face_normals(m)
    _, cy = pmesh.face_centres(m)
    neu = bdy & (np.abs(ny) > 0.5) & (cy > 0.5)
    return np.where(neu, 2, np.where(bdy, 1, 0)).astype(np.int32)


def _relL2(m, u, ue):
    return np.sqrt(np.sum(m.J * (u - ue) ** 2) / np.sum(m.J * ue ** 2))


@pytest.mark.parametrize("N", [1, 2, 3])
def test_apply_is_gradient_of_literal_form(N):
    _, m = unit_square(3, 2, N, jitter=0.1, seed=1)
    rng = np.random.default_rng(N)
    for bc in _bcs(m).values():
        kw_list = [{}, {"kappa": np.exp(0.5 * rng.standard_normal(m.K)),
                        "tau_scale": 200 * (1 + rng.random(m.K))}]
        for kw in kw_list:
            u = rng.standard_normal((m.Np, m.K))
            ud = rng.standard_normal((m.Np, m.K))
            for udir in (None, ud):
                y = S.sipg_apply_2d(m, u, bc_types=bc, u_dir=udir, **kw)
                yy = np.empty_like(y)
                for i in range(m.Np):
                    for k in range(m.K):
                        e = np.zeros_like(u)
                        e[i, k] = 1.0
                        yy[i, k] = S.sipg_bilinear_form(m, u, e, bc_types=bc, u_dir=udir, **kw)
                assert np.abs(y - yy).max() <= 1e-13 * np.abs(y).max()


def test_matrix_symmetric_and_spd():
    _, m = unit_square(6, 6, 3)
    for name, bc in _bcs(m).items():
        A = S.sipg_matrix_2d(m, bc_types=bc)
        assert np.abs(A - A.T).max() < 1e-10
        assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


def test_pure_neumann_nullspace_is_constants():
    _, m = unit_square(4, 4, 3)
    bc = np.where(pmesh.boundary_faces(m), 2, 0).astype(np.int32)
    y = S.sipg_apply_2d(m, np.ones((m.Np, m.K)), bc_types=bc)
    assert np.abs(y).max() < 1e-9


def test_kappa_constant_scales_operator():
    _, m = unit_square(3, 3, 3)
    u = np.random.default_rng(0).standard_normal((m.Np, m.K))
    bc = pmesh.dirichlet_x_neumann_y(m)
    y1 = S.sipg_apply_2d(m, u, bc_types=bc)
    y3 = S.sipg_apply_2d(m, u, bc_types=bc, kappa=np.full(m.K, 3.0))
    assert np.abs(y3 - 3 * y1).max() <= 1e-13 * np.abs(y3).max()


def test_helmholtz_is_alpha_mass_plus_A():
    _, m = unit_square(3, 3, 3)
    bc = pmesh.all_dirichlet(m)
    A = S.sipg_matrix_2d(m, bc_types=bc)
    H = S.sipg_matrix_2d(m, bc_types=bc, alpha=7.0)
    Mb = np.kron(np.diag(m.J[0]), S.local_mass(m))
    assert np.abs(H - A - 7.0 * Mb).max() <= 1e-12 * np.abs(H).max()


@pytest.mark.parametrize("N,Kx", [(3, 6), (4, 8)])
def test_manufactured_full_dirichlet(N, Kx):
    _, m = unit_square(Kx, Kx, N)
    ue = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    u = S.solve_sipg_poisson_2d(m, 2 * np.pi ** 2 * ue, bc_types=pmesh.all_dirichlet(m))
    assert _relL2(m, u, ue) < 1e-3


def test_manufactured_mixed_homog_neumann():
    _, m = unit_square(6, 6, 3)
    ue = np.sin(np.pi * m.x) * np.cos(np.pi * m.y)
    u = S.solve_sipg_poisson_2d(m, 2 * np.pi ** 2 * ue, bc_types=pmesh.dirichlet_x_neumann_y(m))
    assert _relL2(m, u, ue) < 1e-3


def test_manufactured_nonhomog_neumann():
    _, m = unit_square(8, 8, 3)
    ue = m.x * (1 - m.x) * m.y
    xf = m.x.reshape(-1, order="F")[m.vmapM]
    q = (xf * (1 - xf)).reshape((3 * m.Nfp, m.K), order="F")
    u = S.solve_sipg_poisson_2d(m, 2.0 * m.y, bc_types=_top_neumann(m), q_neu=q)
    assert _relL2(m, u, ue) < 1e-3


def test_harmonic_nonhomog_dirichlet():
    _, m = unit_square(8, 8, 3)
    ue = m.x ** 2 - m.y ** 2
    u = S.solve_sipg_poisson_2d(m, np.zeros_like(ue), bc_types=pmesh.all_dirichlet(m), u_dir=ue)
    assert _relL2(m, u, ue) < 1e-6


@pytest.mark.parametrize("alpha", [1.0, 100.0])
def test_helmholtz_manufactured(alpha):
    _, m = unit_square(8, 8, 3)
    ue = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    u = S.solve_sipg_helmholtz_2d(m, (2 * np.pi ** 2 + alpha) * ue, alpha=alpha, nu=1.0,
                                  bc_types=pmesh.all_dirichlet(m))
    assert _relL2(m, u, ue) < 1e-3


def test_h_convergence_order():
    # rate ~ N+1 pins every term of the operator (a dropped face term breaks it)
    errs = []
    for Kx in (4, 8):
        _, m = unit_square(Kx, Kx, 2)
        ue = np.exp(m.x) * np.cos(3 * m.y)
        u = S.solve_sipg_poisson_2d(m, 8 * ue, bc_types=pmesh.all_dirichlet(m), u_dir=ue)
        errs.append(_relL2(m, u, ue))
    assert np.log2(errs[0] / errs[1]) > 2.5


def test_cg_matches_lu():
    _, m = unit_square(8, 8, 3)
    ue = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    bc = pmesh.all_dirichlet(m)
    ulu = S.solve_sipg_poisson_2d(m, 2 * np.pi ** 2 * ue, bc_types=bc)
    ucg = S.solve_sipg_poisson_2d(m, 2 * np.pi ** 2 * ue, bc_types=bc,
                                  iterative=S.IterativeConfig(tol=1e-12, maxiter=4000))
    assert np.abs(ucg - ulu).max() < 1e-8


def test_block_jacobi_pcg_matches_lu_and_blocks_match_dense():
    _, m = unit_square(4, 4, 2)
    bc = pmesh.all_dirichlet(m)
    A = S.sipg_matrix_2d(m, bc_types=bc)
    B_dense = S.extract_block_diag_dense(A, m.Np, m.K)
    B_free = S.extract_block_diag_matfree(S.flat_apply(m, bc), m)
    assert np.abs(B_dense - B_free).max() < 1e-11 * np.abs(B_dense).max()
    f = np.random.default_rng(1).standard_normal((m.Np, m.K))
    ulu = S.solve_sipg_poisson_2d(m, f, bc_types=bc)
    pc = S.make_block_jacobi_preconditioner(m, bc_types=bc)
    upc = S.solve_sipg_poisson_2d(m, f, bc_types=bc, preconditioner=pc,
                                  iterative=S.IterativeConfig(tol=1e-11, maxiter=5000))
    assert np.abs(upc - ulu).max() / np.abs(ulu).max() < 1e-7


def test_colouring_valid():
    _, m = unit_square(5, 5, 1)
    col, nc = S.color_elements(m)
    assert nc <= 4
    for k in range(m.K):
        for nb in m.EToE[k]:
            assert nb == k or col[nb] != col[k]
