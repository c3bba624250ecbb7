"""Pins of the oracle's penalty, coefficient weighting and V-cycle
composition by what the mathematics fixes, independent of the oracle's own
formulas.

* Coefficient consistency across an interface (sipg.py:234-280 with a
  per-element kappa): for kappa = 1 on x < 1/2 and 3 on x > 1/2 and the
  piecewise-linear u with continuous flux kappa du/dx, SIPG is consistent,
  so the discrete solution is u itself to round-off.  A face that took the
  neighbour's (or the same) kappa on both normal derivatives of the flux
  average would leave a consistency error of O(1).
* Penalty closed form (sipg.py:255-263, 276-280): for u constant on each
  element every gradient vanishes and a(u, u) is the penalty alone,
  sum over element sides of w * tau_scale_k (N+1)^2 kappa_k L^2 / A_k * [c]^2,
  with L the side length and A_k the element area taken from the vertex
  coordinates (sJ = L/2, J = A/2, Fscale = sJ/J, 1^T m1d 1 = 2).  A penalty
  scaled by (N+1), by the neighbour's Fscale, or with other face weights
  than 1/2, 1, 0 fails it.
* V-cycle error propagation (multigrid.py:211-246): the dense I - B A of the
  oracle's V-cycle equals S^nu (I - P B_c P^T A) S^nu with
  S = I - omega D^{-1} A, D the dense block diagonal of A and B_c the coarser
  cycle (A_c^{-1} at the coarsest level), built from dense matrices alone.
"""
import numpy as np
import pytest

from oracle import mesh as om
from oracle import multigrid as omg
from oracle import sipg as S


def _tag(mesh, dir_x=True, dir_y=True):
    """Boundary faces by the wall they lie on: Dirichlet or Neumann."""
    E = np.asarray(mesh.EToE)
    bdy = E == np.arange(mesh.K)[:, None]
    nx = mesh.nx.reshape((mesh.Nfp, 3, mesh.K), order="F")[0].T
    on_x = np.abs(nx) > 0.5
    code = np.where(on_x, 1 if dir_x else 2, 1 if dir_y else 2)
    return np.where(bdy, code, 0).astype(np.int32)


def _mesh(n, N, jitter_y=0.0, seed=0):
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, n, n)
    VY = VY.copy()
    if jitter_y:
        rng = np.random.default_rng(seed)
        inner = (VY > 1e-12) & (VY < 1 - 1e-12)
        VY[inner] += rng.uniform(-jitter_y, jitter_y, inner.sum()) / n
    return om.build_mesh_2d(VX, VY, EToV, N)


def _centroid_x(mesh):
    return np.asarray(mesh.VX)[np.asarray(mesh.EToV)].mean(axis=1)


@pytest.mark.parametrize("N", [1, 2, 3])
@pytest.mark.parametrize("dir_y", [True, False])
def test_kappa_interface_consistency(N, dir_y):
    m = _mesh(4, N, jitter_y=0.15, seed=N)
    kap = np.where(_centroid_x(m) < 0.5, 1.0, 3.0)
    b = 0.7
    g = lambda x: np.where(x < 0.5, x, 0.5 + (x - 0.5) / 3.0)
    ue = g(m.x) + b * m.y
    bc = _tag(m, dir_x=True, dir_y=dir_y)
    A = S.sipg_matrix_2d(m, bc_types=bc, kappa=kap)
    rhs = -S.sipg_apply_2d(m, np.zeros_like(ue), bc_types=bc, u_dir=ue, kappa=kap)
    if not dir_y:
        # Neumann flux data kappa * du/dn on the y walls (du/dy = b)
        ny = m.ny.reshape((m.Nfp, 3, m.K), order="F")
        q = (kap[None, None, :] * b * ny).reshape((3 * m.Nfp, m.K), order="F")
        rhs = rhs + S.sipg_neumann_load_2d(m, q, bc_types=bc)
    u = np.linalg.solve(A, rhs.T.reshape(-1)).reshape(m.K, m.Np).T
    assert np.abs(u - ue).max() < 1e-10 * np.abs(ue).max()


def test_kappa_interface_consistency_fails_for_a_uniform_kappa():
    """The same data with kappa = 1 everywhere is not a solution: the check
    has the power to see the coefficient."""
    m = _mesh(4, 2)
    g = lambda x: np.where(x < 0.5, x, 0.5 + (x - 0.5) / 3.0)
    ue = g(m.x)
    bc = _tag(m)
    A = S.sipg_matrix_2d(m, bc_types=bc)
    rhs = -S.sipg_apply_2d(m, np.zeros_like(ue), bc_types=bc, u_dir=ue)
    u = np.linalg.solve(A, rhs.T.reshape(-1)).reshape(m.K, m.Np).T
    assert np.abs(u - ue).max() > 1e-3


def _penalty_closed_form(VX, VY, EToV, N, bc, c, tau, kap):
    """sum_k sum_f w tau_k (N+1)^2 kappa_k L^2 / A_k [c]^2 from vertices."""
    K = EToV.shape[0]
    P = np.stack([VX, VY], axis=1)
    area = np.empty(K)
    for k in range(K):
        a, b_, c_ = P[EToV[k]]
        area[k] = 0.5 * abs((b_[0] - a[0]) * (c_[1] - a[1]) - (b_[1] - a[1]) * (c_[0] - a[0]))
    owner = {}
    for k in range(K):
        for f in range(3):
            key = tuple(sorted((int(EToV[k, f]), int(EToV[k, (f + 1) % 3]))))
            owner.setdefault(key, []).append(k)
    total = 0.0
    for k in range(K):
        for f in range(3):
            v0, v1 = int(EToV[k, f]), int(EToV[k, (f + 1) % 3])
            L = np.hypot(*(P[v1] - P[v0]))
            code = bc[k, f]
            if code == 2:
                continue
            if code == 1:
                w, jump = 1.0, c[k]
            else:
                nb = [j for j in owner[tuple(sorted((v0, v1)))] if j != k]
                w, jump = 0.5, c[k] - c[nb[0]]
            total += w * tau[k] * (N + 1) ** 2 * kap[k] * L * L / area[k] * jump * jump
    return total


@pytest.mark.parametrize("N", [1, 3, 5])
def test_penalty_closed_form(N):
    rng = np.random.default_rng(10 + N)
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 2.0, 3, 4)
    h = 0.5
    inner = (VX > 1e-12) & (VX < 1 - 1e-12) & (VY > 1e-12) & (VY < 2 - 1e-12)
    VX = VX + np.where(inner, rng.uniform(-0.15 * h, 0.15 * h, VX.size), 0.0)
    VY = VY + np.where(inner, rng.uniform(-0.15 * h, 0.15 * h, VY.size), 0.0)
    m = om.build_mesh_2d(VX, VY, EToV, N)
    c = rng.standard_normal(m.K)
    u = np.broadcast_to(c, (m.Np, m.K)).copy()
    tau = 200.0 * (1.0 + rng.random(m.K))
    kap = np.exp(0.5 * rng.standard_normal(m.K))
    for bc in (_tag(m), _tag(m, dir_y=False), _tag(m, dir_x=False, dir_y=False)):
        want = _penalty_closed_form(VX, VY, np.asarray(EToV), N, bc, c, tau, kap)
        got = S.sipg_bilinear_form(m, u, u, bc_types=bc, tau_scale=tau, kappa=kap)
        assert abs(got - want) <= 1e-12 * abs(want)
        # the apply is the gradient of the same form: u^T A u
        y = S.sipg_apply_2d(m, u, bc_types=bc, tau_scale=tau, kappa=kap)
        assert abs(np.sum(u * y) - want) <= 1e-11 * abs(want)
    # scalar tau_scale broadcasts the same way
    bc = _tag(m)
    want = _penalty_closed_form(VX, VY, np.asarray(EToV), N, bc, c, np.full(m.K, 37.0),
                                np.ones(m.K))
    got = S.sipg_bilinear_form(m, u, u, bc_types=bc, tau_scale=37.0)
    assert abs(got - want) <= 1e-12 * abs(want)


# ---------------------------------------------------------------------------
# V-cycle error propagation
# ---------------------------------------------------------------------------
def _dense_of(fn, n):
    out = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        out[:, j] = fn(e)
        e[j] = 0.0
    return out


def _block_diag(A, Np, K):
    D = np.zeros_like(A)
    for k in range(K):
        s = slice(k * Np, (k + 1) * Np)
        D[s, s] = A[s, s]
    return D


def _p_dense(I, K):
    """Element-major prolongation of the per-element (Np_hi, Np_lo) map."""
    return np.kron(np.eye(K), I)


def _h_dense(Ic, Kc):
    """Fine element 4k+c is child c of coarse element k."""
    Np = Ic.shape[1]
    P = np.zeros((4 * Kc * Np, Kc * Np))
    for k in range(Kc):
        for c in range(4):
            P[(4 * k + c) * Np:(4 * k + c + 1) * Np, k * Np:(k + 1) * Np] = Ic[c]
    return P


def _error_propagator(As, Ps, nu, omega, Nps, Ks):
    """E_0 of the V-cycle from dense level matrices (fine to coarse)."""
    B = np.linalg.inv(As[-1])
    for l in range(len(As) - 2, -1, -1):
        A = As[l]
        n = A.shape[0]
        Sm = np.eye(n) - omega * np.linalg.solve(_block_diag(A, Nps[l], Ks[l]), A)
        Snu = np.linalg.matrix_power(Sm, nu)
        E = Snu @ (np.eye(n) - Ps[l] @ B @ Ps[l].T @ A) @ Snu
        B = (np.eye(n) - E) @ np.linalg.inv(A)
    return E


@pytest.mark.parametrize("nu,omega,alpha", [(2, 0.7, 0.0), (1, 0.6, 5.0)])
def test_pmg_vcycle_error_propagation(nu, omega, alpha):
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 2, 2)
    m = om.build_mesh_2d(VX, VY, EToV, 3)
    bc = _tag(m, dir_y=False)
    mg = omg.make_sipg_mg_preconditioner(m, bc_types=bc, alpha=alpha, nu_smooth=nu,
                                         omega=omega)
    degrees = [3, 2, 1]
    meshes = [om.build_mesh_2d(VX, VY, EToV, d) for d in degrees]
    As = [S.sipg_matrix_2d(mm, bc_types=bc, alpha=alpha) for mm in meshes]
    Ps = [_p_dense(omg.p_interp_matrix(meshes[l + 1], meshes[l]), m.K) for l in range(2)]
    E = _error_propagator(As, Ps, nu, omega, [mm.Np for mm in meshes], [m.K] * 3)
    n = m.Np * m.K
    Bd = _dense_of(mg, n)
    got = np.eye(n) - Bd @ As[0]
    assert np.abs(got - E).max() < 1e-9
    assert np.abs(E).max() > 1e-2          # a non-trivial propagator


def test_hpmg_vcycle_error_propagation():
    VX, VY, EToV = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 1, 1)
    mg, finest = omg.make_sipg_hpmg_preconditioner(VX, VY, EToV, 2, 2, bc_fn=_tag)
    hier = omg.forest_hierarchy(VX, VY, EToV, 2)
    fVX, fVY, fE = hier[-1]
    p_meshes = [om.build_mesh_2d(fVX, fVY, fE, d) for d in (2, 1)]
    h_meshes = [om.build_mesh_2d(hx, hy, hE, 1) for hx, hy, hE in hier[-2::-1]]
    meshes = p_meshes + h_meshes                       # K: 32, 32, 8, 2
    As = [S.sipg_matrix_2d(mm, bc_types=_tag(mm)) for mm in meshes]
    Ps = [_p_dense(omg.p_interp_matrix(p_meshes[1], p_meshes[0]), 32)]
    Ic = omg.h_child_interp_matrices(p_meshes[1])
    Ps += [_h_dense(Ic, meshes[l + 1].K) for l in (1, 2)]
    E = _error_propagator(As, Ps, 2, 0.7, [mm.Np for mm in meshes], [mm.K for mm in meshes])
    n = finest.Np * finest.K
    got = np.eye(n) - _dense_of(mg, n) @ As[0]
    assert np.abs(got - E).max() < 1e-9
    assert np.abs(E).max() > 1e-2
