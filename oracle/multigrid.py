"""Oracle: p- and hp-multigrid V-cycle preconditioners (float64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/ddg/dg2d/multigrid.py``:

* level operator ``alpha M + A`` on the element-major flat vector (53-74);
* level = operator + colour-probe block-Jacobi inverse (87-99);
* p-transfer ``I = V_lo(r_hi, s_hi) invV_lo``, prolong ``I u``, restrict
  ``I^T r`` (106-134);
* h-transfer: four red-child interpolation matrices, fine element ``4k+c``
  = child ``c`` of coarse element ``k`` (137-184);
* uniform forest hierarchy (187-204, via ``refinement.py:210-244,
  359-382, 389-455``), restated here as a loop over sorted leaves;
* V-cycle with a dense LU coarsest solve, ``nu`` damped block-Jacobi
  sweeps before and after the coarse correction (211-246);
* ``make_sipg_mg_preconditioner`` (253-294) and
  ``make_sipg_hpmg_preconditioner`` (301-367).

Pinned in ``tests/test_oracle_multigrid.py``: transfers reproduce
polynomials exactly, restriction is the transpose of prolongation, the
forest's children sit where the child maps put them, the V-cycle is a
symmetric positive-definite operator, and the reference's own MG tests
(``test/test_multigrid_2d.py:77-202``: PCG-MG vs direct solve < 1e-7 and
the iteration-count inequalities).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from . import mesh as om
from . import reference_element as re
from . import sipg as osp


class Level:
    """One multigrid level (multigrid.py:77-84)."""

    def __init__(self, mesh, bc_types, tau_scale, alpha, kappa=None):
        self.mesh = mesh
        self.apply = osp.flat_apply(mesh, bc_types, tau_scale, alpha, kappa)
        blocks = osp.extract_block_diag_matfree(self.apply, mesh)
        self.blocks = blocks
        self.binv = np.linalg.inv(blocks)
        self.prolong = None
        self.restrict = None

    def smooth_inv(self, vf):
        K, Np = self.mesh.K, self.mesh.Np
        return np.einsum("kij,kj->ki", self.binv, vf.reshape(K, Np)).reshape(-1)


def p_interp_matrix(mesh_lo, mesh_hi) -> np.ndarray:
    """(Np_hi, Np_lo) nodal interpolation N_lo -> N_hi (multigrid.py:106-114)."""
    return re.vandermonde_2d(mesh_lo.N, mesh_hi.r, mesh_hi.s) @ mesh_lo.invV


def p_prolong(I):
    Np_lo = I.shape[1]
    return lambda f: np.einsum("hl,kl->kh", I, f.reshape(-1, Np_lo)).reshape(-1)


def p_restrict(I):
    Np_hi = I.shape[0]
    return lambda f: np.einsum("hl,kh->kl", I, f.reshape(-1, Np_hi)).reshape(-1)


RED_CHILDREN = (("A", "mAB", "mCA"), ("mAB", "B", "mBC"), ("mCA", "mBC", "C"),
                ("mAB", "mBC", "mCA"))


def h_child_interp_matrices(mesh) -> np.ndarray:
    """(4, Np, Np): parent nodal values -> child c nodal values
    (multigrid.py:137-163)."""
    P = {"A": np.array([-1.0, -1.0]), "B": np.array([1.0, -1.0]),
         "C": np.array([-1.0, 1.0])}
    P["mAB"] = 0.5 * (P["A"] + P["B"])
    P["mBC"] = 0.5 * (P["B"] + P["C"])
    P["mCA"] = 0.5 * (P["C"] + P["A"])
    r, s = mesh.r, mesh.s
    out = []
    for names in RED_CHILDREN:
        P0, P1, P2 = (P[n] for n in names)
        pr = -(r + s) / 2 * P0[0] + (1 + r) / 2 * P1[0] + (1 + s) / 2 * P2[0]
        ps = -(r + s) / 2 * P0[1] + (1 + r) / 2 * P1[1] + (1 + s) / 2 * P2[1]
        out.append(re.vandermonde_2d(mesh.N, pr, ps) @ mesh.invV)
    return np.stack(out)


def h_prolong(Ic):
    Np = Ic.shape[1]
    return lambda f: np.einsum("cij,kj->kci", Ic, f.reshape(-1, Np)).reshape(-1)


def h_restrict(Ic):
    Np = Ic.shape[1]
    return lambda f: np.einsum("cij,kci->kj", Ic, f.reshape(-1, 4, Np)).reshape(-1)


def forest_hierarchy(VX, VY, EToV, n_levels):
    """Uniform refinement forest (refinement.py:210-244, 359-382, 389-455),
    every leaf refined ``n_levels`` times.  Returns ``[(VX, VY, EToV), ...]``
    from the base mesh to the finest, each level's elements being the sorted
    leaves, so element ``4k+c`` is the ``c``-th red child of element ``k``."""
    vx = [float(v) for v in np.asarray(VX)]
    vy = [float(v) for v in np.asarray(VY)]
    verts = [tuple(int(i) for i in t) for t in np.asarray(EToV)]
    children = [()] * len(verts)
    edge_mid = {}

    def mid(a, b):
        key = (min(a, b), max(a, b))
        if key not in edge_mid:
            edge_mid[key] = len(vx)
            vx.append(0.5 * (vx[a] + vx[b]))
            vy.append(0.5 * (vy[a] + vy[b]))
        return edge_mid[key]

    def active():
        leaves = [i for i in range(len(verts)) if not children[i]]
        tris = np.asarray([verts[i] for i in leaves], dtype=np.int64)
        used = np.unique(tris.reshape(-1))
        remap = -np.ones(len(vx), dtype=np.int64)
        remap[used] = np.arange(used.size)
        return (np.asarray(vx)[used], np.asarray(vy)[used], remap[tris])

    out = [active()]
    for _ in range(n_levels):
        for node in sorted(i for i in range(len(verts)) if not children[i]):
            v0, v1, v2 = verts[node]
            m01, m12, m20 = mid(v0, v1), mid(v1, v2), mid(v2, v0)
            kids = []
            for tri in ((v0, m01, m20), (m01, v1, m12), (m20, m12, v2), (m01, m12, m20)):
                kids.append(len(verts))
                verts.append(tri)
                children.append(())
            children[node] = tuple(kids)
        out.append(active())
    return out


def assemble_vcycle(levels, *, nu_smooth=2, omega=0.7, regularize=False, eps=1e-8):
    """V-cycle preconditioner (multigrid.py:211-246)."""
    coarse = levels[-1]
    n_c = coarse.mesh.Np * coarse.mesh.K
    A_c = np.empty((n_c, n_c))
    e = np.zeros(n_c)
    for j in range(n_c):
        e[j] = 1.0
        A_c[:, j] = coarse.apply(e)
        e[j] = 0.0
    if regularize:
        A_c = A_c + eps * np.max(np.abs(np.diag(A_c))) * np.eye(n_c)
    lu = sla.lu_factor(A_c)
    nlev = len(levels)

    def vcycle(li, b):
        if li == nlev - 1:
            return sla.lu_solve(lu, b)
        lvl = levels[li]
        A, Minv = lvl.apply, lvl.smooth_inv
        x = np.zeros_like(b)
        for _ in range(nu_smooth):
            x = x + omega * Minv(b - A(x))
        e_c = vcycle(li + 1, lvl.restrict(b - A(x)))
        x = x + lvl.prolong(e_c)
        for _ in range(nu_smooth):
            x = x + omega * Minv(b - A(x))
        return x

    pc = lambda v: vcycle(0, np.asarray(v, dtype=np.float64))
    pc.A_coarse = A_c
    return pc


def make_sipg_mg_preconditioner(mesh, *, bc_types, tau_scale=200.0, alpha=0.0, p_coarse=1,
                                nu_smooth=2, omega=0.7, regularize=False, eps=1e-8,
                                degrees=None, kappa=None):
    """p-multigrid (multigrid.py:253-294).  ``degrees`` (fine to coarse, an
    extension; default ``N, N-1, ..., p_coarse`` as in the reference)."""
    p_fine = int(mesh.N)
    if p_coarse < 1 or p_coarse > p_fine:
        raise ValueError(f"p_coarse must be in [1, {p_fine}], got {p_coarse}")
    if degrees is None:
        degrees = list(range(p_fine, p_coarse - 1, -1))
    meshes = {d: (mesh if d == p_fine else om.build_mesh_2d(mesh.VX, mesh.VY, mesh.EToV, d))
              for d in degrees}
    levels = [Level(meshes[d], bc_types, tau_scale, alpha, kappa) for d in degrees]
    for li in range(len(degrees) - 1):
        I = p_interp_matrix(meshes[degrees[li + 1]], meshes[degrees[li]])
        levels[li].prolong = p_prolong(I)
        levels[li].restrict = p_restrict(I)
    pc = assemble_vcycle(levels, nu_smooth=nu_smooth, omega=omega, regularize=regularize,
                         eps=eps)
    pc.levels = levels
    return pc


def make_sipg_hpmg_preconditioner(base_VX, base_VY, base_EToV, N, n_h_levels, *, bc_fn,
                                  tau_scale=200.0, alpha=0.0, p_coarse=1, nu_smooth=2,
                                  omega=0.7, regularize=False, eps=1e-8, degrees=None,
                                  fine_vertices=None):
    """hp-multigrid (multigrid.py:301-367).  The base mesh is given by its
    vertex arrays; the degree-``N`` finest mesh is returned with the
    preconditioner.  ``fine_vertices=(VX, VY)`` optionally replaces the
    finest level's vertex coordinates (a jittered mesh); coarser levels keep
    the same coordinates for the vertices they share (an extension)."""
    hier = forest_hierarchy(base_VX, base_VY, base_EToV, n_h_levels)
    if fine_vertices is not None:
        FX, FY = (np.asarray(a, dtype=np.float64) for a in fine_vertices)
        hier = [(FX[:vx.size].copy(), FY[:vx.size].copy(), E) for vx, _, E in hier]
    fVX, fVY, fE = hier[-1]
    finest = om.build_mesh_2d(fVX, fVY, fE, N)
    if p_coarse < 1 or p_coarse > N:
        raise ValueError(f"p_coarse must be in [1, {N}], got {p_coarse}")
    if degrees is None:
        degrees = list(range(N, p_coarse - 1, -1))
    bc_fine = bc_fn(finest)
    p_meshes = {d: (finest if d == N else om.build_mesh_2d(fVX, fVY, fE, d)) for d in degrees}
    levels = [Level(p_meshes[d], bc_fine, tau_scale, alpha) for d in degrees]
    for hVX, hVY, hE in hier[-2::-1]:
        hm = om.build_mesh_2d(hVX, hVY, hE, p_coarse)
        levels.append(Level(hm, bc_fn(hm), tau_scale, alpha))
    n_p = len(degrees)
    for li in range(n_p - 1):
        I = p_interp_matrix(p_meshes[degrees[li + 1]], p_meshes[degrees[li]])
        levels[li].prolong = p_prolong(I)
        levels[li].restrict = p_restrict(I)
    Ic = h_child_interp_matrices(p_meshes[p_coarse])
    for li in range(n_p - 1, len(levels) - 1):
        levels[li].prolong = h_prolong(Ic)
        levels[li].restrict = h_restrict(Ic)
    pc = assemble_vcycle(levels, nu_smooth=nu_smooth, omega=omega, regularize=regularize,
                         eps=eps)
    pc.levels = levels
    return pc, finest


def pcg_count(A, b, Minv, *, tol=1e-8, maxiter=2000):
    """The iteration-counting PCG of ``test/test_multigrid_2d.py:49-68``
    (stop on ``|r|/|b| < tol`` right after the update)."""
    bnorm = np.linalg.norm(b)
    x = np.zeros_like(b)
    r = b.copy()
    z = Minv(r)
    p = z.copy()
    rz = float(r @ z)
    for it in range(1, maxiter + 1):
        Ap = A(p)
        a = rz / float(p @ Ap)
        x = x + a * p
        r = r - a * Ap
        if np.linalg.norm(r) / bnorm < tol:
            return x, it
        z = Minv(r)
        rz_new = float(r @ z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, maxiter
