"""Oracle: cubature on the reference triangle and the per-mesh cubature data
of curved elements (float64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/ddg/dg2d/cubature.py``:

* ``triangle_cubature`` (48-65): Stroud conical product of n-point
  Gauss--Legendre in ``a`` and n-point Gauss--Jacobi(1, 0) in ``b``,
  ``n = ceil((degree+1)/2)``, mapped by ``r = (1+a)(1-b)/2 - 1, s = b``;
  weights ``w_a w_b / 2``;
* ``attach_cubature`` (107-221): Lagrange basis and reference derivatives
  at the points, per-element ``x_r, x_s, y_r, y_s, J`` and metrics there,
  and on each face ``N+1``-point Gauss--Legendre with the face-nodal to
  face-cubature interpolation ``face_L1d`` and ``sJ, nx, ny`` at the points
  (face f runs from vertex f to vertex f+1: ``(t,-1), (-t,t), (-1,-t)``);
* ``per_element_mass`` (228-239) and ``cubature_mass_apply`` (242-245).

The Gauss--Jacobi(1, 0) rule is computed here by Golub--Welsch (the
eigenvalues of its Jacobi matrix), not taken from a library.
Pinned in ``tests/test_oracle_cubature.py``: exactness on monomials up to
the degree against the closed-form integrals over the triangle, affine
meshes (cubature mass = J M_ref, cubature apply = nodal apply), and
curved-element areas against the boundary line integral.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import reference_element as re


def _gauss_legendre(n):
    return np.polynomial.legendre.leggauss(n)


def _gauss_jacobi_10(n):
    """n-point Gauss rule for the weight (1 - b) on [-1, 1] (Golub--Welsch on
    the three-term recurrence of the Jacobi polynomials P^(1,0))."""
    a, b = 1.0, 0.0
    k = np.arange(n, dtype=np.float64)
    diag = (b * b - a * a) / ((2 * k + a + b) * (2 * k + a + b + 2))
    kk = np.arange(1, n, dtype=np.float64)
    off = (2.0 / (2 * kk + a + b) * np.sqrt(kk * (kk + a) * (kk + b) * (kk + a + b)
                                            / ((2 * kk + a + b - 1) * (2 * kk + a + b + 1))))
    T = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    x, V = np.linalg.eigh(T)
    mu0 = 2.0 ** (a + b + 1) / (a + b + 1)      # integral of (1-b) over [-1, 1]
    return x, mu0 * V[0] ** 2


def triangle_cubature(degree: int):
    """(r, s, w) of the conical rule exact to total degree ``degree``."""
    n = max(1, (degree + 2) // 2)
    a, wa = _gauss_legendre(n)
    b, wb = _gauss_jacobi_10(n)
    A, B = np.meshgrid(a, b, indexing="ij")
    r = (1.0 + A) * (1.0 - B) / 2.0 - 1.0
    return r.reshape(-1), B.reshape(-1), (np.outer(wa, wb) / 2.0).reshape(-1)


@dataclass(frozen=True)
class MeshCubature:
    """Field-for-field mirror of the reference ``MeshCubature``
    (cubature.py:72-104)."""
    r: np.ndarray
    s: np.ndarray
    w: np.ndarray
    Lq: np.ndarray
    Drq: np.ndarray
    Dsq: np.ndarray
    J: np.ndarray
    rx: np.ndarray
    sx: np.ndarray
    ry: np.ndarray
    sy: np.ndarray
    face_w: np.ndarray
    face_Lq: np.ndarray
    face_L1d: np.ndarray
    face_sJ: np.ndarray
    face_nx: np.ndarray
    face_ny: np.ndarray


FACE_RS = ((lambda t: (t, -np.ones_like(t))), (lambda t: (-t, t)), (lambda t: (-np.ones_like(t), -t)))
FACE_DRS = ((1.0, 0.0), (-1.0, 1.0), (0.0, -1.0))


def attach_cubature(mesh, degree=None, face_degree=None) -> MeshCubature:
    """cubature.py:107-221."""
    N = int(mesh.N)
    degree = 4 * N - 2 if degree is None else degree
    face_degree = 2 * N + 1 if face_degree is None else face_degree
    rq, sq, wq = triangle_cubature(max(degree, 2 * N))
    invV = np.asarray(mesh.invV)
    Lq = re.vandermonde_2d(N, rq, sq) @ invV
    Vr, Vs = re.grad_vandermonde_2d(N, rq, sq)
    Drq, Dsq = Vr @ invV, Vs @ invV
    x, y = np.asarray(mesh.x), np.asarray(mesh.y)
    xr, xs, yr, ys = Drq @ x, Dsq @ x, Drq @ y, Dsq @ y
    J = xr * ys - xs * yr
    n_face = max(1, (face_degree + 2) // 2)
    tq, fw = _gauss_legendre(n_face)
    Fm = np.asarray(mesh.Fmask)
    r, s = np.asarray(mesh.r), np.asarray(mesh.s)
    t_nodes = (r[Fm[:, 0]], -r[Fm[:, 1]], -s[Fm[:, 2]])
    fLq, fL1, fsJ, fnx, fny = [], [], [], [], []
    for f in range(3):
        rf, sf = FACE_RS[f](tq)
        fLq.append(re.vandermonde_2d(N, rf, sf) @ invV)
        fL1.append(re.vandermonde_1d(N, tq) @ np.linalg.inv(re.vandermonde_1d(N, t_nodes[f])))
        Vrf, Vsf = re.grad_vandermonde_2d(N, rf, sf)
        Drf, Dsf = Vrf @ invV, Vsf @ invV
        dr, ds = FACE_DRS[f]
        xt = dr * (Drf @ x) + ds * (Dsf @ x)
        yt = dr * (Drf @ y) + ds * (Dsf @ y)
        sj = np.sqrt(xt * xt + yt * yt)
        fsJ.append(sj)
        fnx.append(yt / sj)
        fny.append(-xt / sj)
    return MeshCubature(r=rq, s=sq, w=wq, Lq=Lq, Drq=Drq, Dsq=Dsq, J=J, rx=ys / J, sx=-yr / J,
                        ry=-xs / J, sy=xr / J, face_w=fw, face_Lq=np.stack(fLq),
                        face_L1d=np.stack(fL1), face_sJ=np.stack(fsJ), face_nx=np.stack(fnx),
                        face_ny=np.stack(fny))


def per_element_mass(cub: MeshCubature) -> np.ndarray:
    """(K, Np, Np) cubature mass (cubature.py:228-239)."""
    return np.einsum("q,qk,qi,qj->kij", cub.w, cub.J, cub.Lq, cub.Lq)


def cubature_mass_apply(cub: MeshCubature, f) -> np.ndarray:
    """Lq^T (w J (Lq f)) (cubature.py:242-245)."""
    return cub.Lq.T @ (cub.w[:, None] * cub.J * (cub.Lq @ f))
