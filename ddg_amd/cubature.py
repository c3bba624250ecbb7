"""Cubature on the reference triangle and per-mesh cubature data: drop-in for
``ddg.dg2d.cubature``.

Host-side setup (``src/ddg/dg2d/cubature.py``).  ``triangle_cubature`` is the
conical product rule (Gauss--Legendre in ``a`` times Gauss--Jacobi(1, 0) in
``b``, ``r = (1+a)(1-b)/2 - 1, s = b``, weights ``w_a w_b / 2``;
cubature.py:48-65), exact to total degree ``degree``.  ``attach_cubature``
evaluates the basis, its reference derivatives and each element's metrics at
the points, and on every face the ``N+1``-point Gauss--Legendre data
(``face_L1d`` from face-node values to the points, ``sJ``, normals;
cubature.py:107-221).  The ``MeshCubature`` fields and shapes are the
reference's, so either one can be passed as ``cub=`` to ``ddg_amd.sipg``.

For the device operator the per-point data are folded per element
(:func:`volume_weights`: ``w J`` times the metric products) and per face side
(:func:`face_mass`: ``L1d^T diag(w sJ) L1d``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
from scipy.special import roots_jacobi

from .reference import _modes, ref_element, vandermonde_1d


def triangle_cubature(degree: int):
    """``(r, s, w)``, ``ceil((degree+1)/2)^2`` points, weights summing to 2."""
    n = max(1, (degree + 2) // 2)
    a, wa = np.polynomial.legendre.leggauss(n)
    b, wb = roots_jacobi(n, 1.0, 0.0)
    r = np.outer(1.0 + a, 1.0 - b) / 2.0 - 1.0
    s = np.broadcast_to(b, (n, n))
    return r.reshape(-1), np.ascontiguousarray(s).reshape(-1), (np.outer(wa, wb) / 2.0).reshape(-1)


@dataclass(frozen=True)
class MeshCubature:
    """Cubature data of one mesh (cubature.py:72-104)."""
    r: np.ndarray          # (Ncub,)
    s: np.ndarray
    w: np.ndarray
    Lq: np.ndarray         # (Ncub, Np)
    Drq: np.ndarray
    Dsq: np.ndarray
    J: np.ndarray          # (Ncub, K)
    rx: np.ndarray
    sx: np.ndarray
    ry: np.ndarray
    sy: np.ndarray
    face_w: np.ndarray     # (Nfcub,)
    face_Lq: np.ndarray    # (3, Nfcub, Np)
    face_L1d: np.ndarray   # (3, Nfcub, Nfp)
    face_sJ: np.ndarray    # (3, Nfcub, K)
    face_nx: np.ndarray
    face_ny: np.ndarray


def _basis(N, r, s):
    """Lagrange basis and its r, s derivatives at the points (rows)."""
    V, Vr, Vs = _modes(N, r, s)
    invV = ref_element(N).invV
    return V @ invV, Vr @ invV, Vs @ invV


def attach_cubature(mesh, degree: Optional[int] = None,
                    face_degree: Optional[int] = None) -> MeshCubature:
    """Cubature data of ``mesh``: volume rule of degree ``4N-2`` (at least
    ``2N``) and an ``N+1``-point face rule by default (cubature.py:107-221)."""
    N = int(mesh.N)
    ref = ref_element(N)
    degree = 4 * N - 2 if degree is None else int(degree)
    face_degree = 2 * N + 1 if face_degree is None else int(face_degree)
    r, s, w = triangle_cubature(max(degree, 2 * N))
    Lq, Drq, Dsq = _basis(N, r, s)
    x, y = np.asarray(mesh.x, dtype=np.float64), np.asarray(mesh.y, dtype=np.float64)
    xr, xs, yr, ys = Drq @ x, Dsq @ x, Drq @ y, Dsq @ y
    J = xr * ys - xs * yr
    t, fw = np.polynomial.legendre.leggauss(max(1, (face_degree + 2) // 2))
    one = np.ones_like(t)
    # face f runs from vertex f to vertex f+1: (r, s) and d(r, s)/dt
    faces = (((t, -one), (1.0, 0.0)), ((-t, t), (-1.0, 1.0)), ((-one, -t), (0.0, -1.0)))
    Fm = ref.Fmask
    t_nodes = (ref.r[Fm[:, 0]], -ref.r[Fm[:, 1]], -ref.s[Fm[:, 2]])
    fLq, fL1, fsJ, fnx, fny = [], [], [], [], []
    for f, ((rf, sf), (dr, ds)) in enumerate(faces):
        L, Lr, Ls = _basis(N, rf, sf)
        fLq.append(L)
        fL1.append(np.linalg.solve(vandermonde_1d(N, t_nodes[f]).T, vandermonde_1d(N, t).T).T)
        xt = (dr * Lr + ds * Ls) @ x
        yt = (dr * Lr + ds * Ls) @ y
        sj = np.hypot(xt, yt)
        fsJ.append(sj)
        fnx.append(yt / sj)
        fny.append(-xt / sj)
    return MeshCubature(r=r, s=s, w=w, Lq=Lq, Drq=Drq, Dsq=Dsq, J=J, rx=ys / J, sx=-yr / J,
                        ry=-xs / J, sy=xr / J, face_w=fw, face_Lq=np.stack(fLq),
                        face_L1d=np.stack(fL1), face_sJ=np.stack(fsJ), face_nx=np.stack(fnx),
                        face_ny=np.stack(fny))


def per_element_mass(cub) -> np.ndarray:
    """``(K, Np, Np)`` mass ``sum_q w_q J_q L_i L_j`` (cubature.py:228-239)."""
    Lq = np.asarray(cub.Lq)
    wJ = np.asarray(cub.w)[:, None] * np.asarray(cub.J)
    return np.einsum("qi,qk,qj->kij", Lq, wJ, Lq)


def cubature_mass_apply(cub, f) -> np.ndarray:
    """``M_K f`` without forming ``M_K`` (cubature.py:242-245)."""
    Lq = np.asarray(cub.Lq)
    return Lq.T @ (np.asarray(cub.w)[:, None] * np.asarray(cub.J) * (Lq @ np.asarray(f)))


def volume_weights(cub) -> np.ndarray:
    """``(K, 4, Ncub)`` per element and point: ``wJ (rx^2 + ry^2)``,
    ``wJ (rx sx + ry sy)``, ``wJ (sx^2 + sy^2)`` and ``wJ`` -- the volume
    cubature of ``grad u . grad v`` and of the mass (sipg.py:190-201)."""
    w = np.asarray(cub.w)[:, None]
    J = np.asarray(cub.J)
    rx, sx, ry, sy = (np.asarray(a) for a in (cub.rx, cub.sx, cub.ry, cub.sy))
    wJ = w * J
    return np.stack([wJ * (rx * rx + ry * ry), wJ * (rx * sx + ry * sy),
                     wJ * (sx * sx + sy * sy), wJ], axis=0).transpose(2, 0, 1).copy()


def face_mass(cub) -> np.ndarray:
    """``(K, 3, Nfp, Nfp)`` face mass of every element side under the face
    cubature, ``L1d^T diag(face_w sJ_q) L1d`` (sipg.py:281-299 on
    face-node values)."""
    L1 = np.asarray(cub.face_L1d)                                  # (3, Nq, Nfp)
    W = np.asarray(cub.face_w)[None, :, None] * np.asarray(cub.face_sJ)  # (3, Nq, K)
    return np.einsum("fqp,fqk,fqr->kfpr", L1, W, L1)
