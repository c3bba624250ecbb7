"""Curved boundary faces: drop-in for ``ddg.dg2d.curved``.

Host-side mesh setup (``src/ddg/dg2d/curved.py``): boundary faces selected
by a predicate on their mean node position get their nodes moved by a
projection onto the curve; every node of such an element then moves by the
Gordon--Hall blend ``weight_f(r, s) * I_f(offset)`` of the face offsets (the
1D interpolant of the offsets in the face parameter -- ``r`` on faces 0 and
1, ``s`` on face 2 -- evaluated at the node), after which the face nodes are
put back exactly on the curve (curved.py:90-231).  The metrics, normals,
surface Jacobians and ``Fscale`` are recomputed from the new nodes; the
connectivity and node maps stay those of the straight-sided mesh.

The result is a ``Mesh2D`` whose geometric factors vary inside the curved
elements.  ``ddg_amd.sipg`` runs such meshes on its general operator
(per-node metrics; cubature with ``cub=``), not on the affine one.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from .mesh import geometric_factors
from .reference import ref_element, vandermonde_1d


def _blend(r, s, face):
    """Gordon--Hall weight: 1 on face ``face``, 0 at the opposite vertex
    (curved.py:61-87)."""
    return ((1.0 - s) / 2.0, (r + s + 2.0) / 2.0, (1.0 - r) / 2.0)[face]


def project_to_curve_2d(mesh, *, face_on_curve, project_fn, blend_interior=True):
    """Bend the boundary faces that ``face_on_curve(x_mean, y_mean)`` (arrays
    (3, K)) selects onto the curve ``project_fn`` (curved.py:90-231)."""
    N, K = int(mesh.N), int(mesh.K)
    ref = ref_element(N)
    Fm = ref.Fmask
    x0 = np.asarray(mesh.x, dtype=np.float64)
    y0 = np.asarray(mesh.y, dtype=np.float64)
    xm = x0[Fm.T].mean(axis=1)                              # (3, K)
    ym = y0[Fm.T].mean(axis=1)
    bdy = (np.asarray(mesh.EToE) == np.arange(K)[:, None]).T
    sel = np.asarray(face_on_curve(xm, ym), dtype=bool) & bdy
    x, y = x0.copy(), y0.copy()
    for f in range(3):
        ks = np.flatnonzero(sel[f])
        if ks.size == 0:
            continue
        xo, yo = x0[Fm[:, f]][:, ks], y0[Fm[:, f]][:, ks]
        xp, yp = (np.asarray(a, dtype=np.float64) for a in project_fn(xo, yo))
        if blend_interior:
            t = ref.s if f == 2 else ref.r
            interp = np.linalg.solve(vandermonde_1d(N, t[Fm[:, f]]).T,
                                     vandermonde_1d(N, t).T).T   # (Np, Nfp)
            w = _blend(ref.r, ref.s, f)[:, None]
            x[:, ks] += w * (interp @ (xp - xo))
            y[:, ks] += w * (interp @ (yp - yo))
        x[np.ix_(Fm[:, f], ks)] = xp
        y[np.ix_(Fm[:, f], ks)] = yp
    rx, sx, ry, sy, J, nx, ny, sJ, Fscale = geometric_factors(x, y, N)
    return dataclasses.replace(mesh, x=x, y=y, rx=rx, sx=sx, ry=ry, sy=sy, J=J, nx=nx,
                               ny=ny, sJ=sJ, Fscale=Fscale)


def circle_projection(cx: float, cy: float, radius: float):
    """Radial projection onto the circle (curved.py:234-243)."""
    def project(x, y):
        dx, dy = np.asarray(x) - cx, np.asarray(y) - cy
        d = np.hypot(dx, dy)
        scale = radius / np.where(d > 1e-14, d, 1.0)
        return cx + dx * scale, cy + dy * scale
    return project


def apply_circular_hole_2d(mesh, *, cx: float, cy: float, radius: float,
                           blend_interior: bool = True, detection_tol: float = 0.1):
    """Bend the boundary faces within ``radius (1 + detection_tol)`` of the
    centre onto the circle (curved.py:246-268)."""
    lim = (radius * (1.0 + detection_tol)) ** 2
    return project_to_curve_2d(
        mesh, face_on_curve=lambda xf, yf: (xf - cx) ** 2 + (yf - cy) ** 2 < lim,
        project_fn=circle_projection(cx, cy, radius), blend_interior=blend_interior)
