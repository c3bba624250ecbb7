"""Oracle: curved boundary faces by node projection with Gordon--Hall
blending (float64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/ddg/dg2d/curved.py``:

* ``blending_weight`` (61-87): 1 on face f, 0 at the opposite vertex;
* ``project_to_curve_2d`` (90-231): boundary faces whose mean node position
  passes the predicate get their face nodes projected; every node of the
  element then moves by ``weight * I_face->vol(offset)`` (1D interpolation
  of the face offsets in the face parameter ``r, r, s`` for faces 0, 1, 2)
  and the face nodes are set back to their projected positions; metrics,
  normals, sJ and Fscale are recomputed from the new nodes.  Connectivity
  and node maps stay those of the affine mesh;
* ``circle_projection`` (234-243), ``apply_circular_hole_2d`` (246-268).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from . import mesh as om
from . import reference_element as re


def blending_weight(r, s, face):
    if face == 0:
        return (1.0 - s) / 2.0
    if face == 1:
        return (r + s + 2.0) / 2.0
    if face == 2:
        return (1.0 - r) / 2.0
    raise ValueError(f"face must be 0, 1, or 2, got {face}")


def project_to_curve_2d(mesh, *, face_on_curve, project_fn, blend_interior=True):
    K, Nfp = mesh.K, mesh.Nfp
    Fm = np.asarray(mesh.Fmask)
    x, y = np.array(mesh.x, dtype=np.float64), np.array(mesh.y, dtype=np.float64)
    xa = np.stack([x[Fm[:, f], :].mean(axis=0) for f in range(3)])
    ya = np.stack([y[Fm[:, f], :].mean(axis=0) for f in range(3)])
    bdy = (np.asarray(mesh.EToE) == np.arange(K)[:, None]).T
    curved = np.asarray(face_on_curve(xa, ya)) & bdy
    x0, y0 = x.copy(), y.copy()
    r, s = np.asarray(mesh.r), np.asarray(mesh.s)
    for f in range(3):
        ks = np.where(curved[f])[0]
        if ks.size == 0:
            continue
        xo, yo = x0[Fm[:, f]][:, ks], y0[Fm[:, f]][:, ks]
        xp, yp = project_fn(xo, yo)
        x[np.ix_(Fm[:, f], ks)] = xp
        y[np.ix_(Fm[:, f], ks)] = yp
        if not blend_interior:
            continue
        wgt = blending_weight(r, s, f)
        t = s if f == 2 else r
        I = re.vandermonde_1d(mesh.N, t) @ np.linalg.inv(re.vandermonde_1d(mesh.N, t[Fm[:, f]]))
        x[:, ks] += wgt[:, None] * (I @ (xp - xo))
        y[:, ks] += wgt[:, None] * (I @ (yp - yo))
        x[np.ix_(Fm[:, f], ks)] = xp
        y[np.ix_(Fm[:, f], ks)] = yp
    rx, sx, ry, sy, J, nx, ny, sJ, Fscale = om.geometry(x, y, np.asarray(mesh.Dr),
                                                        np.asarray(mesh.Ds), Fm)
    return dataclasses.replace(mesh, x=x, y=y, rx=rx, sx=sx, ry=ry, sy=sy, J=J, nx=nx,
                               ny=ny, sJ=sJ, Fscale=Fscale)


def circle_projection(cx, cy, radius):
    def project(x, y):
        dx, dy = x - cx, y - cy
        d = np.sqrt(dx * dx + dy * dy)
        sc = radius / np.where(d > 1e-14, d, 1.0)
        return cx + dx * sc, cy + dy * sc
    return project


def apply_circular_hole_2d(mesh, *, cx, cy, radius, blend_interior=True, detection_tol=0.1):
    def pred(xf, yf):
        return (xf - cx) ** 2 + (yf - cy) ** 2 < (radius * (1 + detection_tol)) ** 2
    return project_to_curve_2d(mesh, face_on_curve=pred,
                               project_fn=circle_projection(cx, cy, radius),
                               blend_interior=blend_interior)
