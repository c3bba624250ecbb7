"""Oracle: 2D triangle mesh with StartUp2D state (float64, numpy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Literal (loop) restatement of ``src/ddg/dg2d/mesh.py``: the structured
builder (125-156), face connectivity (520-571), node maps (574-627),
geometric factors (634-644), normals (647-684) and ``build_mesh_2d``
(691-760).  The loops make it O(K) Python work; it is meant for the small
meshes of the parity tests.  Pinned by ``test/test_mesh_2d.py:36-109``'s
invariants in ``tests/test_oracle_mesh.py``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import reference_element as re


@dataclass(frozen=True)
class Mesh2D:
    """Field-for-field mirror of the reference ``Mesh2D`` (mesh.py:50-101)."""
    N: int
    Np: int
    K: int
    Nfp: int
    Nfaces: int
    VX: np.ndarray
    VY: np.ndarray
    EToV: np.ndarray
    r: np.ndarray
    s: np.ndarray
    V: np.ndarray
    invV: np.ndarray
    Dr: np.ndarray
    Ds: np.ndarray
    LIFT: np.ndarray
    Fmask: np.ndarray
    x: np.ndarray
    y: np.ndarray
    rx: np.ndarray
    sx: np.ndarray
    ry: np.ndarray
    sy: np.ndarray
    J: np.ndarray
    nx: np.ndarray
    ny: np.ndarray
    sJ: np.ndarray
    Fscale: np.ndarray
    EToE: np.ndarray
    EToF: np.ndarray
    vmapM: np.ndarray
    vmapP: np.ndarray
    vmapB: np.ndarray
    mapB: np.ndarray


def uniform_rectangle_mesh_2d(xmin, xmax, ymin, ymax, Kx, Ky):
    """Two triangles per cell, (v00,v10,v11) then (v00,v11,v01), cells
    row-major in j then i (mesh.py:125-156)."""
    xs = np.linspace(xmin, xmax, Kx + 1)
    ys = np.linspace(ymin, ymax, Ky + 1)
    VX = np.tile(xs, Ky + 1)
    VY = np.repeat(ys, Kx + 1)
    tris = []
    for j in range(Ky):
        for i in range(Kx):
            v00 = j * (Kx + 1) + i
            v10 = v00 + 1
            v01 = v00 + Kx + 1
            v11 = v01 + 1
            tris.append((v00, v10, v11))
            tris.append((v00, v11, v01))
    return VX, VY, np.array(tris, dtype=np.int64)


def _connect(EToV):
    """Face-to-face connectivity by sorted vertex pairs (mesh.py:520-571)."""
    K = EToV.shape[0]
    nv = int(EToV.max()) + 1
    fn = np.vstack([EToV[:, [0, 1]], EToV[:, [1, 2]], EToV[:, [2, 0]]])
    fn = np.sort(fn, axis=1)
    EToE = np.tile(np.arange(K)[:, None], (1, 3))
    EToF = np.tile(np.arange(3)[None, :], (K, 1))
    fid = fn[:, 0] * nv + fn[:, 1]
    elem = np.tile(np.arange(K), 3)
    face = np.repeat(np.arange(3), K)
    order = np.argsort(fid, kind="stable")
    fs, es, ffs = fid[order], elem[order], face[order]
    for i in np.where(fs[:-1] == fs[1:])[0]:
        k1, f1, k2, f2 = es[i], ffs[i], es[i + 1], ffs[i + 1]
        EToE[k1, f1], EToF[k1, f1] = k2, f2
        EToE[k2, f2], EToF[k2, f2] = k1, f1
    return EToE, EToF


def _maps(Np, Nfp, K, Fmask, x, y, EToE, EToF, EToV, VX, VY, tol=1e-10):
    """vmapM/vmapP by nearest-node matching per face pair (mesh.py:574-627)."""
    ids = np.arange(K * Np).reshape(K, Np).T
    vM = np.zeros((Nfp, 3, K), dtype=np.int64)
    vP = np.zeros((Nfp, 3, K), dtype=np.int64)
    for k in range(K):
        for f in range(3):
            vM[:, f, k] = ids[Fmask[:, f], k]
    xf = x.flatten(order="F")
    yf = y.flatten(order="F")
    for k1 in range(K):
        for f1 in range(3):
            k2, f2 = EToE[k1, f1], EToF[k1, f1]
            a, b = EToV[k1, f1], EToV[k1, (f1 + 1) % 3]
            d = np.hypot(VX[a] - VX[b], VY[a] - VY[b])
            iM, iP = vM[:, f1, k1], vM[:, f2, k2]
            D = np.hypot(xf[iM][:, None] - xf[iP][None, :], yf[iM][:, None] - yf[iP][None, :])
            ok = D < tol * d
            for p in range(Nfp):
                hit = np.where(ok[p])[0]
                vP[p, f1, k1] = iP[hit[0]] if hit.size else iM[p]
    vM = vM.flatten(order="F")
    vP = vP.flatten(order="F")
    mapB = np.where(vP == vM)[0]
    return vM, vP, vM[mapB], mapB


def geometry(x, y, Dr, Ds, Fmask):
    """Per-node metrics (mesh.py:634-644), face normals, surface Jacobians
    (mesh.py:647-684) and Fscale = sJ / J at the face nodes."""
    K = x.shape[1]
    Nfp = Fmask.shape[0]
    xr, xs, yr, ys = Dr @ x, Ds @ x, Dr @ y, Ds @ y
    J = xr * ys - xs * yr
    rx, sx, ry, sy = ys / J, -yr / J, -xs / J, xr / J
    Ff = Fmask.flatten(order="F")
    fxr, fxs, fyr, fys = xr[Ff], xs[Ff], yr[Ff], ys[Ff]
    nx = np.zeros((3 * Nfp, K))
    ny = np.zeros((3 * Nfp, K))
    i0, i1, i2 = np.arange(Nfp), np.arange(Nfp, 2 * Nfp), np.arange(2 * Nfp, 3 * Nfp)
    nx[i0], ny[i0] = fyr[i0], -fxr[i0]
    nx[i1], ny[i1] = fys[i1] - fyr[i1], -fxs[i1] + fxr[i1]
    nx[i2], ny[i2] = -fys[i2], fxs[i2]
    sJ = np.sqrt(nx ** 2 + ny ** 2)
    nx, ny = nx / sJ, ny / sJ
    Fscale = sJ / J[Ff]
    return rx, sx, ry, sy, J, nx, ny, sJ, Fscale


def rectangle_with_square_hole_mesh_2d(xmin, xmax, ymin, ymax, Kx, Ky,
                                       hole_xmin, hole_xmax, hole_ymin, hole_ymax):
    """The structured rectangle mesh without the cells whose centre lies in
    the hole box; unused vertices pruned and renumbered (mesh.py:159-206)."""
    xs = np.linspace(xmin, xmax, Kx + 1)
    ys = np.linspace(ymin, ymax, Ky + 1)
    VX = np.tile(xs, Ky + 1)
    VY = np.repeat(ys, Kx + 1)
    tris = []
    for j in range(Ky):
        for i in range(Kx):
            cx, cy = 0.5 * (xs[i] + xs[i + 1]), 0.5 * (ys[j] + ys[j + 1])
            if hole_xmin <= cx <= hole_xmax and hole_ymin <= cy <= hole_ymax:
                continue
            v00 = j * (Kx + 1) + i
            v10, v01 = v00 + 1, v00 + Kx + 1
            tris.append((v00, v10, v01 + 1))
            tris.append((v00, v01 + 1, v01))
    E = np.array(tris, dtype=np.int64)
    used = np.unique(E.reshape(-1))
    remap = -np.ones(VX.size, dtype=np.int64)
    remap[used] = np.arange(used.size)
    return VX[used], VY[used], remap[E]


def build_mesh_2d(VX, VY, EToV, N: int) -> Mesh2D:
    """StartUp2D (mesh.py:691-760)."""
    VX = np.asarray(VX, dtype=np.float64)
    VY = np.asarray(VY, dtype=np.float64)
    EToV = np.asarray(EToV, dtype=np.int64)
    K = EToV.shape[0]
    Nfp = N + 1
    Np = (N + 1) * (N + 2) // 2
    xe, ye = re.equilateral_nodes_2d(N)
    r, s = re.xy_to_rs(xe, ye)
    V = re.vandermonde_2d(N, r, s)
    invV = np.linalg.inv(V)
    Dr, Ds = re.d_matrices_2d(N, r, s, V)
    Fmask = re.face_masks_2d(N, r, s)
    LIFT = re.lift_2d(N, r, s, V, Fmask)
    va, vb, vc = EToV[:, 0], EToV[:, 1], EToV[:, 2]
    x = 0.5 * (np.outer(-(r + s), VX[va]) + np.outer(1 + r, VX[vb]) + np.outer(1 + s, VX[vc]))
    y = 0.5 * (np.outer(-(r + s), VY[va]) + np.outer(1 + r, VY[vb]) + np.outer(1 + s, VY[vc]))
    rx, sx, ry, sy, J, nx, ny, sJ, Fscale = geometry(x, y, Dr, Ds, Fmask)
    EToE, EToF = _connect(EToV)
    vmapM, vmapP, vmapB, mapB = _maps(Np, Nfp, K, Fmask, x, y, EToE, EToF, EToV, VX, VY)
    return Mesh2D(N=N, Np=Np, K=K, Nfp=Nfp, Nfaces=3, VX=VX, VY=VY, EToV=EToV,
                  r=r, s=s, V=V, invV=invV, Dr=Dr, Ds=Ds, LIFT=LIFT, Fmask=Fmask,
                  x=x, y=y, rx=rx, sx=sx, ry=ry, sy=sy, J=J, nx=nx, ny=ny, sJ=sJ,
                  Fscale=Fscale, EToE=EToE, EToF=EToF, vmapM=vmapM, vmapP=vmapP,
                  vmapB=vmapB, mapB=mapB)
