"""2D triangle mesh (``Mesh2D``) and its vectorised builders.

Host-side setup for the SIPG operator.  ``Mesh2D`` has the same fields,
shapes and index conventions as the reference's pytree
(``src/ddg/dg2d/mesh.py:37-118``): volume fields ``(Np, K)``, face fields
``(3*Nfp, K)``, ``vmapM``/``vmapP`` into the element-major flat vector
``u.T.reshape(-1)``.  Any object with those attributes (including the
reference's own ``Mesh2D``) can be handed to :mod:`ddg_amd.sipg`.

The builders are vectorised (the reference's connectivity and node-map
construction, ``mesh.py:520-627``, loops over every face in Python), so a
2x256x256 mesh builds in about a second.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .reference import ref_element

BC_INTERIOR = 0
BC_DIRICHLET = 1
BC_NEUMANN = 2


@dataclass(frozen=True)
class Mesh2D:
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


def uniform_rectangle_mesh_2d(xmin, xmax, ymin, ymax, Kx: int, Ky: int):
    """``2*Kx*Ky`` triangles; cell (i, j) -> elements ``2(j*Kx+i)`` =
    (v00, v10, v11) and ``2(j*Kx+i)+1`` = (v00, v11, v01) (mesh.py:125-156)."""
    xs = np.linspace(xmin, xmax, Kx + 1)
    ys = np.linspace(ymin, ymax, Ky + 1)
    VX = np.tile(xs, Ky + 1)
    VY = np.repeat(ys, Kx + 1)
    j, i = np.meshgrid(np.arange(Ky), np.arange(Kx), indexing="ij")
    v00 = (j * (Kx + 1) + i).reshape(-1)
    v10, v01 = v00 + 1, v00 + Kx + 1
    v11 = v01 + 1
    EToV = np.empty((2 * Kx * Ky, 3), dtype=np.int64)
    EToV[0::2] = np.stack([v00, v10, v11], 1)
    EToV[1::2] = np.stack([v00, v11, v01], 1)
    return VX, VY, EToV


def jitter_vertices(VX, VY, h: float, frac: float, seed: int,
                    bounds=(0.0, 1.0, 0.0, 1.0)):
    """Move every interior vertex by ``U(-frac*h, frac*h)`` per coordinate,
    seeded; boundary vertices stay so the walls stay axis-aligned."""
    xmin, xmax, ymin, ymax = bounds
    rng = np.random.default_rng(seed)
    VX = np.array(VX, dtype=np.float64)
    VY = np.array(VY, dtype=np.float64)
    tol = 1e-12 * max(xmax - xmin, ymax - ymin)
    inner = ((np.abs(VX - xmin) > tol) & (np.abs(VX - xmax) > tol)
             & (np.abs(VY - ymin) > tol) & (np.abs(VY - ymax) > tol))
    d = rng.uniform(-frac * h, frac * h, size=(2, VX.size))
    VX[inner] += d[0, inner]
    VY[inner] += d[1, inner]
    return VX, VY


def _connect(EToV: np.ndarray):
    """Face connectivity by matching sorted vertex pairs (mesh.py:520-571).
    Boundary faces point to themselves."""
    K = EToV.shape[0]
    fv = np.stack([EToV[:, [0, 1]], EToV[:, [1, 2]], EToV[:, [2, 0]]], axis=1)  # (K,3,2)
    fv = np.sort(fv, axis=2).reshape(-1, 2)
    key = fv[:, 0].astype(np.int64) * (int(EToV.max()) + 1) + fv[:, 1]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    EToE = np.repeat(np.arange(K)[:, None], 3, 1)
    EToF = np.repeat(np.arange(3)[None, :], K, 0)
    m = np.where(ks[:-1] == ks[1:])[0]
    a, b = order[m], order[m + 1]
    ka, fa, kb, fb = a // 3, a % 3, b // 3, b % 3
    EToE[ka, fa], EToF[ka, fa] = kb, fb
    EToE[kb, fb], EToF[kb, fb] = ka, fa
    return EToE, EToF


def geometric_factors(x, y, N: int):
    """Per-node metrics ``rx, sx, ry, sy, J`` of the element maps given by
    the nodal coordinates (mesh.py:634-644), outward unit normals and
    surface Jacobians at the face nodes (mesh.py:647-684) and
    ``Fscale = sJ / J`` there.  Works for curved (non-affine) elements."""
    ref = ref_element(N)
    Dr, Ds, Fm, Nfp = ref.Dr, ref.Ds, ref.Fmask, ref.Nfp
    K = x.shape[1]
    xr, xs, yr, ys = Dr @ x, Ds @ x, Dr @ y, Ds @ y
    J = xr * ys - xs * yr
    rx, sx, ry, sy = ys / J, -yr / J, -xs / J, xr / J
    Ff = Fm.flatten(order="F")
    fxr, fxs, fyr, fys = xr[Ff], xs[Ff], yr[Ff], ys[Ff]
    nx = np.empty((3 * Nfp, K))
    ny = np.empty((3 * Nfp, K))
    a0, a1, a2 = slice(0, Nfp), slice(Nfp, 2 * Nfp), slice(2 * Nfp, 3 * Nfp)
    nx[a0], ny[a0] = fyr[a0], -fxr[a0]
    nx[a1], ny[a1] = fys[a1] - fyr[a1], fxr[a1] - fxs[a1]
    nx[a2], ny[a2] = -fys[a2], fxs[a2]
    sJ = np.sqrt(nx * nx + ny * ny)
    nx, ny = nx / sJ, ny / sJ
    return rx, sx, ry, sy, J, nx, ny, sJ, sJ / J[Ff]


def rectangle_with_square_hole_mesh_2d(xmin, xmax, ymin, ymax, Kx: int, Ky: int,
                                       hole_xmin, hole_xmax, hole_ymin, hole_ymax):
    """The structured rectangle mesh of :func:`uniform_rectangle_mesh_2d`
    without the cells whose centre lies in the hole box, unused vertices
    dropped and renumbered in order (mesh.py:159-206).  The hole bounds
    should fall on grid lines."""
    VX, VY, EToV = uniform_rectangle_mesh_2d(xmin, xmax, ymin, ymax, Kx, Ky)
    xs = np.linspace(xmin, xmax, Kx + 1)
    ys = np.linspace(ymin, ymax, Ky + 1)
    cx = 0.5 * (xs[:-1] + xs[1:])
    cy = 0.5 * (ys[:-1] + ys[1:])
    hole = ((cy[:, None] >= hole_ymin) & (cy[:, None] <= hole_ymax)
            & (cx[None, :] >= hole_xmin) & (cx[None, :] <= hole_xmax)).reshape(-1)
    keep = np.repeat(~hole, 2)                      # two triangles per cell
    E = EToV[keep]
    used = np.unique(E)
    remap = np.full(VX.size, -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    return VX[used], VY[used], remap[E]


def build_mesh_2d(VX, VY, EToV, N: int, node_tol: float = 1e-10) -> Mesh2D:
    """StartUp2D (mesh.py:691-760), vectorised."""
    ref = ref_element(N)
    VX = np.asarray(VX, dtype=np.float64)
    VY = np.asarray(VY, dtype=np.float64)
    EToV = np.asarray(EToV, dtype=np.int64)
    K, Np, Nfp = EToV.shape[0], ref.Np, ref.Nfp
    r, s, Dr, Ds, Fm = ref.r, ref.s, ref.Dr, ref.Ds, ref.Fmask
    va, vb, vc = EToV[:, 0], EToV[:, 1], EToV[:, 2]
    x = 0.5 * (np.outer(-(r + s), VX[va]) + np.outer(1 + r, VX[vb]) + np.outer(1 + s, VX[vc]))
    y = 0.5 * (np.outer(-(r + s), VY[va]) + np.outer(1 + r, VY[vb]) + np.outer(1 + s, VY[vc]))
    rx, sx, ry, sy, J, nx, ny, sJ, Fscale = geometric_factors(x, y, N)
    EToE, EToF = _connect(EToV)
    # node maps: face f's nodes run v0->v1, v1->v2, v0->v2 for f = 0, 1, 2
    ids = np.arange(K * Np).reshape(K, Np)                 # ids[k, i] = k*Np+i
    vM = ids[:, Fm.T]                                      # (K, 3, Nfp)
    start_v = EToV[:, [0, 1, 0]]                           # (K, 3)
    kk = EToE
    ff = EToF
    same = start_v == start_v[kk, ff]
    nb = vM[kk, ff]                                        # (K, 3, Nfp)
    vP = np.where(same[..., None], nb, nb[..., ::-1])
    bdy = kk == np.arange(K)[:, None]
    vP = np.where(bdy[..., None], vM, vP)
    xf, yf = x.T.reshape(-1), y.T.reshape(-1)
    edge = np.hypot(VX[EToV] - VX[np.roll(EToV, -1, 1)], VY[EToV] - VY[np.roll(EToV, -1, 1)])
    dist = np.hypot(xf[vM] - xf[vP], yf[vM] - yf[vP])
    if np.any(dist >= node_tol * edge[..., None]):
        raise ValueError("face nodes of neighbouring elements do not coincide; "
                         "EToV must describe a conforming, consistently oriented mesh")
    vmapM = np.transpose(vM, (2, 1, 0)).reshape(-1, order="F")
    vmapP = np.transpose(vP, (2, 1, 0)).reshape(-1, order="F")
    mapB = np.where(vmapP == vmapM)[0]
    return Mesh2D(N=N, Np=Np, K=K, Nfp=Nfp, Nfaces=3, VX=VX, VY=VY, EToV=EToV,
                  r=r, s=s, V=ref.V, invV=ref.invV, Dr=Dr, Ds=Ds, Fmask=Fm,
                  x=x, y=y, rx=rx, sx=sx, ry=ry, sy=sy, J=J, nx=nx, ny=ny, sJ=sJ,
                  Fscale=Fscale, EToE=EToE, EToF=EToF, vmapM=vmapM, vmapP=vmapP,
                  vmapB=vmapM[mapB], mapB=mapB)


def boundary_faces(mesh) -> np.ndarray:
    """(K, 3) bool: faces whose EToE points to the element itself."""
    E = np.asarray(mesh.EToE)
    return E == np.arange(mesh.K)[:, None]


def face_normals(mesh):
    """(K, 3) outward normal components at each face (affine: constant)."""
    Nfp, K = mesh.Nfp, mesh.K
    nx = np.asarray(mesh.nx).reshape((Nfp, 3, K), order="F")[0].T
    ny = np.asarray(mesh.ny).reshape((Nfp, 3, K), order="F")[0].T
    return nx, ny


def face_centres(mesh):
    """(K, 3) mean coordinates of each face's nodes."""
    Fm = np.asarray(mesh.Fmask)
    x, y = np.asarray(mesh.x), np.asarray(mesh.y)
    cx = np.stack([x[Fm[:, f]].mean(0) for f in range(3)], 1)
    cy = np.stack([y[Fm[:, f]].mean(0) for f in range(3)], 1)
    return cx, cy


def bc_by_normal(mesh, dirichlet, neumann) -> np.ndarray:
    """Tag boundary faces from their outward normal, as the reference's
    ``_bc_types_by_normal`` test helper (test_sipg_poisson_2d.py:41-64).
    ``dirichlet``/``neumann`` map ``(nx, ny)`` (K, 3) arrays to bool masks."""
    bdy = boundary_faces(mesh)
    nx, ny = face_normals(mesh)
    bc = np.full((mesh.K, 3), BC_INTERIOR, dtype=np.int32)
    bc = np.where(dirichlet(nx, ny) & bdy, BC_DIRICHLET, bc)
    bc = np.where(neumann(nx, ny) & bdy, BC_NEUMANN, bc)
    return bc.astype(np.int32)


def all_dirichlet(mesh) -> np.ndarray:
    return np.where(boundary_faces(mesh), BC_DIRICHLET, BC_INTERIOR).astype(np.int32)


def dirichlet_x_neumann_y(mesh) -> np.ndarray:
    """Dirichlet on x = const walls, Neumann on y = const walls."""
    return bc_by_normal(mesh, lambda nx, ny: np.abs(nx) > 0.5, lambda nx, ny: np.abs(ny) > 0.5)


def refine_uniform(VX, VY, EToV):
    """One uniform red refinement with the child numbering of the reference's
    refinement forest (refinement.py:359-382, 389-455): fine element 4k+c is
    child c of element k, children (v0,m01,m20), (m01,v1,m12), (m20,m12,v2),
    (m01,m12,m20).  Midpoint vertices are numbered in order of first use
    (element order, edges 01, 12, 20).  Returns (VX, VY, EToV)."""
    VX = np.asarray(VX, dtype=np.float64)
    VY = np.asarray(VY, dtype=np.float64)
    EToV = np.asarray(EToV, dtype=np.int64)
    K, nv = EToV.shape[0], VX.size
    e = np.stack([EToV[:, [0, 1]], EToV[:, [1, 2]], EToV[:, [2, 0]]], axis=1).reshape(-1, 2)
    key = np.sort(e, axis=1)
    key = key[:, 0] * nv + key[:, 1]
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")           # unique edges by first use
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    mid = nv + rank[inv].reshape(K, 3)                 # (K,3): m01, m12, m20
    ea = e.reshape(K, 3, 2)
    first_edges = ea.reshape(-1, 2)[first[order]]
    VX2 = np.concatenate([VX, 0.5 * (VX[first_edges[:, 0]] + VX[first_edges[:, 1]])])
    VY2 = np.concatenate([VY, 0.5 * (VY[first_edges[:, 0]] + VY[first_edges[:, 1]])])
    v0, v1, v2 = EToV[:, 0], EToV[:, 1], EToV[:, 2]
    m01, m12, m20 = mid[:, 0], mid[:, 1], mid[:, 2]
    kids = np.stack([np.stack([v0, m01, m20], 1), np.stack([m01, v1, m12], 1),
                     np.stack([m20, m12, v2], 1), np.stack([m01, m12, m20], 1)], axis=1)
    return VX2, VY2, kids.reshape(4 * K, 3)


def forest_rectangle_hierarchy(n_base: int, levels: int, bounds=(0.0, 1.0, 0.0, 1.0)):
    """Vertex/element arrays of a structured ``2*n_base^2`` rectangle mesh
    refined uniformly ``levels`` times: a list ``[(VX, VY, EToV), ...]``
    from coarsest to finest, element 4k+c of level l+1 being child c of
    element k of level l.  Vertex ids of a coarse level are kept on every
    finer level, so vertex moves made on the finest level can be carried to
    the coarse ones by truncation."""
    xmin, xmax, ymin, ymax = bounds
    out = [uniform_rectangle_mesh_2d(xmin, xmax, ymin, ymax, n_base, n_base)]
    for _ in range(levels):
        out.append(refine_uniform(*out[-1]))
    return out
