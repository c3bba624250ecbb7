"""Device-resident SIPG operator: packs a ``Mesh2D`` + BC tags + penalty
into per-element records and owns the libddgx handle.

Packing (host, float64) follows the reference's per-face quantities
(src/ddg/dg2d/sipg.py:117-139, 211-280): for element k and face f,

* ``tau_k,f = tau_scale * (N+1)^2 * Fscale[face node 0]``     (sipg.py:263)
* interior face (both sides tagged 0, distinct neighbour k', f'):
  ``pen = (kappa_k tau_k sJ_k + kappa_k' tau_k' sJ_k') / 2``,
  ``sJw = (sJ_k + sJ_k') / 4``  -- the two half-weighted sides of the
  reference's face loop (w = 0.5) combined;
* Dirichlet: ``pen = kappa_k tau_k sJ_k``, ``sJw = sJ_k``     (w = 1);
* Neumann, or a boundary face tagged interior: no coupling (w = 0 in the
  reference; a self-neighbour face has zero jump and its averaged terms
  cancel in the gradient).

Element-constant geometry is read at node 0 of each element (affine
elements: the reference's per-node arrays are constant up to rounding).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .reference import mfma_fragments, ref_element

BC_INTERIOR, BC_DIRICHLET, BC_NEUMANN = 0, 1, 2


def _face_rows(a, Nfp, K):
    """(3*Nfp, K) face array -> (K, 3) values at face node 0."""
    return np.asarray(a, dtype=np.float64).reshape((Nfp, 3, K), order="F")[0].T


def pack(mesh, bc_types, tau_scale=200.0, kappa=None):
    """Return ``(rec (K,16), nbr (K,4) int32, fsj (K,4))`` float64/int32."""
    N, K, Nfp = int(mesh.N), int(mesh.K), int(mesh.Nfp)
    ref = ref_element(N)
    if not np.array_equal(np.asarray(mesh.Fmask), ref.Fmask):
        raise ValueError("mesh.Fmask does not match the degree-N reference element")
    bc = np.asarray(bc_types).astype(np.int64)
    if bc.shape != (K, 3):
        raise ValueError(f"bc_types must have shape (K, 3) = {(K, 3)}, got {bc.shape}")
    if np.any((bc < 0) | (bc > 2)):
        raise ValueError("bc_types entries must be BC_INTERIOR, BC_DIRICHLET or BC_NEUMANN")
    kap = np.ones(K) if kappa is None else np.asarray(kappa, dtype=np.float64).reshape(K)
    J = np.asarray(mesh.J, dtype=np.float64)[0]
    rx, sx = np.asarray(mesh.rx)[0], np.asarray(mesh.sx)[0]
    ry, sy = np.asarray(mesh.ry)[0], np.asarray(mesh.sy)[0]
    nx, ny = _face_rows(mesh.nx, Nfp, K), _face_rows(mesh.ny, Nfp, K)
    sJ, Fs = _face_rows(mesh.sJ, Nfp, K), _face_rows(mesh.Fscale, Nfp, K)
    tau = np.broadcast_to(np.asarray(tau_scale, dtype=np.float64), (3, K)).T
    tau_side = tau * (N + 1) ** 2 * Fs * kap[:, None]             # (K, 3)

    EToE = np.asarray(mesh.EToE, dtype=np.int64)
    EToF = np.asarray(mesh.EToF, dtype=np.int64)
    selfn = EToE == np.arange(K)[:, None]
    interior = (bc == BC_INTERIOR) & ~selfn
    kn, fn = EToE, EToF
    if np.any(interior & (bc[kn, fn] != BC_INTERIOR)):
        raise ValueError("interior faces must be tagged BC_INTERIOR on both sides")
    # neighbour node order from vmapP: same or reversed face-node order
    vM = np.asarray(mesh.vmapM).reshape((Nfp, 3, K), order="F")
    vP = np.asarray(mesh.vmapP).reshape((Nfp, 3, K), order="F")
    nbM = vM[:, fn, kn]                                           # (Nfp, K, 3)
    Pk = np.transpose(vP, (0, 2, 1))                              # (Nfp, K, 3)
    fwd = np.all(Pk == nbM, axis=0)
    rev = np.all(Pk == nbM[::-1], axis=0)
    if np.any(interior & ~(fwd | rev)):
        raise ValueError("vmapP does not match the neighbour face in either node order")
    r = (interior & rev & ~fwd).astype(np.int64)

    code = np.where(bc == BC_DIRICHLET, -1, -2).astype(np.int64)
    code = np.where(interior, (kn << 3) | (fn << 1) | r, code)
    sJn = sJ[kn, fn]
    tn = tau_side[kn, fn]
    isd = bc == BC_DIRICHLET
    sjw = np.where(interior, 0.25 * (sJ + sJn), np.where(isd, sJ, 0.0))
    pen = np.where(interior, 0.5 * (tau_side * sJ + tn * sJn), np.where(isd, tau_side * sJ, 0.0))

    rec = np.zeros((K, _lib.REC_WORDS))
    rec[:, 0] = kap * J * (rx * rx + ry * ry)
    rec[:, 1] = kap * J * (rx * sx + ry * sy)
    rec[:, 2] = kap * J * (sx * sx + sy * sy)
    rec[:, 3] = J
    rec[:, 4:7] = kap[:, None] * (nx * rx[:, None] + ny * ry[:, None])
    rec[:, 7:10] = kap[:, None] * (nx * sx[:, None] + ny * sy[:, None])
    rec[:, 10:13] = sjw
    rec[:, 13:16] = pen
    nbr = np.zeros((K, 4), dtype=np.int32)
    nbr[:, :3] = code
    fsj = np.zeros((K, 4))
    fsj[:, :3] = np.where(bc == BC_NEUMANN, sJ, 0.0)
    return rec, nbr, fsj


def is_affine(mesh, rtol=1e-9):
    """Whether every element's metrics and Jacobian are constant over its
    nodes (straight-sided elements), to ``rtol`` of their magnitude."""
    for name in ("rx", "sx", "ry", "sy", "J"):
        a = np.asarray(getattr(mesh, name), dtype=np.float64)
        if np.abs(a - a[0]).max() > rtol * max(np.abs(a).max(), 1e-300):
            return False
    return True


def is_conforming(mesh, rtol=1e-9):
    """Whether the node pairs vmapM/vmapP of every face sit at the same place
    (a curved-boundary projection moves nodes of one side only)."""
    x = np.asarray(mesh.x, dtype=np.float64).T.reshape(-1)
    y = np.asarray(mesh.y, dtype=np.float64).T.reshape(-1)
    vM, vP = np.asarray(mesh.vmapM), np.asarray(mesh.vmapP)
    scale = max(np.ptp(x), np.ptp(y), 1e-300)
    return bool(np.hypot(x[vM] - x[vP], y[vM] - y[vP]).max() <= rtol * scale)


def _side_codes(mesh, bc_types):
    """(K, 3) int32 face codes of every element side: ``(k'<<3)|(f'<<1)|rev``
    for sides coupled to a neighbour (a boundary face tagged interior couples
    to itself), -1 Dirichlet, -2 Neumann; and the bc array."""
    N, K, Nfp = int(mesh.N), int(mesh.K), int(mesh.Nfp)
    bc = np.asarray(bc_types).astype(np.int64)
    if bc.shape != (K, 3):
        raise ValueError(f"bc_types must have shape (K, 3) = {(K, 3)}, got {bc.shape}")
    if np.any((bc < 0) | (bc > 2)):
        raise ValueError("bc_types entries must be BC_INTERIOR, BC_DIRICHLET or BC_NEUMANN")
    kn = np.asarray(mesh.EToE, dtype=np.int64)
    fn = np.asarray(mesh.EToF, dtype=np.int64)
    inner = bc == BC_INTERIOR
    selfn = kn == np.arange(K)[:, None]
    if np.any(inner & ~selfn & (bc[kn, fn] != BC_INTERIOR)):
        raise ValueError("interior faces must be tagged BC_INTERIOR on both sides")
    vM = np.asarray(mesh.vmapM).reshape((Nfp, 3, K), order="F")
    vP = np.asarray(mesh.vmapP).reshape((Nfp, 3, K), order="F")
    nbM = vM[:, fn, kn]                                           # (Nfp, K, 3)
    Pk = np.transpose(vP, (0, 2, 1))
    fwd = np.all(Pk == nbM, axis=0)
    rev = np.all(Pk == nbM[::-1], axis=0) & ~fwd
    if np.any(inner & ~(fwd | rev)):
        raise ValueError("vmapP does not match the neighbour face in either node order")
    code = np.where(bc == BC_DIRICHLET, -1, -2)
    code = np.where(inner, (kn << 3) | (fn << 1) | rev.astype(np.int64), code)
    return code.astype(np.int32), bc


def pack_curved(mesh, bc_types, tau_scale=200.0, cub=None):
    """Host arrays of ``ddgx_op_create_curved``: per-node metrics, per-side
    normals, weights and penalty, face codes, and with ``cub`` the face
    cubature masses, volume cubature weights and point tables
    (sipg.py:142-301; cubature.py:107-221)."""
    from .cubature import face_mass, volume_weights
    N, K, Nfp = int(mesh.N), int(mesh.K), int(mesh.Nfp)
    ref = ref_element(N)
    if not np.array_equal(np.asarray(mesh.Fmask), ref.Fmask):
        raise ValueError("mesh.Fmask does not match the degree-N reference element")
    codes, bc = _side_codes(mesh, bc_types)
    reftab = np.concatenate([ref.Dr.ravel(), ref.Ds.ravel(), ref.M.ravel(), ref.m1d.ravel()])
    geo = np.stack([np.asarray(getattr(mesh, n), dtype=np.float64).T
                    for n in ("rx", "sx", "ry", "sy", "J")], axis=1)       # (K, 5, Np)
    face = lambda a: np.transpose(np.asarray(a, dtype=np.float64).reshape((Nfp, 3, K),
                                                                          order="F"), (2, 1, 0))
    fnorm = np.stack([face(mesh.nx), face(mesh.ny)], axis=-1)             # (K, 3, Nfp, 2)
    Fs0 = face(mesh.Fscale)[:, :, 0]
    sJ0 = face(mesh.sJ)[:, :, 0]
    tau = np.broadcast_to(np.asarray(tau_scale, dtype=np.float64), (3, K)).T
    w = np.where(bc == BC_DIRICHLET, 1.0, np.where(bc == BC_NEUMANN, 0.0, 0.5))
    fpar = np.stack([w, w * tau * (N + 1) ** 2 * Fs0, sJ0], axis=-1)      # (K, 3, 3)
    out = dict(reftab=reftab, geo=geo, fnorm=fnorm, fpar=fpar, codes=codes, fmass=None,
               estiff=None, emass=None)
    if cub is not None:
        Lq = np.asarray(cub.Lq, dtype=np.float64)
        if Lq.shape[1] != ref.Np or np.asarray(cub.J).shape[1] != K:
            raise ValueError("cub was built for a different mesh or order")
        out.update(fmass=face_mass(cub), **element_matrices(cub))
    return out


def element_matrices(cub):
    """Each element's volume cubature stiffness ``sum_q w J grad L_i . grad
    L_j`` and mass ``sum_q w J L_i L_j`` (sipg.py:190-201,
    cubature.py:228-239), both symmetric: (K, Np(Np+1)/2) upper triangles,
    row after row, from the per-point weights of
    :func:`ddg_amd.cubature.volume_weights`."""
    from .cubature import volume_weights
    vw = volume_weights(cub)                                   # (K, 4, Ncub)
    Dr = np.asarray(cub.Drq, dtype=np.float64)
    Ds = np.asarray(cub.Dsq, dtype=np.float64)
    Lq = np.asarray(cub.Lq, dtype=np.float64)
    Nq, Np = Lq.shape
    outer = lambda A, B: (A[:, :, None] * B[:, None, :]).reshape(Nq, Np * Np)
    rs = outer(Dr, Ds)
    stiff = (vw[:, 0] @ outer(Dr, Dr) + vw[:, 1] @ (rs + outer(Ds, Dr))
             + vw[:, 2] @ outer(Ds, Ds))
    mass = vw[:, 3] @ outer(Lq, Lq)
    K = vw.shape[0]
    iu = np.triu_indices(Np)
    up = lambda a: np.ascontiguousarray(a.reshape(K, Np, Np)[:, iu[0], iu[1]])
    return {"estiff": up(stiff), "emass": up(mass)}


TILE = 64          # elements per matrix-core workgroup
HALO_MAX = 64      # out-of-tile neighbour elements a tile may have


def max_tile_halo(EToE, perm=None, tile=TILE):
    """Largest number of distinct out-of-tile face neighbours of a tile when
    the elements are processed in the order ``perm``."""
    E = np.asarray(EToE, dtype=np.int64)
    K = E.shape[0]
    perm = np.arange(K) if perm is None else np.asarray(perm, dtype=np.int64)
    inv = np.empty(K, dtype=np.int64)
    inv[perm] = np.arange(K)
    En = E[perm]                                   # neighbours of each position
    pn = inv[En]
    t = np.arange(K) // tile
    out = (pn // tile != t[:, None]) & (En != perm[:, None])
    keys = np.unique(t[:, None].repeat(3, 1)[out] * K + pn[out])
    return int(np.bincount(keys // K, minlength=1).max()) if keys.size else 0


def morton_order(mesh):
    """Element order along a Z-order curve of the element centroids."""
    cx = np.asarray(mesh.x).mean(axis=0)
    cy = np.asarray(mesh.y).mean(axis=0)
    span = max(np.ptp(cx), np.ptp(cy), 1e-300)
    qx = ((cx - cx.min()) / span * 65535).astype(np.uint64)
    qy = ((cy - cy.min()) / span * 65535).astype(np.uint64)

    def spread(v):
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
        return v
    return np.argsort(spread(qx) | (spread(qy) << np.uint64(1)), kind="stable")


def tile_order(mesh):
    """Processing order of the elements: None (the mesh's own order) when it
    keeps every 64-element tile's halo at 32 elements or fewer (e.g.
    refinement-forest meshes), else a Morton order of the centroids when that
    is more compact.  Each 16 halo elements cost one extra gradient product."""
    natural = max_tile_halo(mesh.EToE)
    if natural <= HALO_MAX // 2:
        return None
    perm = morton_order(mesh)
    m = max_tile_halo(mesh.EToE, perm)
    if m < natural:
        return perm.astype(np.int32)
    if natural > HALO_MAX:
        raise ValueError("no element order keeps the 64-element tiles compact enough")
    return None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class SipgOperator:
    """One SIPG operator on the current HIP device, in float64 or float32.

    ``apply`` / ``rhs`` / ``pcg`` take and return torch tensors on the
    device in the element-major flat layout (``(K*Np,)`` or ``(K, Np)``).
    """

    def __init__(self, mesh, bc_types, tau_scale=200.0, kappa=None, dtype=torch.float64,
                 order="auto", cub=None):
        """``order``: element processing order of the apply's 64-element tiles:
        ``"auto"`` (``tile_order``), ``None`` (the mesh's own numbering) or a
        permutation of ``range(K)``.  A mesh with curved elements, or a
        ``cub`` (``MeshCubature``), gets the general operator
        (``ddgx_op_create_curved``) instead of the affine one."""
        if not torch.cuda.is_available():
            raise RuntimeError("SipgOperator needs a HIP device (no CPU fallback)")
        self.N, self.K, self.Np, self.Nfp = int(mesh.N), int(mesh.K), int(mesh.Np), int(mesh.Nfp)
        self.dtype = dtype
        self.code = _lib.DDGX_F64 if dtype == torch.float64 else _lib.DDGX_F32
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be torch.float64 or torch.float32")
        self._bj = {}
        self.cub = cub
        # the affine kernel merges both sides of a face and reads element
        # constants: it needs straight-sided elements whose face nodes meet
        self.curved = cub is not None or not (is_affine(mesh) and is_conforming(mesh))
        if self.curved:
            if kappa is not None:
                raise NotImplementedError("kappa is not supported on curved operators")
            self._create_curved(mesh, bc_types, tau_scale, cub)
            return
        rec, nbr, fsj = pack(mesh, bc_types, tau_scale, kappa)
        ref = ref_element(self.N)
        tab = np.concatenate([ref.kernel_tables(),
                              mfma_fragments(ref, "f64" if dtype == torch.float64 else "f32")])
        L = _lib.lib()
        if tab.size != L.ddgx_table_size(self.N):
            raise RuntimeError(
                f"libddgx.so expects {L.ddgx_table_size(self.N)} table words at order {self.N}, "
                f"the host built {tab.size}: the library and ddg_amd/reference.py disagree "
                "(rebuild with python -m ddg_amd.build)")
        if isinstance(order, str):
            if order != "auto":
                raise ValueError("order must be 'auto', None or a permutation of range(K)")
            order = tile_order(mesh)
        elif order is not None:
            order = np.ascontiguousarray(np.asarray(order, dtype=np.int32).reshape(-1))
            if order.size != self.K:
                raise ValueError("order must hold K entries")
        self.order = order
        h = C.c_void_p()
        torch.cuda.current_device()  # make sure the HIP context exists
        _lib.check(L.ddgx_op_create(self.N, self.K, self.code,
                                    tab.ctypes.data_as(C.c_void_p),
                                    np.ascontiguousarray(rec).ctypes.data_as(C.c_void_p),
                                    np.ascontiguousarray(nbr).ctypes.data_as(C.c_void_p),
                                    np.ascontiguousarray(fsj).ctypes.data_as(C.c_void_p),
                                    None if order is None else order.ctypes.data_as(C.c_void_p),
                                    C.byref(h)), "ddgx_op_create")
        self._h = h
        self.device = torch.device("cuda", torch.cuda.current_device())

    def _create_curved(self, mesh, bc_types, tau_scale, cub):
        d = pack_curved(mesh, bc_types, tau_scale, cub)
        self.order = None
        keep = {k: (None if v is None else np.ascontiguousarray(v)) for k, v in d.items()}
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        torch.cuda.current_device()
        _lib.check(_lib.lib().ddgx_op_create_curved(
            self.N, self.K, self.code, int(cub is not None), ptr(keep["reftab"]),
            ptr(keep["geo"]), ptr(keep["fnorm"]), ptr(keep["fpar"]), ptr(keep["codes"]),
            ptr(keep["fmass"]), ptr(keep["estiff"]), ptr(keep["emass"]), C.byref(h)),
            "ddgx_op_create_curved")
        self._h = h
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                for pc in self._bj.values():
                    pc.close()
                _lib.lib().ddgx_op_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def _vec(self, t, name):
        if t is None:
            return None
        if not (isinstance(t, torch.Tensor) and t.is_cuda):
            raise TypeError(f"{name} must be a torch tensor on the HIP device")
        if t.dtype != self.dtype or t.numel() != self.K * self.Np or not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous {self.dtype} with K*Np = "
                             f"{self.K * self.Np} entries (got {t.dtype}, {t.numel()})")
        return t

    def new(self):
        return torch.empty(self.K * self.Np, dtype=self.dtype, device=self.device)

    def apply(self, u, y=None, *, u_dir=None, alpha=0.0, shift=0.0):
        u = self._vec(u, "u")
        y = self.new() if y is None else self._vec(y, "y")
        u_dir = self._vec(u_dir, "u_dir")
        _lib.check(_lib.lib().ddgx_sipg_apply(self._h, _ptr(u), _ptr(y), _ptr(u_dir),
                                              float(alpha), float(shift), _stream()),
                   "ddgx_sipg_apply")
        return y

    def rhs(self, f=None, *, u_dir=None, q_neu=None, out=None):
        f = self._vec(f, "f")
        u_dir = self._vec(u_dir, "u_dir")
        if q_neu is not None:
            if q_neu.dtype != self.dtype or q_neu.numel() != self.K * 3 * self.Nfp:
                raise ValueError("q_neu must hold K*3*Nfp values of the operator dtype")
        out = self.new() if out is None else self._vec(out, "out")
        _lib.check(_lib.lib().ddgx_sipg_rhs(self._h, _ptr(f), _ptr(u_dir), _ptr(q_neu),
                                            _ptr(out), _stream()), "ddgx_sipg_rhs")
        return out

    def dot(self, a, b):
        out = C.c_double()
        _lib.check(_lib.lib().ddgx_dot(self._h, _ptr(self._vec(a, "a")),
                                       _ptr(self._vec(b, "b")), C.byref(out), _stream()),
                   "ddgx_dot")
        return out.value

    def block_jacobi(self, alpha=0.0, dtype=None):
        """Block-Jacobi preconditioner of A + alpha*M (needs a float64 operator)."""
        dtype = self.dtype if dtype is None else dtype
        key = (float(alpha), dtype)
        if key not in self._bj:
            self._bj[key] = BlockJacobi(self, alpha, dtype)
        return self._bj[key]

    def pcg(self, b, x=None, *, alpha=0.0, shift=0.0, pc=None, tol=1e-9, atol=0.0,
            maxiter=1000, check_every=16):
        """CG with JAX's stopping rule; returns ``(x, info_dict)``."""
        b = self._vec(b, "b")
        x = torch.zeros_like(b) if x is None else self._vec(x, "x")
        cfg = _lib.CGConfig(float(tol), float(atol), int(maxiter), int(check_every))
        info = _lib.CGInfo()
        pch = None
        if pc is not None:
            if not isinstance(pc, DevicePreconditioner):
                raise TypeError("pc must be a device preconditioner (block-Jacobi or multigrid)")
            if pc.K != self.K or pc.N != self.N:
                raise ValueError("preconditioner was built for a different mesh/order")
            if pc.dtype != self.dtype and not (
                    self.dtype == torch.float64 and pc.dtype == torch.float32
                    and getattr(pc, "is_multigrid", False)):
                raise ValueError("preconditioner dtype differs from the operator dtype (only a "
                                 "float32 multigrid may precondition a float64 solve)")
            pch = pc.handle
        _lib.check(_lib.lib().ddgx_pcg_solve(self._h, float(alpha), float(shift), pch,
                                             _ptr(b), _ptr(x), C.byref(cfg), C.byref(info),
                                             _stream()), "ddgx_pcg_solve")
        return x, {"iterations": info.iterations, "converged": bool(info.converged),
                   "residual2": info.residual2, "atol2": info.atol2}


class BatchSolver:
    """``ddgx_batch``: ``nseg`` problems of one class -- the float64 operator
    ``op`` and the multigrid ``pc`` (float32 or float64, or None) -- solved
    together, one kernel launch per step for all of them.  Each problem has
    its own right-hand side and Helmholtz coefficient; its solution is bitwise
    the one ``op.pcg(b, alpha=a, pc=pc)`` returns for it alone.  The batch
    keeps device copies of ``op`` and ``pc`` (``nbytes`` in all)."""

    def __init__(self, op: SipgOperator, pc=None, nseg: int = 1):
        if op.dtype != torch.float64:
            raise ValueError("batched solves take a float64 operator")
        if pc is not None and not getattr(pc, "is_multigrid", False):
            raise ValueError("batched solves take a multigrid preconditioner or none")
        self.nseg, self.K, self.Np = int(nseg), op.K, op.Np
        self.n = op.K * op.Np
        self.device = op.device
        h = C.c_void_p()
        _lib.check(_lib.lib().ddgx_batch_create(op.handle, None if pc is None else pc.handle,
                                                self.nseg, _stream(), C.byref(h)),
                   "ddgx_batch_create")
        self._h = h
        nb = C.c_int64()
        _lib.check(_lib.lib().ddgx_batch_bytes(h, C.byref(nb)), "ddgx_batch_bytes")
        self.nbytes = int(nb.value)

    def pcg(self, b, alpha, x=None, *, tol=1e-9, atol=0.0, maxiter=1000, check_every=16,
            stream=None):
        """Solve ``(A + alpha[s] M) x_s = b_s`` for every segment s.  ``b``, ``x``:
        contiguous float64 device tensors of ``nseg * K * Np`` (segment s at
        offset ``s * K * Np``); returns ``(x, [info_dict per segment])``."""
        if not (isinstance(b, torch.Tensor) and b.is_cuda and b.dtype == torch.float64
                and b.is_contiguous() and b.numel() == self.nseg * self.n):
            raise ValueError(f"b must be contiguous float64 on the device with "
                             f"{self.nseg} x {self.n} entries")
        x = torch.zeros_like(b) if x is None else x
        if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()
                and x.numel() == b.numel()):
            raise ValueError("x must match b")
        a = [float(v) for v in alpha]
        if len(a) != self.nseg:
            raise ValueError(f"need {self.nseg} Helmholtz coefficients")
        av = (C.c_double * self.nseg)(*a)
        cfg = _lib.CGConfig(float(tol), float(atol), int(maxiter), int(check_every))
        infos = (_lib.CGInfo * self.nseg)()
        st = _stream() if stream is None else C.c_void_p(stream)
        _lib.check(_lib.lib().ddgx_batch_pcg(self._h, av, _ptr(b), _ptr(x), C.byref(cfg),
                                             infos, st), "ddgx_batch_pcg")
        return x, [{"iterations": i.iterations, "converged": bool(i.converged),
                    "residual2": i.residual2, "atol2": i.atol2} for i in infos]

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().ddgx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def refine_solve(op64: SipgOperator, op32: SipgOperator, b, x=None, *, alpha=0.0, shift=0.0,
                 pc32=None, tol=1e-10, atol=0.0, max_outer=20, inner_tol=1e-4,
                 inner_maxiter=1000, check_every=16):
    """Mixed-precision iterative refinement: float64 residual and update
    around float32 PCG solves (``ddgx_ir_solve``).  Returns ``(x, info)``."""
    if op64.dtype != torch.float64 or op32.dtype != torch.float32:
        raise ValueError("refine_solve needs a float64 and a float32 operator")
    if op64.K != op32.K or op64.N != op32.N:
        raise ValueError("operators differ in mesh or order")
    b = op64._vec(b, "b")
    x = torch.zeros_like(b) if x is None else op64._vec(x, "x")
    pch = None
    if pc32 is not None:
        if not isinstance(pc32, DevicePreconditioner) or pc32.dtype != torch.float32:
            raise ValueError("pc32 must be a float32 device preconditioner")
        if pc32.K != op32.K or pc32.N != op32.N:
            raise ValueError("preconditioner was built for a different mesh/order")
        pch = pc32.handle
    cfg = _lib.IRConfig(float(tol), float(atol), int(max_outer), float(inner_tol),
                        int(inner_maxiter), int(check_every))
    info = _lib.IRInfo()
    _lib.check(_lib.lib().ddgx_ir_solve(op64.handle, op32.handle, float(alpha), float(shift),
                                        pch, _ptr(b), _ptr(x), C.byref(cfg), C.byref(info),
                                        _stream()), "ddgx_ir_solve")
    return x, {"outer": info.outer, "iterations": info.inner_total,
               "converged": bool(info.converged), "at_floor": bool(info.at_floor),
               "residual2": info.residual2, "atol2": info.atol2}


class DevicePreconditioner:
    """A preconditioner handle of libddgx: callable ``z = M^{-1} r`` on device
    vectors (the reference's preconditioner protocol, sipg.py:726-729,
    multigrid.py:246) and accepted by ``SipgOperator.pcg``."""

    _h = None
    K = N = 0
    dtype = torch.float64
    is_multigrid = False

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _lib.lib().ddgx_pc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, r, z=None):
        if not (isinstance(r, torch.Tensor) and r.is_cuda):
            raise TypeError("r must be a torch tensor on the HIP device")
        if r.dtype != self.dtype or r.numel() != self.K * ((self.N + 1) * (self.N + 2) // 2) \
                or not r.is_contiguous():
            raise ValueError("vector dtype/size differs from the preconditioner's")
        z = torch.empty_like(r) if z is None else z
        _lib.check(_lib.lib().ddgx_pc_apply(self._h, _ptr(r), _ptr(z), _stream()),
                   "ddgx_pc_apply")
        return z


class BlockJacobi(DevicePreconditioner):
    """Per-element inverse diagonal blocks on the device (sipg.py:707-805)."""

    def __init__(self, op64: SipgOperator, alpha=0.0, dtype=None):
        if op64.dtype != torch.float64:
            raise ValueError("block-Jacobi blocks are built from a float64 operator")
        self.op = op64
        self.K, self.N = op64.K, op64.N
        self.alpha = float(alpha)
        self.dtype = op64.dtype if dtype is None else dtype
        code = _lib.DDGX_F64 if self.dtype == torch.float64 else _lib.DDGX_F32
        h = C.c_void_p()
        _lib.check(_lib.lib().ddgx_block_jacobi_create(op64.handle, self.alpha, code, _stream(),
                                                       C.byref(h)), "ddgx_block_jacobi_create")
        self._h = h

    def blocks(self):
        K, Np = self.op.K, self.op.Np
        out = torch.empty((K, Np, Np), dtype=torch.float64, device=self.op.device)
        _lib.check(_lib.lib().ddgx_block_jacobi_blocks(self._h, _ptr(out), _stream()),
                   "ddgx_block_jacobi_blocks")
        return out
