"""Drop-in for ``ddg.dg2d.sipg``: SIPG Poisson/Helmholtz apply, loads and
matrix-free CG solves, executed by the gfx950 HIP kernels of libddgx.

Same names, keyword arguments, array shapes and error behaviour as
``src/ddg/dg2d/sipg.py`` (BC codes 54-56, ``IterativeConfig`` 59-71,
``sipg_apply_2d`` 304-337, loads 398-453, solves 536-612 and 812-918,
block-Jacobi builders 718-805), so a caller can switch by import path.

Inputs may be numpy arrays (copied to the device, results copied back as
numpy) or torch tensors on the HIP device (results stay on the device).
Operator handles are cached per (mesh, bc_types, tau_scale, dtype), so
repeated calls do not repeat the setup that the reference recomputes on
every call (sipg.py:412-413, 849-850).

Differences from the reference, by design:
* Curved meshes (per-node metrics) and ``cub=`` (a ``MeshCubature`` from
  ``ddg_amd.cubature`` or the reference's) run on the general operator
  (``ddgx_op_create_curved``).  The preconditioner builders take such a
  mesh as the reference's do: block-Jacobi blocks of the nodal form by a
  colour probe of the general apply, p-multigrid with the curved finest
  level and straight-sided lower orders (multigrid.py:253-294).
* ``iterative=None`` (direct LU) assembles the dense matrix from device
  applies and factors it with torch on the device; like the reference it
  is only practical for small meshes.
* Solves return the solution; ``return_info=True`` also returns the CG
  iteration count and final residual (JAX's ``cg`` returns ``info=None``).
"""

from __future__ import annotations

import weakref
from dataclasses import dataclass

import numpy as np
import torch

from .operator import BlockJacobi, DevicePreconditioner, SipgOperator, refine_solve
from .reference import ref_element

BC_INTERIOR = 0
BC_DIRICHLET = 1
BC_NEUMANN = 2


@dataclass(frozen=True)
class RefineConfig:
    """Mixed-precision iterative refinement (an extension: the reference
    solves in float64 only).  The outer loop keeps the float64 residual and
    stops by ``IterativeConfig.tol``/``atol``; each correction is a float32
    PCG solve to ``inner_tol`` with a float32 preconditioner.

    The faster mixed-precision form needs no ``refine``: pass a float32
    multigrid (``multigrid.make_sipg_*_preconditioner(..., dtype=
    torch.float32)``) as ``preconditioner`` of a float64 solve, which then
    runs float64 CG with the V-cycle in float32 (the float64 solve's
    iteration count; about 1.4x its speed at 2x256x256, N=4)."""
    inner_tol: float = 1e-4
    max_outer: int = 20
    inner_maxiter: int = 1000


@dataclass(frozen=True)
class IterativeConfig:
    """Iterative solver configuration (sipg.py:59-71)."""
    method: str = "cg"
    tol: float = 1e-9
    atol: float = 0.0
    maxiter: int = 1000


# ---------------------------------------------------------------------------
# operator cache
# ---------------------------------------------------------------------------
_CACHE: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def _key_of(bc_types, tau_scale, kappa, dtype, cub=None):
    bc = np.ascontiguousarray(np.asarray(bc_types, dtype=np.int32))
    tau = np.ascontiguousarray(np.asarray(tau_scale, dtype=np.float64))
    kap = None if kappa is None else np.ascontiguousarray(np.asarray(kappa, dtype=np.float64))
    return (bc.tobytes(), bc.shape, tau.tobytes(), tau.shape,
            None if kap is None else kap.tobytes(), str(dtype), None if cub is None else id(cub))


def get_operator(mesh, bc_types, tau_scale=200.0, kappa=None, dtype=torch.float64,
                 cub=None) -> SipgOperator:
    """Cached device operator for this mesh / BC tagging / penalty / cubature
    (the operator keeps ``cub`` alive, so its identity is a stable key)."""
    try:
        per_mesh = _CACHE.setdefault(mesh, {})
    except TypeError:  # mesh not weak-referenceable: no caching
        return SipgOperator(mesh, bc_types, tau_scale, kappa, dtype, cub=cub)
    key = _key_of(bc_types, tau_scale, kappa, dtype, cub)
    op = per_mesh.get(key)
    if op is None:
        op = SipgOperator(mesh, bc_types, tau_scale, kappa, dtype, cub=cub)
        per_mesh[key] = op
    return op


# ---------------------------------------------------------------------------
# reference-element helpers the reference's callers import (sipg.py:78-139)
# ---------------------------------------------------------------------------
def _local_mass(mesh) -> np.ndarray:
    """Reference-element mass ``M_ref = inv(V V^T)`` (sipg.py:78-80)."""
    V = np.asarray(mesh.V, dtype=np.float64)
    return np.linalg.inv(V @ V.T)


def _face_param_for_face(mesh, f: int) -> np.ndarray:
    """1D parameter of face ``f``'s nodes in Fmask order: ``r`` on faces 0
    and 1, ``s`` on face 2 (sipg.py:83-95)."""
    Fm = np.asarray(mesh.Fmask)[:, f]
    return np.asarray(mesh.r if f < 2 else mesh.s, dtype=np.float64)[Fm]


def _face_1d_mass_matrices(mesh) -> np.ndarray:
    """(3, Nfp, Nfp) per-face 1D mass ``inv(V1D V1D^T)`` (sipg.py:98-110)."""
    from .reference import vandermonde_1d
    out = []
    for f in range(3):
        V1 = vandermonde_1d(int(mesh.N), _face_param_for_face(mesh, f))
        out.append(np.linalg.inv(V1 @ V1.T))
    return np.stack(out)


def _bc_face_arrays(mesh, bc_types):
    """``(bc_per_facenode_flat, bc_per_face, w_per_face)`` (sipg.py:117-139):
    face-node tags in F order, (3, K) tags, and face weights 0.5 interior,
    1 Dirichlet, 0 Neumann."""
    bpf = np.asarray(bc_types).T.astype(np.int32)
    Nfp, K = int(mesh.Nfp), int(mesh.K)
    flat = np.broadcast_to(bpf[None], (Nfp, 3, K)).reshape(-1, order="F")
    w = np.where(bpf == BC_DIRICHLET, 1.0, np.where(bpf == BC_NEUMANN, 0.0, 0.5))
    return flat, bpf, w


def _check_ref_mats(mesh, M_ref, m1d_per_face):
    ref = ref_element(int(mesh.N))
    if M_ref is not None and not np.allclose(np.asarray(M_ref), ref.M, rtol=1e-12, atol=1e-12):
        raise ValueError("M_ref differs from the reference-element mass matrix")
    if m1d_per_face is not None and not np.allclose(np.asarray(m1d_per_face), ref.m1d,
                                                    rtol=1e-12, atol=1e-12):
        raise ValueError("m1d_per_face differs from the reference-element face mass")


class _Field:
    """(Np, K) field <-> device flat element-major vector."""

    def __init__(self, x, dtype, device, shape):
        self.was_torch = isinstance(x, torch.Tensor)
        self.shape = shape
        if self.was_torch:
            t = x.to(device=device, dtype=dtype)
        else:
            t = torch.as_tensor(np.asarray(x, dtype=np.float64), device=device).to(dtype)
        self.flat = t.reshape(shape).T.contiguous().reshape(-1) if t.dim() == 2 else t.contiguous()

    def back(self, flat, like_torch=None):
        Np, K = self.shape
        out = flat.reshape(K, Np).T
        keep = self.was_torch if like_torch is None else like_torch
        return out.contiguous() if keep else out.double().cpu().numpy()


def _field(x, op):
    return _Field(x, op.dtype, op.device, (op.Np, op.K))


def _face_field(q, op):
    """(3*Nfp, K) face array -> device flat (K, 3, Nfp) = F-order flat."""
    t = q if isinstance(q, torch.Tensor) else torch.as_tensor(np.asarray(q, dtype=np.float64))
    t = t.to(device=op.device, dtype=op.dtype)
    return t.reshape(3 * op.Nfp, op.K).T.contiguous().reshape(-1)


# ---------------------------------------------------------------------------
# apply and loads
# ---------------------------------------------------------------------------
def sipg_apply_2d(mesh, u, *, bc_types, u_dir=None, tau_scale=200.0, M_ref=None,
                  m1d_per_face=None, cub=None, kappa=None):
    """``M (-Delta_h) u`` plus the Dirichlet-load offset when ``u_dir`` is
    given (sipg.py:304-337).  ``kappa`` (K,) is an extension for
    ``-div(kappa grad u)``; ``None`` is the reference operator."""
    _check_ref_mats(mesh, M_ref, m1d_per_face)
    op = get_operator(mesh, bc_types, tau_scale, kappa, cub=cub)
    fu = _field(u, op)
    ud = None if u_dir is None else _field(u_dir, op).flat
    return fu.back(op.apply(fu.flat, u_dir=ud))


def sipg_bilinear_form(mesh, u, v, *, bc_types, u_dir=None, tau_scale=200.0, M_ref=None,
                       m1d_per_face=None, cub=None, kappa=None):
    """``a(u, v)`` (sipg.py:142-301).  The form is linear in ``v``, whose
    ghost values are homogeneous, so ``a(u, v) = sum(v * apply(u; u_dir))``:
    one device apply and one float64 dot product.  Returns a Python float
    (numpy inputs) or a 0-d device tensor (torch inputs)."""
    _check_ref_mats(mesh, M_ref, m1d_per_face)
    op = get_operator(mesh, bc_types, tau_scale, kappa, cub=cub)
    fu = _field(u, op)
    fv = _field(v, op)
    ud = None if u_dir is None else _field(u_dir, op).flat
    y = op.apply(fu.flat, u_dir=ud)
    val = torch.dot(fv.flat.double(), y.double())
    return val if fu.was_torch else float(val)


def sipg_dirichlet_load_2d(mesh, u_dir, *, bc_types, tau_scale=200.0):
    """A_aff(0; u_dir) (sipg.py:398-417)."""
    op = get_operator(mesh, bc_types, tau_scale)
    fd = _field(u_dir, op)
    return fd.back(op.apply(torch.zeros_like(fd.flat), u_dir=fd.flat))


def sipg_neumann_load_2d(mesh, q_neu, *, bc_types):
    """Neumann surface load (sipg.py:420-453); ``q_neu`` is (3*Nfp, K)."""
    op = get_operator(mesh, bc_types)
    out = op.rhs(None, q_neu=_face_field(q_neu, op))
    Np, K = op.Np, op.K
    res = out.reshape(K, Np).T
    return res.contiguous() if isinstance(q_neu, torch.Tensor) else res.double().cpu().numpy()


def _rhs(op, source, u_dir, q_neu):
    f = _field(source, op)
    ud = None if u_dir is None else _field(u_dir, op).flat
    q = None if q_neu is None else _face_field(q_neu, op)
    return f, op.rhs(f.flat, u_dir=ud, q_neu=q)


# ---------------------------------------------------------------------------
# dense path (small meshes), sipg.py:344-391
# ---------------------------------------------------------------------------
def _dense(op, alpha=0.0, shift=0.0):
    n = op.K * op.Np
    A = torch.empty((n, n), dtype=torch.float64, device=op.device)
    e = torch.zeros(n, dtype=op.dtype, device=op.device)
    col = op.new()
    for j in range(n):
        e.zero_()
        e[j] = 1.0
        op.apply(e, col, alpha=alpha, shift=shift)
        A[:, j] = col
    return A


def sipg_matrix_2d(mesh, *, bc_types, tau_scale=200.0, cub=None):
    """Dense (Np K)^2 SIPG matrix from device applies (sipg.py:344-364)."""
    return _dense(get_operator(mesh, bc_types, tau_scale, cub=cub)).cpu().numpy()


def factor_sipg_2d(mesh, *, bc_types, tau_scale=200.0, regularize=False, eps=1.0e-8, cub=None):
    """(A, lu, piv) with the LU factors on the device (sipg.py:367-391)."""
    A = _dense(get_operator(mesh, bc_types, tau_scale, cub=cub))
    if regularize:
        A = A + eps * A.diagonal().abs().max() * torch.eye(A.shape[0], dtype=A.dtype, device=A.device)
    lu, piv = torch.linalg.lu_factor(A)
    return A, lu, piv


def sipg_helmholtz_matrix_2d(mesh, *, alpha, nu, bc_types, tau_scale=200.0, cub=None):
    """Dense ``alpha M + A`` from device applies (sipg.py:460-496); ``nu``
    is unused, as in the reference."""
    return _dense(get_operator(mesh, bc_types, tau_scale, cub=cub), alpha=alpha).cpu().numpy()


def _block_diag_mass(mesh, M_ref) -> np.ndarray:
    """Dense block-diagonal physical mass, block k = ``J_k M_ref`` at rows
    ``k Np:(k+1) Np`` (sipg.py:499-515)."""
    Np, K = int(mesh.Np), int(mesh.K)
    J = np.asarray(mesh.J, dtype=np.float64)[0]
    out = np.zeros((Np * K, Np * K))
    idx = np.arange(K)[:, None, None] * Np
    rows = idx + np.arange(Np)[None, :, None]
    cols = idx + np.arange(Np)[None, None, :]
    out[rows, cols] = J[:, None, None] * np.asarray(M_ref, dtype=np.float64)[None]
    return out


def factor_sipg_helmholtz_2d(mesh, *, alpha, nu, bc_types, tau_scale=200.0, cub=None):
    """alpha M + A factored (sipg.py:518-533); ``nu`` is unused as there."""
    A = _dense(get_operator(mesh, bc_types, tau_scale, cub=cub), alpha=alpha)
    lu, piv = torch.linalg.lu_factor(A)
    return A, lu, piv


def _lu_solve(factored, b):
    _, lu, piv = factored
    return torch.linalg.lu_solve(lu, piv, b.double().reshape(-1, 1)).reshape(-1)


# ---------------------------------------------------------------------------
# preconditioners (sipg.py:636-805)
# ---------------------------------------------------------------------------
def _extract_block_diag_dense(A_dense, Np, K):
    """(K, Np, Np) diagonal blocks of a dense element-major matrix
    (sipg.py:636-648).  numpy or torch, as given."""
    idx = np.arange(K)[:, None] * Np + np.arange(Np)[None, :]
    if isinstance(A_dense, torch.Tensor):
        t = torch.as_tensor(idx, device=A_dense.device)
        return A_dense[t[:, :, None], t[:, None, :]]
    return np.asarray(A_dense)[idx[:, :, None], idx[:, None, :]]


def _color_elements(mesh):
    """Greedy colouring of the face-adjacency graph in element order
    (sipg.py:651-668): ``(colour (K,), ncolours)``; at most four colours
    on a triangulation."""
    E = np.asarray(mesh.EToE, dtype=np.int64)
    K = E.shape[0]
    colour = np.full(K, -1, dtype=np.int64)
    for k in range(K):
        taken = 0
        for nb in E[k]:
            if nb != k and colour[nb] >= 0:
                taken |= 1 << int(colour[nb])
        c = 0
        while taken >> c & 1:
            c += 1
        colour[k] = c
    return colour, int(colour.max()) + 1 if K else 0


def extract_block_diag_matfree(apply_flat, mesh):
    """Per-element (K, Np, Np) diagonal blocks of any element-local-plus-
    face-neighbour operator from ``ncolours * Np`` probe applies
    (sipg.py:671-704): probing node j of every element of one colour is
    zero on all their face neighbours, so the result on those elements is
    column j of their own blocks.

    ``apply_flat`` takes and returns flat element-major vectors.  The probes
    are float64 tensors on the HIP device; an ``apply_flat`` that answers
    with numpy arrays gets numpy blocks back, otherwise device tensors.
    (The device preconditioner builders use a closed form instead of
    probing on affine meshes, and the same colour probe on curved ones:
    ``make_sipg_block_jacobi_preconditioner``.)"""
    Np, K = int(mesh.Np), int(mesh.K)
    colour, nc = _color_elements(mesh)
    dev = torch.device("cuda", torch.cuda.current_device())
    col = torch.as_tensor(colour, device=dev)
    blocks = torch.zeros((K, Np, Np), dtype=torch.float64, device=dev)
    as_numpy = False
    for c in range(nc):
        mask = col == c
        for j in range(Np):
            probe = torch.zeros((K, Np), dtype=torch.float64, device=dev)
            probe[:, j] = mask.double()
            r = apply_flat(probe.reshape(-1))
            if not isinstance(r, torch.Tensor):
                as_numpy = True
            r = torch.as_tensor(np.asarray(r) if as_numpy else r, device=dev).double()
            blocks[:, :, j] += torch.where(mask[:, None], r.reshape(K, Np), 0.0)
    return blocks.cpu().numpy() if as_numpy else blocks


def _block_jacobi_apply(blocks, Np, K):
    """``M_inv(v_flat)`` from per-element blocks (sipg.py:707-715): inverse
    blocks once, then one batched product per call (numpy or torch)."""
    if isinstance(blocks, torch.Tensor):
        binv = torch.linalg.inv(blocks)
        return lambda v: torch.einsum("kij,kj->ki", binv, v.reshape(K, Np).to(binv.dtype)) \
            .reshape(-1)
    binv = np.linalg.inv(np.asarray(blocks))
    return lambda v: np.einsum("kij,kj->ki", binv, np.asarray(v).reshape(K, Np)).reshape(-1)


def make_sipg_block_jacobi_preconditioner(mesh, *, bc_types, tau_scale=200.0, A_dense=None,
                                          matfree=False) -> BlockJacobi:
    """Block-Jacobi ``M^{-1}`` for the SIPG matrix.  Blocks are always built
    on the device in closed form (the reference's matfree result); the
    returned object is callable on device vectors and is accepted by the
    solves as ``preconditioner``."""
    return get_operator(mesh, bc_types, tau_scale).block_jacobi(0.0)


def make_helmholtz_block_jacobi_preconditioner(mesh, *, alpha, nu, bc_types, tau_scale=200.0,
                                               A_dense=None, matfree=False) -> BlockJacobi:
    return get_operator(mesh, bc_types, tau_scale).block_jacobi(float(alpha))


# ---------------------------------------------------------------------------
# solves (sipg.py:536-612, 812-918)
# ---------------------------------------------------------------------------
def _run_cg(op, b, x0, iterative, preconditioner, alpha=0.0, shift=0.0, refine=None, key=None):
    if iterative.method != "cg":
        raise ValueError(f"iterative method {iterative.method!r} not supported for SPD")
    if refine is not None:
        mesh, bc_types, tau_scale = key
        op32 = get_operator(mesh, bc_types, tau_scale, dtype=torch.float32, cub=op.cub)
        if preconditioner is not None and (not isinstance(preconditioner, DevicePreconditioner)
                                           or preconditioner.dtype != torch.float32):
            raise TypeError("refine needs a float32 device preconditioner (dtype=torch.float32)")
        return refine_solve(op, op32, b, x0, alpha=alpha, shift=shift, pc32=preconditioner,
                            tol=iterative.tol, atol=iterative.atol, max_outer=refine.max_outer,
                            inner_tol=refine.inner_tol, inner_maxiter=refine.inner_maxiter)
    pc = None
    if preconditioner is not None:
        if not isinstance(preconditioner, DevicePreconditioner):
            raise TypeError("preconditioner must come from make_*_block_jacobi_preconditioner "
                            "or ddg_amd.multigrid")
        if preconditioner.K != op.K or preconditioner.N != op.N:
            raise ValueError("preconditioner was built for a different mesh")
        pc = preconditioner
    return op.pcg(b, x0, alpha=alpha, shift=shift, pc=pc, tol=iterative.tol,
                  atol=iterative.atol, maxiter=iterative.maxiter)


def solve_sipg_poisson_2d(mesh, source, *, bc_types, u_dir=None, q_neu=None, tau_scale=200.0,
                          A_factored=None, iterative=None, x0=None, preconditioner=None,
                          cub=None, return_info=False, refine=None):
    """Solve ``-Delta u = source`` with mixed BCs (sipg.py:812-918).
    ``refine=RefineConfig(...)`` runs the mixed-precision refinement."""
    op = get_operator(mesh, bc_types, tau_scale, cub=cub)
    f, b = _rhs(op, source, u_dir, q_neu)
    if iterative is None:
        if A_factored is None:
            A_factored = factor_sipg_2d(mesh, bc_types=bc_types, tau_scale=tau_scale, cub=cub)
        x = _lu_solve(A_factored, b).to(op.dtype)
        u = f.back(x)
        return (u, None) if return_info else u
    shift = 0.0
    if not np.any(np.asarray(bc_types) == BC_DIRICHLET):
        # singular pure-Neumann operator: eps-shift from one probe apply and a
        # mean-zero right-hand side (sipg.py:885-906)
        n = op.K * op.Np
        probe = torch.sin(torch.arange(n, dtype=torch.float64, device=op.device)).to(op.dtype)
        probe = probe - probe.mean()
        Ap = op.apply(probe)
        shift = 1.0e-6 * float(torch.linalg.norm(Ap) / torch.linalg.norm(probe))
        b = b - b.mean()
    xi = None if x0 is None else _field(x0, op).flat.clone()
    x, info = _run_cg(op, b, xi, iterative, preconditioner, 0.0, shift, refine,
                      (mesh, bc_types, tau_scale))
    u = f.back(x)
    return (u, info) if return_info else u


def solve_sipg_helmholtz_2d(mesh, rhs_field, *, alpha, nu, bc_types, u_dir=None, q_neu=None,
                            tau_scale=200.0, A_factored=None, iterative=None, x0=None,
                            preconditioner=None, cub=None, return_info=False, refine=None):
    """Solve ``(alpha M - Delta) u = rhs_field`` (sipg.py:536-612).
    ``refine=RefineConfig(...)`` runs the mixed-precision refinement."""
    op = get_operator(mesh, bc_types, tau_scale, cub=cub)
    f, b = _rhs(op, rhs_field, u_dir, q_neu)
    if iterative is None:
        if A_factored is None:
            A_factored = factor_sipg_helmholtz_2d(mesh, alpha=alpha, nu=nu, bc_types=bc_types,
                                                  tau_scale=tau_scale, cub=cub)
        u = f.back(_lu_solve(A_factored, b).to(op.dtype))
        return (u, None) if return_info else u
    xi = None if x0 is None else _field(x0, op).flat.clone()
    x, info = _run_cg(op, b, xi, iterative, preconditioner, float(alpha), 0.0, refine,
                      (mesh, bc_types, tau_scale))
    u = f.back(x)
    return (u, info) if return_info else u
