"""Drop-in for ``ddg.dg2d.multigrid``: p- and hp-multigrid V-cycle
preconditioners for the SIPG Poisson/Helmholtz operators, run on the device
by libddgx (``ddgx_mg_create``).

Same builders, keywords and defaults as ``src/ddg/dg2d/multigrid.py:253-367``:
geometric (rediscretised) levels, damped block-Jacobi smoothing
(``nu_smooth`` sweeps, ``omega``), exact nodal p-transfers, red-child
h-transfers on a uniform forest hierarchy, and a direct solve of the
coarsest level.  The returned object is callable on device vectors (the
reference's ``M^{-1}(v_flat)`` protocol) and is accepted by
``ddg_amd.sipg.solve_*`` and ``SipgOperator.pcg`` as ``preconditioner``.

Extensions (defaults reproduce the reference):
* ``degrees`` -- explicit fine-to-coarse degree list, e.g. ``[4, 2, 1]``;
* ``dtype`` -- precision of the cycle (``torch.float32`` for the mixed-
  precision solver; blocks and the coarsest inverse are always built in
  float64 and the coarsest solve runs in float64);
* ``kappa`` -- per-element coefficient (p-multigrid only);
* ``fine_vertices`` -- replacement vertex coordinates of the finest
  hp-level mesh (a jittered mesh); coarser levels use the same coordinates
  for the vertices they share.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import mesh as pm
from .operator import DevicePreconditioner, _stream
from .reference import h_child_interp_matrices, p_interp_matrix


class Multigrid(DevicePreconditioner):
    """V-cycle over levels ``[(mesh, bc_types, kappa), ...]`` (fine to coarse)
    with transfer kinds ``kinds[l]`` ("p" or "h") and host transfer matrices
    ``mats[l]``.  A float32 cycle preconditions float32 solves, the inner
    solves of ``refine=`` and, directly, float64 solves (mixed-precision
    preconditioning)."""

    is_multigrid = True

    def __init__(self, levels, kinds, mats, *, tau_scale=200.0, alpha=0.0, nu_smooth=2,
                 omega=0.7, regularize=False, eps=1.0e-8, dtype=torch.float64, ops=None):
        """``ops``: optional prebuilt float64 level operators (then ``dtype``
        must be float64 and ``levels`` only names the meshes); otherwise the
        operators come from the per-mesh cache of ``ddg_amd.sipg``."""
        from .sipg import get_operator
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be torch.float64 or torch.float32")
        self._levels = levels  # keeps the level meshes (and their cached operators) alive
        if ops is not None:
            if dtype != torch.float64 or len(ops) != len(levels):
                raise ValueError("prebuilt level operators must be float64, one per level")
            self.ops = list(ops)
        else:
            self.ops = [get_operator(m, bc, tau_scale, kap, dtype) for m, bc, kap in levels]
        self.ops64 = (self.ops if dtype == torch.float64 else
                      [get_operator(m, bc, tau_scale, kap, torch.float64) for m, bc, kap in levels])
        self.dtype = dtype
        self.K, self.N = self.ops[0].K, self.ops[0].N
        self.op = self.ops[0]
        self.alpha = float(alpha)
        n = len(levels)
        H = (C.c_void_p * n)(*[o.handle.value for o in self.ops])
        H64 = (C.c_void_p * n)(*[o.handle.value for o in self.ops64])
        self._mats = [np.ascontiguousarray(m, dtype=np.float64) for m in mats]
        kind = (C.c_int32 * max(1, n - 1))(*[{"p": 0, "h": 1}[k] for k in kinds])
        P = (C.POINTER(C.c_double) * max(1, n - 1))(
            *[m.ctypes.data_as(C.POINTER(C.c_double)) for m in self._mats])
        cfg = _lib.MGConfig(int(nu_smooth), float(omega), float(alpha), int(bool(regularize)),
                            float(eps))
        h = C.c_void_p()
        _lib.check(_lib.lib().ddgx_mg_create(n, H, H64, kind, P, C.byref(cfg), _stream(),
                                             C.byref(h)), "ddgx_mg_create")
        self._h = h
        self.degrees = [int(m.N) for m, _, _ in levels]
        self.sizes = [int(m.K) for m, _, _ in levels]


def _operator_apply(mesh, bc_types, tau_scale, alpha):
    """Flat apply of ``alpha M + A`` on device vectors (multigrid.py:53-74),
    one ``ddgx_sipg_apply`` per call."""
    from .sipg import get_operator
    op = get_operator(mesh, bc_types, tau_scale)
    return lambda u_flat: op.apply(u_flat.to(op.dtype).contiguous(), alpha=float(alpha))


def _p_interp_matrix(mesh_lo, mesh_hi) -> np.ndarray:
    """(Np_hi, Np_lo) nodal interpolation (multigrid.py:106-114)."""
    return p_interp_matrix(int(mesh_lo.N), int(mesh_hi.N))


def _h_child_interp_matrices(mesh) -> np.ndarray:
    """(4, Np, Np) red-child interpolation (multigrid.py:137-163)."""
    return h_child_interp_matrices(int(mesh.N))


def _build_h_meshes(base_mesh, n_levels):
    """Meshes of the uniform-refinement hierarchy at the base mesh's order,
    coarse to fine (multigrid.py:187-204)."""
    return [pm.build_mesh_2d(vx, vy, e, int(base_mesh.N))
            for vx, vy, e in build_h_meshes(base_mesh, n_levels)]


def _check_p_coarse(p_coarse, p_fine):
    if p_coarse < 1 or p_coarse > p_fine:
        raise ValueError(f"p_coarse must be in [1, {p_fine}], got {p_coarse}")


def _degrees(p_fine, p_coarse, degrees):
    if degrees is None:
        return list(range(p_fine, p_coarse - 1, -1))
    d = [int(x) for x in degrees]
    if d[0] != p_fine or d[-1] != p_coarse or any(b >= a for a, b in zip(d, d[1:])):
        raise ValueError(f"degrees must fall strictly from {p_fine} to {p_coarse}, got {d}")
    return d


def make_sipg_mg_preconditioner(mesh, *, bc_types, tau_scale=200.0, alpha=0.0, p_coarse=1,
                                nu_smooth=2, omega=0.7, regularize=False, eps=1.0e-8,
                                degrees=None, kappa=None, dtype=torch.float64) -> Multigrid:
    """p-multigrid V-cycle preconditioner (multigrid.py:253-294)."""
    p_fine = int(mesh.N)
    _check_p_coarse(p_coarse, p_fine)
    degs = _degrees(p_fine, p_coarse, degrees)
    bc = np.asarray(bc_types, dtype=np.int32)
    meshes = [mesh if d == p_fine else pm.build_mesh_2d(mesh.VX, mesh.VY, mesh.EToV, d)
              for d in degs]
    levels = [(m, bc, kappa) for m in meshes]
    mats = [p_interp_matrix(lo, hi) for hi, lo in zip(degs, degs[1:])]
    return Multigrid(levels, ["p"] * (len(degs) - 1), mats, tau_scale=tau_scale, alpha=alpha,
                     nu_smooth=nu_smooth, omega=omega, regularize=regularize, eps=eps,
                     dtype=dtype)


def build_h_meshes(base_mesh, n_levels):
    """Uniform-refinement hierarchy ``[(VX, VY, EToV), ...]`` from the base
    mesh to the finest (multigrid.py:187-204): element 4k+c of a level is the
    c-th red child of element k of the level below."""
    out = [(np.asarray(base_mesh.VX, dtype=np.float64), np.asarray(base_mesh.VY, dtype=np.float64),
            np.asarray(base_mesh.EToV, dtype=np.int64))]
    for _ in range(n_levels):
        out.append(pm.refine_uniform(*out[-1]))
    return out


def make_sipg_hpmg_preconditioner(base_mesh, n_h_levels, *, bc_fn, tau_scale=200.0, alpha=0.0,
                                  p_coarse=1, nu_smooth=2, omega=0.7, regularize=False,
                                  eps=1.0e-8, degrees=None, fine_vertices=None,
                                  dtype=torch.float64):
    """hp-multigrid V-cycle preconditioner (multigrid.py:301-367).
    Returns ``(preconditioner, finest_mesh)``."""
    if np.ndim(tau_scale) != 0:
        raise ValueError("hp-multigrid takes a scalar tau_scale (levels differ in K)")
    hier = build_h_meshes(base_mesh, n_h_levels)
    if fine_vertices is not None:
        FX, FY = (np.asarray(a, dtype=np.float64) for a in fine_vertices)
        if FX.shape != hier[-1][0].shape or FY.shape != hier[-1][1].shape:
            raise ValueError("fine_vertices must match the finest level's vertex count")
        hier = [(FX[:vx.size].copy(), FY[:vx.size].copy(), E) for vx, _, E in hier]
    N = int(base_mesh.N)
    fVX, fVY, fE = hier[-1]
    finest = pm.build_mesh_2d(fVX, fVY, fE, N)
    _check_p_coarse(p_coarse, N)
    degs = _degrees(N, p_coarse, degrees)
    bc_fine = np.asarray(bc_fn(finest), dtype=np.int32)
    p_meshes = [finest if d == N else pm.build_mesh_2d(fVX, fVY, fE, d) for d in degs]
    levels = [(m, bc_fine, None) for m in p_meshes]
    for hVX, hVY, hE in hier[-2::-1]:
        hm = pm.build_mesh_2d(hVX, hVY, hE, p_coarse)
        levels.append((hm, np.asarray(bc_fn(hm), dtype=np.int32), None))
    mats = [p_interp_matrix(lo, hi) for hi, lo in zip(degs, degs[1:])]
    Ic = h_child_interp_matrices(p_coarse)
    mats += [Ic] * n_h_levels
    kinds = ["p"] * (len(degs) - 1) + ["h"] * n_h_levels
    pc = Multigrid(levels, kinds, mats, tau_scale=tau_scale, alpha=alpha, nu_smooth=nu_smooth,
                   omega=omega, regularize=regularize, eps=eps, dtype=dtype)
    return pc, finest
