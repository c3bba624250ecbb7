"""Oracle: SIPG Poisson/Helmholtz operator, loads and solves (float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/ddg/dg2d/sipg.py``.  The reference obtains its apply as
``jax.grad`` of the scalar bilinear form (sipg.py:304-337); JAX is absent, so
this module keeps two independent statements:

* :func:`sipg_bilinear_form` -- literal transcription of the form
  (sipg.py:142-301, nodal quadrature and the ``cub`` cubature branch);
* :func:`sipg_apply_2d` -- the reverse-mode gradient of that form with
  respect to ``v``, written out by hand step by step (every gather becomes a
  scatter-add, every matrix a transpose), i.e. what ``jax.grad`` evaluates.

``tests/test_oracle_sipg.py`` pins the second against the first on unit
vectors (``apply(u)_i == a(u, e_i)``: the form is linear in ``v``) and both
against symmetry, SPD-ness and the manufactured solutions of
``test/test_sipg_poisson_2d.py``.

Extension (not in the reference): an optional per-element coefficient
``kappa`` (K,) for ``-div(kappa grad u)``.  Each side of a face uses its own
``kappa`` on its normal derivative and penalty, so ``kappa == 1`` is the
reference exactly and ``kappa == c`` gives ``c * A``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla

from . import reference_element as re
from .cg import cg as jax_cg

BC_INTERIOR = 0
BC_DIRICHLET = 1
BC_NEUMANN = 2


@dataclass(frozen=True)
class IterativeConfig:
    """sipg.py:59-71."""
    method: str = "cg"
    tol: float = 1e-9
    atol: float = 0.0
    maxiter: int = 1000


def local_mass(mesh) -> np.ndarray:
    """M_ref = inv(V V^T) (sipg.py:78-80)."""
    return np.linalg.inv(mesh.V @ mesh.V.T)


def face_1d_mass_matrices(mesh) -> np.ndarray:
    """(3, Nfp, Nfp) per-face 1D mass in Fmask order (sipg.py:83-110)."""
    Fm = np.asarray(mesh.Fmask)
    params = [mesh.r[Fm[:, 0]], mesh.r[Fm[:, 1]], mesh.s[Fm[:, 2]]]
    return np.stack([re.face_mass_1d(mesh.N, p) for p in params])


def _bc_face_arrays(mesh, bc_types):
    """sipg.py:117-139."""
    bpf = np.asarray(bc_types).T.astype(np.int32)                  # (3, K)
    bpfn = np.broadcast_to(bpf[None], (mesh.Nfp, 3, mesh.K)).reshape(-1, order="F")
    w = np.where(bpf == BC_DIRICHLET, 1.0, np.where(bpf == BC_NEUMANN, 0.0, 0.5))
    return bpfn, bpf, w


def _tau_per_face(mesh, tau_scale):
    """tau_scale * (N+1)^2 * Fscale at face node 0 (sipg.py:261-263)."""
    Fs = mesh.Fscale.reshape((mesh.Nfp, 3, mesh.K), order="F")[0]
    return np.asarray(tau_scale, dtype=np.float64) * (mesh.N + 1) ** 2 * Fs


def _grad_xy(mesh, u):
    ur, us = mesh.Dr @ u, mesh.Ds @ u
    return mesh.rx * ur + mesh.sx * us, mesh.ry * ur + mesh.sy * us


def _check_kappa_cub(kappa, cub):
    if kappa is not None and cub is not None:
        raise NotImplementedError("kappa with cubature is not restated")


def face_cub_mass(cub):
    """(3, Nfp, Nfp, K) face mass of every element side under the face
    cubature: L1d^T diag(face_w sJ_q) L1d (the cubature branch's face
    integral, sipg.py:281-299, as a matrix on face-node values)."""
    W = cub.face_w[None, :, None] * cub.face_sJ                   # (3, Nfcub, K)
    return np.einsum("fqp,fqk,fqr->fprk", cub.face_L1d, W, cub.face_L1d)


def sipg_bilinear_form(mesh, u, v, *, bc_types, u_dir=None, tau_scale=200.0,
                       M_ref=None, m1d_per_face=None, kappa=None, cub=None) -> float:
    """Literal a(u, v) (sipg.py:142-301): nodal quadrature (``cub is None``)
    or cubature (``cub``: volume sipg.py:190-201, faces 281-299)."""
    _check_kappa_cub(kappa, cub)
    Np, K, Nfp = mesh.Np, mesh.K, mesh.Nfp
    M = local_mass(mesh) if M_ref is None else M_ref
    m1d = face_1d_mass_matrices(mesh) if m1d_per_face is None else m1d_per_face
    kap = np.ones(K) if kappa is None else np.asarray(kappa, dtype=np.float64)
    ux, uy = _grad_xy(mesh, u)
    vx, vy = _grad_xy(mesh, v)
    ux, uy = kap * ux, kap * uy           # kappa-weighted flux (extension)
    if cub is None:
        vol = np.sum(mesh.J * (ux * (M @ vx) + uy * (M @ vy)))
    else:
        urq, usq, vrq, vsq = cub.Drq @ u, cub.Dsq @ u, cub.Drq @ v, cub.Dsq @ v
        uxq, uyq = cub.rx * urq + cub.sx * usq, cub.ry * urq + cub.sy * usq
        vxq, vyq = cub.rx * vrq + cub.sx * vsq, cub.ry * vrq + cub.sy * vsq
        vol = np.sum(cub.w[:, None] * cub.J * (uxq * vxq + uyq * vyq))
    uf, vf = u.T.reshape(-1), v.T.reshape(-1)
    uxf, uyf = ux.T.reshape(-1), uy.T.reshape(-1)
    vxf, vyf = (kap * vx).T.reshape(-1), (kap * vy).T.reshape(-1)
    nx, ny = mesh.nx.reshape(-1, order="F"), mesh.ny.reshape(-1, order="F")
    vM, vP = mesh.vmapM, mesh.vmapP
    uM, uPr, vMv, vPr = uf[vM], uf[vP], vf[vM], vf[vP]
    DuM = nx * uxf[vM] + ny * uyf[vM]
    DuPr = nx * uxf[vP] + ny * uyf[vP]
    DvM = nx * vxf[vM] + ny * vyf[vM]
    DvPr = nx * vxf[vP] + ny * vyf[vP]
    bfn, _, w = _bc_face_arrays(mesh, bc_types)
    isd, isn = bfn == BC_DIRICHLET, bfn == BC_NEUMANN
    isb = isd | isn
    uPd = np.zeros_like(uM) if u_dir is None else u_dir.T.reshape(-1)[vM]
    uP = np.where(isd, uPd, np.where(isn, uM, uPr))
    vP_ = np.where(isd, 0.0, np.where(isn, vMv, vPr))
    DuP = np.where(isb, DuM, DuPr)
    DvP = np.where(isb, DvM, DvPr)
    face = lambda a: a.reshape((Nfp, 3, K), order="F")
    avgu, avgv = face(0.5 * (DuM + DuP)), face(0.5 * (DvM + DvP))
    ju, jv = face(uM - uP), face(vMv - vP_)
    sJ = face(mesh.sJ.reshape(-1, order="F"))[0]
    tau = _tau_per_face(mesh, tau_scale) * kap[None, :]
    if cub is None:
        wsJ = w * sJ
        cons = -np.einsum("fk,pfk,fpq,qfk->", wsJ, avgu, m1d, jv)
        adj = -np.einsum("fk,pfk,fpq,qfk->", wsJ, ju, m1d, avgv)
        pen = np.einsum("fk,pfk,fpq,qfk->", w * tau * sJ, ju, m1d, jv)
    else:
        to_q = lambda a: np.einsum("fqp,pfk->qfk", cub.face_L1d, a)
        wts = cub.face_w[:, None, None] * np.transpose(cub.face_sJ, (1, 0, 2)) * w[None]
        cons = -np.sum(wts * to_q(avgu) * to_q(jv))
        adj = -np.sum(wts * to_q(ju) * to_q(avgv))
        pen = np.sum(wts * tau[None] * to_q(ju) * to_q(jv))
    return float(vol + cons + adj + pen)


def sipg_apply_2d(mesh, u, *, bc_types, u_dir=None, tau_scale=200.0,
                  M_ref=None, m1d_per_face=None, kappa=None, cub=None) -> np.ndarray:
    """d a(u, v) / d v, written as the reverse sweep of the form
    (sipg.py:304-337: ``jax.grad(form)(zeros)``).  Returns (Np, K)."""
    _check_kappa_cub(kappa, cub)
    Np, K, Nfp = mesh.Np, mesh.K, mesh.Nfp
    M = local_mass(mesh) if M_ref is None else M_ref
    m1d = face_1d_mass_matrices(mesh) if m1d_per_face is None else m1d_per_face
    kap = np.ones(K) if kappa is None else np.asarray(kappa, dtype=np.float64)
    # forward quantities that the gradient needs (u side only)
    ux, uy = _grad_xy(mesh, u)
    ux, uy = kap * ux, kap * uy
    uf = u.T.reshape(-1)
    uxf, uyf = ux.T.reshape(-1), uy.T.reshape(-1)
    nx, ny = mesh.nx.reshape(-1, order="F"), mesh.ny.reshape(-1, order="F")
    vM, vP = mesh.vmapM, mesh.vmapP
    uM, uPr = uf[vM], uf[vP]
    DuM = nx * uxf[vM] + ny * uyf[vM]
    DuPr = nx * uxf[vP] + ny * uyf[vP]
    bfn, _, w = _bc_face_arrays(mesh, bc_types)
    isd, isn = bfn == BC_DIRICHLET, bfn == BC_NEUMANN
    isb = isd | isn
    uPd = np.zeros_like(uM) if u_dir is None else u_dir.T.reshape(-1)[vM]
    uP = np.where(isd, uPd, np.where(isn, uM, uPr))
    DuP = np.where(isb, DuM, DuPr)
    face = lambda a: a.reshape((Nfp, 3, K), order="F")
    flat = lambda a: a.reshape(-1, order="F")
    avgu = face(0.5 * (DuM + DuP))
    ju = face(uM - uP)
    sJ = face(mesh.sJ.reshape(-1, order="F"))[0]
    tau = _tau_per_face(mesh, tau_scale) * kap[None, :]
    # face mass of each element side: sJ m1d (nodal) or the face cubature's
    # L1d^T diag(w sJ_q) L1d; cotangents of jmp_v and avg_Dn_v
    if cub is None:
        mf = sJ[:, None, None, :] * m1d[:, :, :, None]              # (3, Nfp, Nfp, K)
    else:
        mf = face_cub_mass(cub)
    g_jv = (-w[None] * np.einsum("fpqk,pfk->qfk", mf, avgu)
            + (w * tau)[None] * np.einsum("fpqk,pfk->qfk", mf, ju))
    g_av = -w[None] * np.einsum("fpqk,pfk->qfk", mf, ju)
    g_jv, g_av = flat(g_jv), flat(g_av)
    # jmp_v = v_M - v_P_eff ; v_P_eff = where(dir, 0, where(neu, v_M, v_P_raw))
    g_vM = g_jv + np.where(isn, -g_jv, 0.0)
    g_vPr = np.where(isb, 0.0, -g_jv)
    # avg = (Dn_v_M + Dn_v_P_eff)/2 ; Dn_v_P_eff = where(bdy, Dn_v_M, Dn_v_P_raw)
    g_DvM = 0.5 * g_av + np.where(isb, 0.5 * g_av, 0.0)
    g_DvPr = np.where(isb, 0.0, 0.5 * g_av)
    n = Np * K
    sc = lambda idx, val: np.bincount(idx, weights=val, minlength=n)
    g_vf = sc(vM, g_vM) + sc(vP, g_vPr)
    g_vxf = sc(vM, nx * g_DvM) + sc(vP, nx * g_DvPr)
    g_vyf = sc(vM, ny * g_DvM) + sc(vP, ny * g_DvPr)
    unflat = lambda a: a.reshape(K, Np).T
    # vx, vy were kappa-weighted in the face terms only
    g_vx = kap * unflat(g_vxf)
    g_vy = kap * unflat(g_vyf)
    out = unflat(g_vf)
    if cub is None:
        g_vx = g_vx + M.T @ (mesh.J * ux)
        g_vy = g_vy + M.T @ (mesh.J * uy)
    else:
        # volume cubature: d/dv of sum w J (ux_q vx_q + uy_q vy_q)
        urq, usq = cub.Drq @ u, cub.Dsq @ u
        cx = cub.w[:, None] * cub.J * (cub.rx * urq + cub.sx * usq)
        cy = cub.w[:, None] * cub.J * (cub.ry * urq + cub.sy * usq)
        out = out + cub.Drq.T @ (cub.rx * cx + cub.ry * cy) + cub.Dsq.T @ (cub.sx * cx + cub.sy * cy)
    g_vr = mesh.rx * g_vx + mesh.ry * g_vy
    g_vs = mesh.sx * g_vx + mesh.sy * g_vy
    return out + mesh.Dr.T @ g_vr + mesh.Ds.T @ g_vs


def mass_apply(mesh, u, M_ref=None, cub=None) -> np.ndarray:
    """J * (M_ref @ u) (sipg.py:565, 597, 853); cubature mass with ``cub``
    (cubature.py:242-245)."""
    if cub is not None:
        return cub.Lq.T @ (cub.w[:, None] * cub.J * (cub.Lq @ u))
    M = local_mass(mesh) if M_ref is None else M_ref
    return mesh.J * (M @ u)


def flat_apply(mesh, bc_types, tau_scale=200.0, alpha=0.0, kappa=None, cub=None):
    """Element-major flat operator alpha*M + A (multigrid.py:53-74;
    sipg.py:589-601 with ``cub``)."""
    M = local_mass(mesh)
    m1d = face_1d_mass_matrices(mesh)
    Np, K = mesh.Np, mesh.K

    def apply(uf):
        u = uf.reshape(K, Np).T
        out = sipg_apply_2d(mesh, u, bc_types=bc_types, tau_scale=tau_scale,
                            M_ref=M, m1d_per_face=m1d, kappa=kappa, cub=cub)
        if alpha != 0.0:
            out = out + alpha * mass_apply(mesh, u, M, cub)
        return out.T.reshape(-1)
    return apply


def sipg_matrix_2d(mesh, *, bc_types, tau_scale=200.0, alpha=0.0, kappa=None,
                   cub=None) -> np.ndarray:
    """Dense (Np K)^2 matrix, column by column (sipg.py:344-364, 460-496)."""
    n = mesh.Np * mesh.K
    ap = flat_apply(mesh, bc_types, tau_scale, alpha, kappa, cub)
    A = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = ap(e)
        e[j] = 0.0
    return A


def factor_sipg_2d(mesh, *, bc_types, tau_scale=200.0, regularize=False,
                   eps=1.0e-8, alpha=0.0, cub=None):
    """sipg.py:367-391 (and 518-533 with alpha)."""
    A = sipg_matrix_2d(mesh, bc_types=bc_types, tau_scale=tau_scale, alpha=alpha, cub=cub)
    if regularize:
        A = A + eps * np.max(np.abs(np.diag(A))) * np.eye(A.shape[0])
    lu, piv = sla.lu_factor(A)
    return A, lu, piv


def sipg_dirichlet_load_2d(mesh, u_dir, *, bc_types, tau_scale=200.0) -> np.ndarray:
    """A_aff(0; u_dir) (sipg.py:398-417)."""
    return sipg_apply_2d(mesh, np.zeros_like(u_dir), bc_types=bc_types,
                         u_dir=u_dir, tau_scale=tau_scale)


def sipg_neumann_load_2d(mesh, q_neu, *, bc_types) -> np.ndarray:
    """sum over Neumann faces of sJ * m1d q scattered to Fmask rows
    (sipg.py:420-453)."""
    Np, K, Nfp = mesh.Np, mesh.K, mesh.Nfp
    m1d = face_1d_mass_matrices(mesh)
    _, bpf, _ = _bc_face_arrays(mesh, bc_types)
    isn = (bpf == BC_NEUMANN).astype(np.float64)
    q = q_neu.reshape((Nfp, 3, K), order="F")
    sJ = mesh.sJ.reshape((Nfp, 3, K), order="F")[0]
    c = np.einsum("fpq,qfk->pfk", m1d, q) * (sJ * isn)[None]
    out = np.zeros((Np, K))
    Fm = np.asarray(mesh.Fmask)
    for f in range(3):
        out[Fm[:, f], :] += c[:, f, :]
    return out


def _rhs(mesh, source, bc_types, u_dir, q_neu, tau_scale, cub=None):
    """Mf - L_D + L_N (sipg.py:849-864); the mass is the cubature mass with
    ``cub``, the loads are always the nodal ones (sipg.py:398-453)."""
    rhs = mass_apply(mesh, source, cub=cub)
    if u_dir is not None:
        rhs = rhs - sipg_dirichlet_load_2d(mesh, u_dir, bc_types=bc_types, tau_scale=tau_scale)
    if q_neu is not None:
        rhs = rhs + sipg_neumann_load_2d(mesh, q_neu, bc_types=bc_types)
    return rhs.T.reshape(-1)


def solve_sipg_poisson_2d(mesh, source, *, bc_types, u_dir=None, q_neu=None,
                          tau_scale=200.0, A_factored=None, iterative=None,
                          x0=None, preconditioner=None, return_info=False, cub=None):
    """sipg.py:812-918."""
    Np, K = mesh.Np, mesh.K
    b = _rhs(mesh, source, bc_types, u_dir, q_neu, tau_scale, cub)
    if iterative is None:
        if A_factored is None:
            A_factored = factor_sipg_2d(mesh, bc_types=bc_types, tau_scale=tau_scale, cub=cub)
        _, lu, piv = A_factored
        return sla.lu_solve((lu, piv), b).reshape(K, Np).T
    raw = flat_apply(mesh, bc_types, tau_scale, cub=cub)
    if not np.any(np.asarray(bc_types) == BC_DIRICHLET):
        probe = np.sin(np.arange(Np * K, dtype=np.float64))
        probe = probe - probe.mean()
        shift = 1.0e-6 * np.linalg.norm(raw(probe)) / np.linalg.norm(probe)
        matvec = lambda x: raw(x) + shift * x
        b = b - b.mean()
    else:
        matvec = raw
    if iterative.method != "cg":
        raise ValueError(f"iterative method {iterative.method!r} not supported for SPD")
    xi = np.zeros(Np * K) if x0 is None else x0.T.reshape(-1)
    x, info = jax_cg(matvec, b, xi, tol=iterative.tol, atol=iterative.atol,
                     maxiter=iterative.maxiter, M=preconditioner)
    u = x.reshape(K, Np).T
    return (u, info) if return_info else u


def solve_sipg_helmholtz_2d(mesh, rhs_field, *, alpha, nu, bc_types, u_dir=None,
                            q_neu=None, tau_scale=200.0, A_factored=None,
                            iterative=None, x0=None, preconditioner=None,
                            return_info=False, cub=None):
    """(alpha M + A) u = M f - L_D + L_N; ``nu`` unused (sipg.py:536-612)."""
    Np, K = mesh.Np, mesh.K
    b = _rhs(mesh, rhs_field, bc_types, u_dir, q_neu, tau_scale, cub)
    if iterative is None:
        if A_factored is None:
            A_factored = factor_sipg_2d(mesh, bc_types=bc_types, tau_scale=tau_scale,
                                        alpha=alpha, cub=cub)
        _, lu, piv = A_factored
        return sla.lu_solve((lu, piv), b).reshape(K, Np).T
    if iterative.method != "cg":
        raise ValueError(f"iterative method {iterative.method!r} not supported for SPD")
    matvec = flat_apply(mesh, bc_types, tau_scale, alpha, cub=cub)
    xi = np.zeros(Np * K) if x0 is None else x0.T.reshape(-1)
    x, info = jax_cg(matvec, b, xi, tol=iterative.tol, atol=iterative.atol,
                     maxiter=iterative.maxiter, M=preconditioner)
    u = x.reshape(K, Np).T
    return (u, info) if return_info else u


# --------------------------------------------------------------------------
# Block-Jacobi (sipg.py:636-805)
# --------------------------------------------------------------------------

def color_elements(mesh):
    """Greedy colouring of face adjacency (sipg.py:651-668)."""
    K = mesh.K
    E = np.asarray(mesh.EToE)
    col = -np.ones(K, dtype=np.int64)
    for k in range(K):
        used = {col[nb] for nb in E[k] if nb != k and col[nb] >= 0}
        c = 0
        while c in used:
            c += 1
        col[k] = c
    return col, int(col.max()) + 1


def extract_block_diag_matfree(apply_flat, mesh) -> np.ndarray:
    """Colour-probe diagonal blocks (sipg.py:671-704).  (K, Np, Np)."""
    Np, K = mesh.Np, mesh.K
    col, nc = color_elements(mesh)
    blocks = np.zeros((K, Np, Np))
    for c in range(nc):
        mask = (col == c).astype(np.float64)
        for j in range(Np):
            probe = np.zeros((K, Np))
            probe[:, j] = mask
            r = apply_flat(probe.reshape(-1)).reshape(K, Np)
            blocks[:, :, j] += np.where(mask[:, None] > 0, r, 0.0)
    return blocks


def extract_block_diag_dense(A, Np, K) -> np.ndarray:
    """sipg.py:636-648."""
    rows = np.arange(K)[:, None] * Np + np.arange(Np)[None, :]
    return A[rows[:, :, None], rows[:, None, :]]


def block_jacobi_apply(blocks):
    """M^-1 v = einsum('kij,kj->ki', inv(blocks), v) (sipg.py:707-715)."""
    Binv = np.linalg.inv(blocks)
    K, Np, _ = blocks.shape
    return lambda vf: np.einsum("kij,kj->ki", Binv, vf.reshape(K, Np)).reshape(-1)


def make_block_jacobi_preconditioner(mesh, *, bc_types, tau_scale=200.0, alpha=0.0,
                                     matfree=True, A_dense=None):
    """sipg.py:718-805 (Poisson alpha=0, Helmholtz alpha>0)."""
    if matfree:
        blocks = extract_block_diag_matfree(flat_apply(mesh, bc_types, tau_scale, alpha), mesh)
    else:
        if A_dense is None:
            A_dense = sipg_matrix_2d(mesh, bc_types=bc_types, tau_scale=tau_scale, alpha=alpha)
        blocks = extract_block_diag_dense(A_dense, mesh.Np, mesh.K)
    return block_jacobi_apply(blocks)
