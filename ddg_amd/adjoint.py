"""Differentiable SIPG solves: gradients of a loss of the solution with
respect to the per-element diffusion coefficient ``kappa``, the per-element
penalty scale ``tau_scale`` and the source ``f``, by the adjoint method.

The reference differentiates its solves with ``jax.grad`` through
``lax.custom_linear_solve(symmetric=True)`` (src/ddg/dg2d/sipg.py:604-609,
910-915): the backward pass is one more CG solve with the same operator.
Here ``SipgSolve`` (a ``torch.autograd.Function``) does the same on the
device -- forward ``A(kappa, tau) u = b``, backward ``A lam = dL/du`` with
the same operator and preconditioner -- and turns ``(u, lam)`` into the
parameter gradients with one kernel (``ddgx_sipg_param_grad``):

    dL/dkappa_k = -d a(u, lam; u_dir)/d kappa_k
    dL/dtau_k   = -d a(u, lam; u_dir)/d tau_scale_k
    dL/df       = (J M)^T lam          (b = J M f + boundary data)

Everything runs in float64.  JAX is not installed in this environment, so
the ``jax.custom_vjp`` twin of this function is described (not built) in
the integration notes.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .operator import SipgOperator, _ptr, _stream


@dataclass
class SolveSpec:
    """Everything of a solve that is not differentiated."""
    mesh: object
    bc_types: np.ndarray
    u_dir: torch.Tensor | None = None     # (K*Np,) float64 device, element-major
    q_neu: torch.Tensor | None = None     # (K*3*Nfp,) float64 device
    preconditioner: str | None = "mg"     # "mg", "bj" or None
    tol: float = 1e-12
    atol: float = 0.0
    maxiter: int = 2000


def param_grad(op: SipgOperator, u, lam, kappa, tau_scale, u_dir=None):
    """``(d a(u, lam; u_dir)/d kappa, d a/d tau_scale)`` per element."""
    if op.dtype != torch.float64:
        raise ValueError("parameter gradients need a float64 operator")
    K = op.K
    kap = None if kappa is None else kappa.detach().to(op.device, torch.float64).contiguous()
    tau = torch.as_tensor(tau_scale, dtype=torch.float64, device=op.device)
    tau = torch.broadcast_to(tau, (K,)).contiguous()
    if kap is not None and kap.numel() != K:
        raise ValueError("kappa must have K entries")
    gk = torch.empty(K, dtype=torch.float64, device=op.device)
    gt = torch.empty_like(gk)
    u, lam = op._vec(u, "u"), op._vec(lam, "lam")
    _lib.check(_lib.lib().ddgx_sipg_param_grad(op.handle, _ptr(u), _ptr(lam),
                                               _ptr(op._vec(u_dir, "u_dir")), _ptr(kap),
                                               _ptr(tau), _ptr(gk), _ptr(gt), _stream()),
               "ddgx_sipg_param_grad")
    return gk, gt


def _preconditioner(op, spec):
    if spec.preconditioner is None:
        return None
    if spec.preconditioner == "bj":
        return op.block_jacobi(0.0)
    if spec.preconditioner == "mg":
        # p-multigrid N -> N/2 -> 1 on operators of this kappa / tau; the fine
        # level is the solve's own operator
        from .multigrid import Multigrid
        from . import mesh as pm
        from .reference import p_interp_matrix
        N = op.N
        degs = sorted({N, max(1, N // 2), 1}, reverse=True)
        m = spec.mesh
        meshes = [m if d == N else pm.build_mesh_2d(m.VX, m.VY, m.EToV, d) for d in degs]
        ops = [op] + [SipgOperator(mm, spec.bc_types, op.tau_scale, op.kappa)
                      for mm in meshes[1:]]
        levels = [(mm, spec.bc_types, op.kappa) for mm in meshes]
        mats = [p_interp_matrix(lo, hi) for hi, lo in zip(degs, degs[1:])]
        return Multigrid(levels, ["p"] * (len(degs) - 1), mats, ops=ops)
    raise ValueError(f"unknown preconditioner {spec.preconditioner!r}")


class SipgSolve(torch.autograd.Function):
    """``u = A(kappa, tau_scale)^{-1} (J M f + boundary data)``, flat (K*Np,)."""

    @staticmethod
    def forward(ctx, kappa, tau_scale, f, spec: SolveSpec):
        m = spec.mesh
        K = int(m.K)
        kap_np = kappa.detach().cpu().numpy().astype(np.float64)
        tau_np = np.broadcast_to(tau_scale.detach().cpu().numpy().astype(np.float64), (K,)).copy()
        op = SipgOperator(m, spec.bc_types, tau_np, kap_np)
        op.kappa, op.tau_scale = kap_np, tau_np
        pc = _preconditioner(op, spec)
        b = op.rhs(f.detach().contiguous(), u_dir=spec.u_dir, q_neu=spec.q_neu)
        u, info = op.pcg(b, pc=pc, tol=spec.tol, atol=spec.atol, maxiter=spec.maxiter)
        ctx.op, ctx.pc, ctx.spec, ctx.info = op, pc, spec, info
        ctx.save_for_backward(kappa, tau_scale, u)
        return u

    @staticmethod
    def backward(ctx, gu):
        kappa, tau_scale, u = ctx.saved_tensors
        op, spec = ctx.op, ctx.spec
        lam, _ = op.pcg(gu.detach().contiguous().to(torch.float64), pc=ctx.pc, tol=spec.tol,
                        atol=spec.atol, maxiter=spec.maxiter)
        gk, gt = param_grad(op, u, lam, kappa, torch.as_tensor(op.tau_scale, device=op.device),
                            spec.u_dir)
        gf = op.rhs(lam)
        g_tau = -gt
        if tau_scale.dim() == 0:
            g_tau = g_tau.sum()
        return -gk, g_tau, gf, None


def sipg_solve(kappa, tau_scale, f, spec: SolveSpec):
    """Differentiable solve (see ``SipgSolve``)."""
    return SipgSolve.apply(kappa, tau_scale, f, spec)


class MassApply(torch.autograd.Function):
    """``J M u`` per element (symmetric: the backward is the same apply)."""

    @staticmethod
    def forward(ctx, u, op):
        ctx.op = op
        return op.rhs(u.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.op.rhs(g.detach().contiguous()), None


def mass_apply(u, op):
    return MassApply.apply(u, op)
