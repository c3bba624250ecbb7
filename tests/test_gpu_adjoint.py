"""Adjoint gradients on the device.

* The parameter-gradient kernel against the oracle's exact unit-perturbation
  derivatives of the literal form: 1e-9 relative (max norm).
* The autograd solve (forward solve, adjoint solve, kernel) against the
  oracle's LU-based adjoint gradient: 1e-7 relative.
* The kappa/tau gradient of a misfit loss on 2x64x64, N=3, per-element
  log-normal kappa and per-element tau against centred finite differences
  (eps = 1e-6) on 20 seeded random unit directions: |fd - g.d| <= 5e-5 |g|
  (measured ~5e-6; the differences of two solves to tol 1e-14 carry
  round-off / eps).  The solves must converge: the iterate of an unconverged
  CG is not a smooth function of the parameters.  At tau ~ 200 the loss
  hardly depends on tau (|dL/dtau| ~ 1e-7 |dL/dkappa|), so these directions
  test kappa; tau is pinned by the two tests above.
"""
import numpy as np
import pytest
import torch

from conftest import rel
from oracle import adjoint as OA
from oracle import mesh as om
from oracle import sipg as S
from ddg_amd import mesh as pmesh
from ddg_amd.adjoint import SolveSpec, mass_apply, param_grad, sipg_solve
from ddg_amd.operator import SipgOperator

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _flat(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a).T).reshape(-1), device=dev)


def _meshes(n, N, seed):
    VX, VY, EToV = pmesh.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, n, n)
    VX, VY = pmesh.jitter_vertices(VX, VY, 1.0 / n, 0.1, seed)
    return pmesh.build_mesh_2d(VX, VY, EToV, N), om.build_mesh_2d(VX, VY, EToV, N)


@pytest.mark.parametrize("N", [1, 3, 5])
def test_param_grad_kernel_matches_oracle(dev, N):
    pm, omesh = _meshes(4, N, 30 + N)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    rng = np.random.default_rng(N)
    kap = np.exp(0.5 * rng.standard_normal(pm.K))
    tau = rng.uniform(100.0, 300.0, pm.K)
    u, lam = rng.standard_normal((pm.Np, pm.K)), rng.standard_normal((pm.Np, pm.K))
    ud = rng.standard_normal((pm.Np, pm.K))
    _say(f"param grad kernel N={N}")
    op = SipgOperator(pm, bc, tau, kap)
    gk, gt = param_grad(op, _flat(u, dev), _flat(lam, dev), torch.as_tensor(kap, device=dev),
                        torch.as_tensor(tau, device=dev), _flat(ud, dev))
    rk, rt = OA.form_param_grad(omesh, u, lam, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud)
    assert rel(gk.cpu().numpy(), rk) < 1e-9
    assert rel(gt.cpu().numpy(), rt) < 1e-9


def test_autograd_matches_oracle_adjoint(dev):
    pm, omesh = _meshes(3, 2, 0)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    rng = np.random.default_rng(0)
    kap = np.exp(0.5 * rng.standard_normal(pm.K))
    tau = rng.uniform(8.0, 12.0, pm.K)
    f = np.sin(2 * pm.x + 1) * np.cos(pm.y)
    ud = np.exp(pm.x) * np.cos(3 * pm.y)
    xf = pm.x.reshape(-1, order="F")[pm.vmapM]
    q = (0.3 + xf).reshape((3 * pm.Nfp, pm.K), order="F")
    kap_t = kap * np.exp(0.2 * rng.standard_normal(pm.K))
    target, _ = OA.solve_dense(omesh, f, bc_types=bc, kappa=kap_t, tau_scale=tau, u_dir=ud,
                               q_neu=q)
    L0, u = OA.misfit_loss(omesh, f, target, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud,
                           q_neu=q)
    _, A = OA.solve_dense(omesh, f, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud, q_neu=q)
    lam = np.linalg.solve(A, S.mass_apply(omesh, u - target).T.reshape(-1)).reshape(pm.K, pm.Np).T
    rk, rt = OA.form_param_grad(omesh, u, lam, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud)
    rf = S.mass_apply(omesh, lam)
    _say("autograd small")
    spec = SolveSpec(pm, bc, u_dir=_flat(ud, dev), q_neu=_flat(q, dev), preconditioner="bj",
                     tol=1e-14, maxiter=3000)
    kt = torch.tensor(kap, device=dev, requires_grad=True)
    tt = torch.tensor(tau, device=dev, requires_grad=True)
    ft = _flat(f, dev).requires_grad_(True)
    uu = sipg_solve(kt, tt, ft, spec)
    e = uu - _flat(target, dev)
    L = 0.5 * torch.dot(e, mass_apply(e, spec_op := SipgOperator(pm, bc)))
    L.backward()
    assert abs(L.item() - L0) < 1e-9 * L0
    assert rel(kt.grad.cpu().numpy(), -rk) < 1e-7
    assert rel(tt.grad.cpu().numpy(), -rt) < 1e-7
    assert rel(ft.grad.cpu().numpy(), rf.T.reshape(-1)) < 1e-7
    del spec_op


def test_config3_gradient_vs_finite_differences(dev):
    """2x64x64, N=3: log-normal kappa (sigma 0.5), per-element tau, misfit to
    the solution of a perturbed kappa; 20 centred differences, eps = 1e-6."""
    pm, _ = _meshes(64, 3, 3)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    rng = np.random.default_rng(33)
    kap = np.exp(0.5 * rng.standard_normal(pm.K))
    tau = 200.0 * rng.uniform(0.8, 1.2, pm.K)
    f = 8.0 * np.exp(pm.x) * np.cos(3 * pm.y)
    ud = np.exp(pm.x) * np.cos(3 * pm.y)
    spec = SolveSpec(pm, bc, u_dir=_flat(ud, dev), preconditioner="bj", tol=1e-14,
                     maxiter=30000)
    ft = _flat(f, dev)
    mop = SipgOperator(pm, bc)
    _say("config3 target")
    with torch.no_grad():
        target = sipg_solve(torch.tensor(kap * np.exp(0.3 * rng.standard_normal(pm.K)),
                                         device=dev), torch.tensor(tau, device=dev), ft, spec)

    def loss(k, t):
        e = sipg_solve(k, t, ft, spec) - target
        return 0.5 * torch.dot(e, mass_apply(e, mop))

    kt = torch.tensor(kap, device=dev, requires_grad=True)
    tt = torch.tensor(tau, device=dev, requires_grad=True)
    _say("config3 gradient")
    L = loss(kt, tt)
    L.backward()
    g = torch.cat([kt.grad, tt.grad]).cpu().numpy()
    gn = np.linalg.norm(g)
    assert np.isfinite(g).all() and gn > 0
    eps = 1e-6
    errs = []
    with torch.no_grad():
        for i in range(20):
            d = rng.standard_normal(2 * pm.K)
            d /= np.linalg.norm(d)
            dk = torch.tensor(d[:pm.K], device=dev)
            dt = torch.tensor(d[pm.K:], device=dev)
            k0, t0 = kt.detach(), tt.detach()
            fd = (loss(k0 + eps * dk, t0 + eps * dt) - loss(k0 - eps * dk, t0 - eps * dt)) / (2 * eps)
            errs.append(abs(fd.item() - g @ d) / gn)
    _say(f"config3 fd: L={L.item():.6e} |g|={gn:.3e} max err/|g| = {max(errs):.3e}")
    assert max(errs) < 5e-5
