"""Pins for the adjoint oracle: the unit-perturbation derivatives of the
form are exact (the form is affine in each parameter), and the adjoint
identity dL/dtheta = -d a(u, lam; u_dir)/d theta matches centred finite
differences of the LU-solved misfit loss, for kappa and tau, with Dirichlet
data and Neumann data present.  The penalty scale is ~10 here: at the
default 200 the loss barely depends on tau (relative sensitivity ~1e-7) and
a finite difference of it is round-off."""

import numpy as np
import pytest

from oracle import adjoint as OA
from oracle import mesh as om
from oracle import sipg as S


def _problem(N=2, n=3, seed=0):
    from ddg_amd import mesh as pm
    VX, VY, EToV = pm.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, n, n)
    VX, VY = pm.jitter_vertices(VX, VY, 1.0 / n, 0.1, seed)
    m = om.build_mesh_2d(VX, VY, EToV, N)
    bc = pm.dirichlet_x_neumann_y(m)
    rng = np.random.default_rng(seed)
    kap = np.exp(0.5 * rng.standard_normal(m.K))
    tau = rng.uniform(8.0, 12.0, m.K)
    f = np.sin(2 * m.x + 1) * np.cos(m.y)
    ud = np.exp(m.x) * np.cos(3 * m.y)
    xf = m.x.reshape(-1, order="F")[m.vmapM]
    q = (0.3 + xf).reshape((3 * m.Nfp, m.K), order="F")
    return m, bc, kap, tau, f, ud, q


def test_form_derivative_is_exact_for_affine_dependence():
    m, bc, kap, tau, f, ud, q = _problem()
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal((m.Np, m.K)), rng.standard_normal((m.Np, m.K))
    gk, gt = OA.form_param_grad(m, u, v, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud)
    a = lambda kp, tp: S.sipg_bilinear_form(m, u, v, bc_types=bc, u_dir=ud,
                                            tau_scale=tp[None, :], kappa=kp)
    for k in (0, 5, m.K - 1):
        e = np.zeros(m.K)
        e[k] = 1.0
        # half steps give half the change: the dependence is affine
        assert np.isclose(a(kap + 0.5 * e, tau) - a(kap, tau), 0.5 * gk[k], rtol=1e-9, atol=1e-12)
        assert np.isclose(a(kap, tau - 2.0 * e) - a(kap, tau), -2.0 * gt[k], rtol=1e-9, atol=1e-12)
    # the symmetric form: swapping u and v leaves the derivatives unchanged
    # without Dirichlet data
    g1 = OA.form_param_grad(m, u, v, bc_types=bc, kappa=kap, tau_scale=tau)
    g2 = OA.form_param_grad(m, v, u, bc_types=bc, kappa=kap, tau_scale=tau)
    np.testing.assert_allclose(g1[0], g2[0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g1[1], g2[1], rtol=1e-9, atol=1e-12)


def test_adjoint_identity_matches_finite_differences():
    m, bc, kap, tau, f, ud, q = _problem()
    rng = np.random.default_rng(2)
    kap_t = kap * np.exp(0.2 * rng.standard_normal(m.K))
    target, _ = OA.solve_dense(m, f, bc_types=bc, kappa=kap_t, tau_scale=tau, u_dir=ud, q_neu=q)
    kw = dict(bc_types=bc, u_dir=ud, q_neu=q)
    L0, u = OA.misfit_loss(m, f, target, kappa=kap, tau_scale=tau, **kw)
    _, A = OA.solve_dense(m, f, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud, q_neu=q)
    dLdu = S.mass_apply(m, u - target).T.reshape(-1)
    lam = np.linalg.solve(A, dLdu).reshape(m.K, m.Np).T
    gk, gt = OA.form_param_grad(m, u, lam, bc_types=bc, kappa=kap, tau_scale=tau, u_dir=ud)
    gk, gt = -gk, -gt
    # step sizes where the centred difference is accurate: its error falls as
    # eps^2 down to ~1e-7 relative, below that LU round-off / eps dominates
    for which, eps, bar in (("kappa", 1e-4, 1e-6), ("tau", 1e-3, 1e-5), ("kappa", 1e-4, 1e-6),
                            ("tau", 1e-3, 1e-5)):
        d = rng.standard_normal(m.K)
        d /= np.linalg.norm(d)
        dk, dt = (d, 0 * d) if which == "kappa" else (0 * d, d)
        Lp, _ = OA.misfit_loss(m, f, target, kappa=kap + eps * dk, tau_scale=tau + eps * dt, **kw)
        Lm, _ = OA.misfit_loss(m, f, target, kappa=kap - eps * dk, tau_scale=tau - eps * dt, **kw)
        fd = (Lp - Lm) / (2 * eps)
        ad = gk @ dk + gt @ dt
        assert abs(fd - ad) <= bar * abs(ad), (which, fd, ad)
    # the derivative is not trivially zero in either parameter
    assert np.abs(gk).max() > 1e-6 * L0 and np.abs(gt).max() > 1e-9 * L0
