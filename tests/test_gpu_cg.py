"""Device block-Jacobi and CG against the oracle (solution bar: 1e-8 relative
at tol 1e-12; iteration counts must agree within 1% or 2 -- exact counts
depend on the reduction order over ~10^3 iterations and are not pinned)."""
import numpy as np
import pytest
import torch

from conftest import unit_square, rel
from oracle import sipg as S
from ddg_amd import mesh as pmesh
from ddg_amd import sipg as D
from ddg_amd.operator import SipgOperator

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _flat(u, dev):
    return torch.as_tensor(np.ascontiguousarray(u.T).reshape(-1), device=dev)


@pytest.mark.parametrize("N", [2, 4, 6])
def test_block_jacobi_blocks(dev, N):
    pm, om = unit_square(4, 3, N, jitter=0.1, seed=1)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    _say(f"bj blocks N={N}")
    op = SipgOperator(pm, bc)
    for alpha in (0.0, 37.0):
        pc = op.block_jacobi(alpha)
        ref = S.extract_block_diag_matfree(S.flat_apply(om, bc, alpha=alpha), om)
        assert rel(pc.blocks().cpu().numpy(), ref) < 1e-12
        r = np.random.default_rng(0).standard_normal(pm.K * pm.Np)
        z = pc(torch.as_tensor(r, device=dev)).cpu().numpy()
        zref = np.einsum("kij,kj->ki", np.linalg.inv(ref), r.reshape(pm.K, pm.Np)).reshape(-1)
        assert rel(z, zref) < 1e-10


@pytest.mark.parametrize("use_pc", [False, True])
def test_pcg_matches_oracle_cg(dev, use_pc):
    pm, om = unit_square(8, 8, 3, jitter=0.1, seed=2)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    f = np.random.default_rng(3).standard_normal((pm.Np, pm.K))
    cfg = S.IterativeConfig(tol=1e-12, maxiter=5000)
    pco = S.make_block_jacobi_preconditioner(om, bc_types=bc) if use_pc else None
    uo, info_o = S.solve_sipg_poisson_2d(om, f, bc_types=bc, iterative=cfg, preconditioner=pco,
                                         return_info=True)
    _say(f"pcg use_pc={use_pc} K={pm.K}")
    pcd = D.make_sipg_block_jacobi_preconditioner(pm, bc_types=bc) if use_pc else None
    ud, info_d = D.solve_sipg_poisson_2d(pm, f, bc_types=bc,
                                         iterative=D.IterativeConfig(tol=1e-12, maxiter=5000),
                                         preconditioner=pcd, return_info=True)
    assert info_d["converged"]
    assert abs(info_d["iterations"] - info_o["iterations"]) <= max(2, 0.01 * info_o["iterations"])
    assert rel(ud, uo) < 1e-8


def test_helmholtz_cg_and_manufactured(dev):
    pm, om = unit_square(8, 8, 3)
    bc = pmesh.all_dirichlet(pm)
    ue = np.sin(np.pi * om.x) * np.sin(np.pi * om.y)
    alpha = 100.0
    _say("helmholtz cg")
    pc = D.make_helmholtz_block_jacobi_preconditioner(pm, alpha=alpha, nu=1.0, bc_types=bc)
    u = D.solve_sipg_helmholtz_2d(pm, (2 * np.pi ** 2 + alpha) * ue, alpha=alpha, nu=1.0,
                                  bc_types=bc, preconditioner=pc,
                                  iterative=D.IterativeConfig(tol=1e-12, maxiter=2000))
    ulu = S.solve_sipg_helmholtz_2d(om, (2 * np.pi ** 2 + alpha) * ue, alpha=alpha, nu=1.0,
                                    bc_types=bc)
    assert np.abs(u - ulu).max() < 1e-8
    err = np.sqrt(np.sum(om.J * (u - ue) ** 2) / np.sum(om.J * ue ** 2))
    assert err < 1e-3


def test_dirichlet_and_neumann_data(dev):
    pm, om = unit_square(8, 8, 3)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    ue = np.exp(om.x) * np.cos(3 * om.y)
    nx, ny = om.nx, om.ny
    xf, yf = om.x.reshape(-1, order="F")[om.vmapM], om.y.reshape(-1, order="F")[om.vmapM]
    gx, gy = np.exp(xf) * np.cos(3 * yf), -3 * np.exp(xf) * np.sin(3 * yf)
    q = (nx.reshape(-1, order="F") * gx + ny.reshape(-1, order="F") * gy).reshape(
        (3 * om.Nfp, om.K), order="F")
    _say("dirichlet+neumann data solve")
    cfg = D.IterativeConfig(tol=1e-12, maxiter=3000)
    u = D.solve_sipg_poisson_2d(pm, 8 * ue, bc_types=bc, u_dir=ue, q_neu=q, iterative=cfg,
                                preconditioner=D.make_sipg_block_jacobi_preconditioner(pm, bc_types=bc))
    uo = S.solve_sipg_poisson_2d(om, 8 * ue, bc_types=bc, u_dir=ue, q_neu=q)
    assert rel(u, uo) < 1e-8


def test_singular_neumann(dev):
    pm, om = unit_square(6, 6, 2)
    bc = np.where(pmesh.boundary_faces(pm), 2, 0).astype(np.int32)
    f = np.cos(np.pi * om.x) * np.cos(np.pi * om.y)
    cfg = S.IterativeConfig(tol=1e-10, maxiter=4000)
    _say("singular neumann")
    u = D.solve_sipg_poisson_2d(pm, f, bc_types=bc, iterative=D.IterativeConfig(tol=1e-10, maxiter=4000))
    uo = S.solve_sipg_poisson_2d(om, f, bc_types=bc, iterative=cfg)
    assert rel(u, uo) < 1e-6


def test_maxiter_not_converged_is_reported(dev):
    pm, _ = unit_square(8, 8, 3)
    bc = pmesh.all_dirichlet(pm)
    _say("maxiter=5")
    u, info = D.solve_sipg_poisson_2d(pm, np.ones((pm.Np, pm.K)), bc_types=bc,
                                      iterative=D.IterativeConfig(tol=1e-12, maxiter=5),
                                      return_info=True)
    assert info["iterations"] == 5 and not info["converged"]
    assert np.all(np.isfinite(u))


def test_bad_method_raises(dev):
    pm, _ = unit_square(2, 2, 1)
    with pytest.raises(ValueError):
        D.solve_sipg_poisson_2d(pm, np.ones((pm.Np, pm.K)), bc_types=pmesh.all_dirichlet(pm),
                                iterative=D.IterativeConfig(method="gmres"))


def test_direct_lu_path(dev):
    pm, om = unit_square(4, 4, 2)
    bc = pmesh.all_dirichlet(pm)
    f = np.random.default_rng(5).standard_normal((pm.Np, pm.K))
    _say("direct lu")
    u = D.solve_sipg_poisson_2d(pm, f, bc_types=bc)
    assert rel(u, S.solve_sipg_poisson_2d(om, f, bc_types=bc)) < 1e-10


@pytest.mark.parametrize("where", ["default", "context"])
def test_concurrent_calls_on_busy_state_are_refused(dev, where):
    """A solve holds its Krylov state until it returns: on the operator's
    default context (ddgx_pcg_solve) a dot or a solve on the same handle from
    another thread meanwhile, on a solver context a second solve on it, is
    refused (DDGX_ERR_BUSY) instead of racing on that state, and runs after."""
    import threading
    import time
    from ddg_amd import _lib
    from ddg_amd._lib import DdgxBusyError
    from ddg_amd.serve import SolverContext
    pm, _ = unit_square(96, 96, 4)
    op = SipgOperator(pm, pmesh.dirichlet_x_neumann_y(pm))
    b = torch.ones(pm.K * pm.Np, dtype=torch.float64, device=dev)
    started, refused, ran = threading.Event(), [], []
    # tol 1e-30 is never met: a solve runs all maxiter iterations (> 0.2 s)
    long_cfg = _lib.CGConfig(1e-30, 0.0, 20000, 64)
    if where == "default":
        probes = (("dot", lambda: op.dot(b, b)), ("pcg", lambda: op.pcg(b, maxiter=1)))

        def long_solve():
            _, info = op.pcg(b, tol=1e-30, maxiter=20000, check_every=64)
            return info["iterations"]
    else:
        ctx = SolverContext(op)
        side = torch.cuda.Stream()
        x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
        probes = (("ctx.pcg", lambda: ctx.pcg(b, x2, 0.0, _lib.CGConfig(1e-9, 0.0, 1, 16),
                                              side.cuda_stream)),)

        def long_solve():
            return ctx.pcg(b, x1, 0.0, long_cfg, side.cuda_stream).iterations

    def probe():
        started.wait()
        time.sleep(0.05)
        for name, fn in probes:
            try:
                fn()
                ran.append(name)
            except DdgxBusyError as e:
                refused.append(str(e))

    torch.cuda.synchronize()  # b and the solution vectors exist before the side stream runs
    th = threading.Thread(target=probe)
    th.start()
    _say(f"busy guard {where} K={pm.K}")
    t0 = time.perf_counter()
    started.set()
    iterations = long_solve()
    dt = time.perf_counter() - t0
    th.join()
    torch.cuda.synchronize()
    _say(f"solve {dt:.3f} s, {iterations} iterations, refused {len(refused)}, ran {ran}")
    assert dt > 0.2 and iterations == 20000
    assert ran == [] and len(refused) == len(probes)
    assert all("in use by a concurrent call" in m for m in refused)
    for _, fn in probes:
        fn()
    assert op.dot(b, b) == float(pm.K * pm.Np)
