"""Solver contexts on a side stream replay captured chunks of CG iterations
(HIP graphs).  Their results must equal the eager launches of the default
context bit for bit, on every replay, with eager work and synchronisations
in between (stated tolerance: 0)."""
import numpy as np
import pytest
import torch

from conftest import unit_square
from ddg_amd import _lib
from ddg_amd import mesh as pmesh
from ddg_amd import multigrid as MG
from ddg_amd import sipg as D
from ddg_amd.operator import refine_solve
from ddg_amd.serve import SolverContext

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def problem(dev):
    pm, _ = unit_square(12, 12, 3, jitter=0.1, seed=41)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    op64 = D.get_operator(pm, bc)
    op32 = D.get_operator(pm, bc, dtype=torch.float32)
    mg64 = MG.make_sipg_mg_preconditioner(pm, bc_types=bc)
    mg32 = MG.make_sipg_mg_preconditioner(pm, bc_types=bc, dtype=torch.float32)
    b = torch.as_tensor(np.random.default_rng(42).standard_normal(pm.K * pm.Np), device=dev)
    return pm, op64, op32, mg64, mg32, b


@pytest.mark.parametrize("alpha", [0.0, 7.5])
def test_graph_pcg_equals_eager(problem, alpha):
    pm, op64, op32, mg64, mg32, b = problem
    _say(f"graph pcg alpha={alpha}")
    x_e, info_e = op64.pcg(b, alpha=alpha, pc=mg64, tol=1e-11, maxiter=500, check_every=8)
    ctx = SolverContext(op64, pc=mg64)
    s = torch.cuda.Stream()
    cfg = _lib.CGConfig(1e-11, 0.0, 500, 8)
    for rep in range(3):
        x = torch.zeros_like(b)
        with torch.cuda.stream(s):
            info = ctx.pcg(b, x, alpha, cfg, s.cuda_stream)
        s.synchronize()
        assert info.iterations == info_e["iterations"] and bool(info.converged)
        assert torch.equal(x, x_e), f"replay {rep} differs from the eager solve"
        op64.apply(x)              # eager work between replays
        torch.cuda.synchronize()


def test_graph_refinement_equals_eager(problem):
    pm, op64, op32, mg64, mg32, b = problem
    _say("graph refinement")
    x_e, info_e = refine_solve(op64, op32, b, pc32=mg32, tol=1e-10, inner_tol=1e-4,
                               inner_maxiter=500, check_every=8)
    ctx = SolverContext(op64, op32, mg32)
    s = torch.cuda.Stream()
    cfg = _lib.IRConfig(1e-10, 0.0, 20, 1e-4, 500, 8)
    for alpha in (0.0, 0.0, 3.0, 0.0):   # a coefficient change forces a new capture
        x = torch.zeros_like(b)
        with torch.cuda.stream(s):
            info = ctx.solve(b, x, alpha, cfg, s.cuda_stream)
        s.synchronize()
        if alpha == 0.0:
            assert info.inner_total == info_e["iterations"] and info.outer == info_e["outer"]
            assert torch.equal(x, x_e)
        else:
            x_a, _ = refine_solve(op64, op32, b, alpha=alpha, pc32=mg32, tol=1e-10,
                                  inner_tol=1e-4, inner_maxiter=500, check_every=8)
            assert torch.equal(x, x_a)


def test_graph_mixed_preconditioned_pcg_equals_eager(problem):
    """Float64 CG with the float32 V-cycle: the captured chunk (conversions
    around the cycle included) replays the eager solve bit for bit."""
    pm, op64, op32, mg64, mg32, b = problem
    _say("graph pcg with float32 V-cycle")
    x_e, info_e = op64.pcg(b, pc=mg32, tol=1e-11, maxiter=500, check_every=8)
    assert info_e["converged"]
    ctx = SolverContext(op64, pc=mg32)
    s = torch.cuda.Stream()
    cfg = _lib.CGConfig(1e-11, 0.0, 500, 8)
    for rep in range(3):
        x = torch.zeros_like(b)
        with torch.cuda.stream(s):
            info = ctx.pcg(b, x, 0.0, cfg, s.cuda_stream)
        s.synchronize()
        assert info.iterations == info_e["iterations"] and bool(info.converged)
        assert torch.equal(x, x_e), f"replay {rep} differs from the eager solve"
        op64.apply(x)
        torch.cuda.synchronize()


@pytest.mark.parametrize("case", ["odd_chunk", "block_jacobi", "maxiter_13", "bj_odd_chunk",
                                  "bj_maxiter_13", "plain_maxiter_13"])
def test_graph_chunk_edge_cases(problem, case):
    """An odd check_every (raised to an even chunk under capture), the
    block-Jacobi update fused into the captured chunk, and a maxiter that is
    not a multiple of the chunk (the captured chunk runs past it; the
    kernels after the stop flag return at once): replays equal the eager
    solve bit for bit.  The bj_ and plain_ cases run the float64 iteration whose
    direction update is inside the apply (two launches per iteration and a
    stop test per chunk) under the same edge cases."""
    pm, op64, op32, mg64, mg32, b = problem
    every, tol, maxiter, pc = 8, 1e-11, 500, mg64
    if case == "odd_chunk":
        every = 7
    elif case == "block_jacobi":
        pc = op64.block_jacobi()
        tol, maxiter = 1e-9, 5000
    elif case == "bj_odd_chunk":
        pc, every, tol, maxiter = op64.block_jacobi(), 7, 1e-9, 5000
    elif case == "bj_maxiter_13":
        pc, tol, maxiter = op64.block_jacobi(), 1e-30, 13
    elif case == "plain_maxiter_13":
        pc, tol, maxiter = None, 1e-30, 13
    else:
        tol, maxiter = 1e-30, 13
    _say(f"graph edge case {case}: every={every} maxiter={maxiter}")
    x_e, info_e = op64.pcg(b, pc=pc, tol=tol, maxiter=maxiter, check_every=every)
    ctx = SolverContext(op64, pc=pc)
    s = torch.cuda.Stream()
    cfg = _lib.CGConfig(tol, 0.0, maxiter, every)
    for rep in range(3):
        x = torch.zeros_like(b)
        with torch.cuda.stream(s):
            info = ctx.pcg(b, x, 0.0, cfg, s.cuda_stream)
        s.synchronize()
        assert info.iterations == info_e["iterations"]
        assert bool(info.converged) == info_e["converged"]
        assert torch.equal(x, x_e), f"replay {rep} differs from the eager solve"
        op64.apply(x)
        torch.cuda.synchronize()
    if case.endswith("maxiter_13"):
        assert info_e["iterations"] == 13 and not info_e["converged"]
    else:
        assert info_e["converged"]
