"""Batched solves (ddgx_batch through ddg_amd.operator.BatchSolver) against
the single solve (ddgx_pcg_solve through SipgOperator.pcg).

A batch advances several problems of one class (mesh, order, multigrid)
with one launch per step; each request keeps its own Helmholtz coefficient,
right-hand side, scalars and stopping test.  Bar: every request's solution
and CG information are bitwise those of its solve alone (stated tolerance:
0), for float64 CG with a float32 or float64 hp-multigrid and without a
preconditioner, with requests that stop at different iterations, and
whatever the batch's other members and their order.  The single solve it is
compared with is itself checked against the oracle in test_gpu_mg.py and
test_gpu_serve.py."""
import ctypes

import numpy as np
import pytest
import torch

from ddg_amd import mesh as pm
from ddg_amd import multigrid as MG
from ddg_amd import sipg as D
from ddg_amd import _lib
from ddg_amd.operator import BatchSolver

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _class(n, N, mg_dtype):
    """The request loop's class layout: a 2x8x8 forest base refined to 2 n^2,
    homogeneous Dirichlet walls, hp-multigrid N, N/2 (or N-1 ...), 1."""
    lb = 2
    lv = int(round(np.log2(n))) - 1
    VX, VY, E = pm.forest_rectangle_hierarchy(2, lb)[-1]
    base = pm.build_mesh_2d(VX, VY, E, N)
    degs = [N, N // 2, 1] if N >= 4 else list(range(N, 0, -1))
    mg, mesh = MG.make_sipg_hpmg_preconditioner(base, lv - lb, bc_fn=pm.all_dirichlet,
                                                degrees=degs, dtype=mg_dtype)
    op = D.get_operator(mesh, pm.all_dirichlet(mesh))
    return op, mg, mesh


def _rhs(op, mesh, seed, dev):
    rng = np.random.default_rng(seed)
    if seed % 3 == 0:   # smooth: converges sooner than the rough ones
        f = np.sin(np.pi * mesh.x) * np.sin(2 * np.pi * mesh.y)
    else:
        f = rng.standard_normal((mesh.Np, mesh.K))
    return op.rhs(torch.as_tensor(np.ascontiguousarray(f.T).reshape(-1), device=dev))


CASES = [(16, 2, "f32"), (32, 4, "f32"), (16, 5, "f32"), (16, 3, "f64"), (16, 4, "none")]


@pytest.mark.parametrize("n,N,kind", CASES)
def test_batch_equals_single_solves(dev, n, N, kind):
    op, mg, mesh = _class(n, N, torch.float64 if kind == "f64" else torch.float32)
    pc = None if kind == "none" else mg
    tol, maxiter = (1e-10, 4000) if pc is not None else (1e-8, 20000)
    n1 = op.K * op.Np
    alphas = [0.0, 12.5, 37.0, 3.0]
    bs = [_rhs(op, mesh, s, dev) for s in range(len(alphas))]
    _say(f"batch n={n} N={N} pc={kind}: single solves")
    single = [op.pcg(b, alpha=a, pc=pc, tol=tol, maxiter=maxiter, check_every=8)
              for a, b in zip(alphas, bs)]
    for x, info in single:
        assert info["converged"], info
    _say("iterations alone:", [info["iterations"] for _, info in single])
    for order in ([0, 1, 2, 3], [3, 1], [2]):
        _say(f"batch n={n} N={N} pc={kind}: batch {order}")
        bt = BatchSolver(op, pc, len(order))
        b = torch.cat([bs[i] for i in order])
        x, infos = bt.pcg(b, [alphas[i] for i in order], tol=tol, maxiter=maxiter,
                          check_every=8)
        torch.cuda.synchronize()
        for j, i in enumerate(order):
            xs, info = single[i]
            assert torch.equal(x[j * n1:(j + 1) * n1], xs), (order, i)
            assert infos[j] == info, (order, i, infos[j], info)
        bt.close()


def test_batch_graph_replay_equals_eager(dev):
    """On a non-null stream the batch replays captured chunks of iterations;
    results equal the eager launches of the null stream, on two solves in a
    row with different coefficients (the chunk is replayed, not recaptured)."""
    op, mg, mesh = _class(32, 3, torch.float32)
    bs = torch.cat([_rhs(op, mesh, s, dev) for s in (1, 2)])
    bt = BatchSolver(op, mg, 2)
    s = torch.cuda.Stream()
    for alphas in ([4.0, 20.0], [9.0, 1.0]):
        _say(f"batch graph replay alphas={alphas}")
        xe, ie = bt.pcg(bs, alphas, tol=1e-10, maxiter=4000, check_every=7)
        xe = xe.clone()
        with torch.cuda.stream(s):
            xg, ig = bt.pcg(bs, alphas, tol=1e-10, maxiter=4000, check_every=7,
                            stream=s.cuda_stream)
        s.synchronize()
        assert torch.equal(xg, xe) and ig == ie
    bt.close()


def test_batch_refuses_what_it_cannot_run(dev):
    op, mg, mesh = _class(16, 2, torch.float32)
    with pytest.raises(ValueError):
        BatchSolver(op, op.block_jacobi(), 2)          # block-Jacobi: not batched
    op32 = D.get_operator(mesh, pm.all_dirichlet(mesh), dtype=torch.float32)
    with pytest.raises(ValueError):
        BatchSolver(op32, mg, 2)                       # float32 operator
    h = ctypes.c_void_p()
    rc = _lib.lib().ddgx_batch_create(op.handle, op.block_jacobi().handle, 2, None,
                                      ctypes.byref(h))
    assert rc != 0 and b"multigrid" in _lib.lib().ddgx_last_error()
    # a mesh whose element count is not a multiple of the 64-element tile
    m = pm.build_mesh_2d(*pm.uniform_rectangle_mesh_2d(0, 1, 0, 1, 5, 5), 2)
    opx = D.get_operator(m, pm.all_dirichlet(m))
    with pytest.raises(ValueError, match="multiple of 64"):
        BatchSolver(opx, None, 2)
    bt = BatchSolver(op, mg, 2)
    with pytest.raises(ValueError):
        bt.pcg(torch.zeros(op.K * op.Np, dtype=torch.float64, device=dev), [1.0, 2.0])
    bt.close()
