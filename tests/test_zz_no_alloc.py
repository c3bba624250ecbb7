"""ddgx_pcg_solve_noalloc (Krylov state in the library's fixed device array)
against ddgx_pcg_solve (the operator's allocated workspace), through the C
ABI.  Same kernels and reduction order, so the bar is bitwise equality of
the solution and equal CG information, on two calls in a row, the second on
a smaller problem; float64 multigrid, float32 multigrid in a float64 solve,
block-Jacobi and no preconditioner.  A solve larger than the array is
refused before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import unit_square
from ddg_amd import _lib
from ddg_amd import mesh as pmesh
from ddg_amd import multigrid as MG
from ddg_amd import sipg as D

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _solve(entry, op, pc, b, cfg, stream):
    x = torch.zeros_like(b)
    info = _lib.CGInfo()
    rc = getattr(_lib.lib(), entry)(op.handle, 0.0, 0.0, None if pc is None else pc.handle,
                                    C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                    C.byref(cfg), C.byref(info), C.c_void_p(stream))
    _lib.check(rc, entry)
    torch.cuda.synchronize()
    return x, (info.iterations, info.converged, info.residual2, info.atol2)


def _problem(n, N, pc_kind, dev):
    pm, _ = unit_square(n, n, N, jitter=0.1, seed=50 + n + N)
    bc = pmesh.dirichlet_x_neumann_y(pm)
    op = D.get_operator(pm, bc)
    if pc_kind == "mg64":
        pc = MG.make_sipg_mg_preconditioner(pm, bc_types=bc)
    elif pc_kind == "mg32":
        pc = MG.make_sipg_mg_preconditioner(pm, bc_types=bc, dtype=torch.float32)
    elif pc_kind == "bj":
        pc = op.block_jacobi()
    else:
        pc = None
    b = torch.as_tensor(np.random.default_rng(n + N).standard_normal(pm.K * pm.Np), device=dev)
    return pm, op, pc, b


@pytest.mark.parametrize("pc_kind", ["mg64", "mg32", "bj", None])
def test_noalloc_equals_allocating_solve(dev, pc_kind):
    s = torch.cuda.current_stream().cuda_stream
    sizes = {"mg64": ((32, 4), (16, 3)), "mg32": ((32, 4), (16, 3)), "bj": ((16, 3), (8, 2)),
             None: ((8, 3), (6, 2))}[pc_kind]
    for n, N in sizes:                         # second call: a smaller problem
        pm, op, pc, b = _problem(n, N, pc_kind, dev)
        maxiter = 20000 if pc_kind in ("bj", None) else 500
        cfg = _lib.CGConfig(1e-10, 0.0, maxiter, 16)
        _say(f"noalloc pcg pc={pc_kind} 2x{n}x{n} N={N} K*Np={pm.K * pm.Np}")
        x_na, i_na = _solve("ddgx_pcg_solve_noalloc", op, pc, b, cfg, s)
        x_ws, i_ws = _solve("ddgx_pcg_solve", op, pc, b, cfg, s)
        assert i_na == i_ws, (i_na, i_ws)
        assert i_na[1] == 1, i_na
        assert torch.equal(x_na, x_ws)


def test_noalloc_rejects_a_solve_larger_than_its_array(dev):
    VX, VY, EToV = pmesh.forest_rectangle_hierarchy(2, 7)[-1]
    m = pmesh.build_mesh_2d(VX, VY, EToV, 6)                   # 3.67 M float64 unknowns
    op = D.get_operator(m, pmesh.dirichlet_x_neumann_y(m))
    b = torch.ones(m.K * m.Np, dtype=torch.float64, device=dev)
    x = torch.full_like(b, 7.0)
    cfg = _lib.CGConfig(1e-10, 0.0, 10, 16)
    info = _lib.CGInfo()
    _say(f"noalloc oversize K*Np={m.K * m.Np}")
    rc = _lib.lib().ddgx_pcg_solve_noalloc(op.handle, 0.0, 0.0, None, C.c_void_p(b.data_ptr()),
                                           C.c_void_p(x.data_ptr()), C.byref(cfg),
                                           C.byref(info), C.c_void_p(None))
    assert rc == 1
    assert b"fixed array" in _lib.lib().ddgx_last_error()
    assert torch.all(x == 7.0)
