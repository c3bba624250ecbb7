"""Full-size solves against the oracle's float64 solutions
(tests/golden/config_solutions.npz, written by
tests/golden/make_solution_fixtures.py on the CPU).

The workloads are the benchmark's: 2x256x256 N=4 (e^x cos 3y with Dirichlet x
/ Neumann y walls, and a 64-mode Fourier right-hand side) and 2x128x128 N=6
(Fourier).  Each is solved to a relative residual of 1e-13 by float64 hp-MG
PCG, float64 CG with the float32 V-cycle, and mixed-precision refinement.
Bar: relative L2 error over the fixture's 4096 seeded nodes below 1e-10 and
every seeded projection within 1e-10 relative."""
import numpy as np
import pytest
import torch

import bench as B
from ddg_amd import sipg as D
from ddg_amd.operator import refine_solve

pytestmark = pytest.mark.gpu

CASES = {"config1": (256, 4, "exact"), "config2_n4": (256, 4, 1), "config2_n6": (128, 6, 2)}


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", list(CASES))
def test_solves_match_oracle_fixture(dev, name):
    n, N, kind = CASES[name]
    _say(f"fixture {name}: building 2x{n}x{n} N={N}")
    m, bc = B.workload_mesh(n, N=N)
    mg64, fin = B.hpmg_for(m, N, torch.float64)
    mg32, _ = B.hpmg_for(fin, N, torch.float32)
    op64 = D.get_operator(fin, bc)
    op32 = D.get_operator(fin, bc, dtype=torch.float32)
    if kind == "exact":
        f, ud, q = B.problem_data(fin)
        b = op64.rhs(B._flat(f, dev), u_dir=B._flat(ud, dev), q_neu=B._flat(q, dev))
    else:
        b = op64.rhs(B._flat(B.fourier_rhs(fin, seed=kind), dev))
    solvers = {
        "fp64": lambda: op64.pcg(b, pc=mg64, tol=1e-13, maxiter=5000),
        "pcg32": lambda: op64.pcg(b, pc=mg32, tol=1e-13, maxiter=5000),
        "refine": lambda: refine_solve(op64, op32, b, pc32=mg32, tol=1e-13, inner_tol=1e-4,
                                       inner_maxiter=2000),
    }
    for label, fn in solvers.items():
        _say(f"fixture {name}: {label}")
        x, info = fn()
        assert info["converged"] or info.get("at_floor"), (label, info)
        e = B.fixture_error(x, name)
        assert e is not None
        assert e["rel_l2_sampled"] < 1e-10, (label, e)
        assert e["rel_proj"] < 1e-10, (label, e)
