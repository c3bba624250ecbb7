"""Request loop on the GPU, with both solve methods (iterative refinement and
float64 CG with the float32 V-cycle, the latter in batches of one class):
every result is bitwise the same alone and in a mix of concurrent, batched
requests (stated tolerance: 0), served results meet the float64 tolerance
against the oracle's dense solve, and admission keeps the slot and HBM
limits."""
import numpy as np
import pytest
import torch

from conftest import rel
from oracle import mesh as om
from oracle import sipg as S
from ddg_amd.serve import (Request, ServeConfig, SolveServer, make_requests,
                           poisson_arrivals, fourier_field)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def server(dev):
    srv = SolveServer(ServeConfig(method="refine", slots=4, keep_solutions=True, tol=1e-10))
    yield srv
    srv.close()


def test_alone_equals_mix(server):
    reqs = make_requests(24, seed=5, sizes=(8, 16, 32), orders=(2, 3, 4, 5), k2_range=(0, 50))
    _say("serve: classes", sorted({(r.n, r.N) for r in reqs}))
    server.prepare(reqs)
    alone = {}
    for r in reqs:
        x, info = server.solve_one(r, server.streams[0])
        torch.cuda.synchronize()
        assert info.converged or info.at_floor, (r, info.residual2, info.atol2)
        alone[r.rid] = x.clone()
    _say("serve: mix")
    poisson_arrivals(reqs, None)
    rep = server.run(reqs)
    assert rep.peak_slots > 1, "requests never overlapped"
    assert len(rep.results) == len(reqs)
    for res in rep.results:
        assert res.converged or res.at_floor
        assert torch.equal(res.x, alone[res.rid]), f"request {res.rid} differs in the mix"


def _oracle_solution(srv, r):
    pc = srv.problem_class(r.n, r.N)
    m = pc.mesh
    mo = om.build_mesh_2d(m.VX, m.VY, m.EToV, m.N)
    f = fourier_field(pc.x, pc.y, r.seed).cpu().numpy().reshape(m.K, m.Np).T
    b = S.mass_apply(mo, f).T.reshape(-1)
    return np.linalg.solve(S.sipg_matrix_2d(mo, bc_types=pc.bc, alpha=r.k2), b)


def test_results_match_oracle(server):
    rng = np.random.default_rng(9)
    reqs = [Request(rid=i, n=8, N=N, k2=float(rng.uniform(0, 50)), seed=int(rng.integers(2 ** 31)))
            for i, N in enumerate((2, 3, 4, 5, 6))]
    for r in reqs:
        pc = server.problem_class(r.n, r.N)
        m = pc.mesh
        mo = om.build_mesh_2d(m.VX, m.VY, m.EToV, m.N)
        _say(f"serve oracle n={r.n} N={r.N} k2={r.k2:.2f}")
        x, info = server.solve_one(r, server.streams[1])
        torch.cuda.synchronize()
        f = fourier_field(pc.x, pc.y, r.seed).cpu().numpy().reshape(m.K, m.Np).T
        b = S.mass_apply(mo, f).T.reshape(-1)
        A = S.sipg_matrix_2d(mo, bc_types=pc.bc, alpha=r.k2)
        xo = np.linalg.solve(A, b)
        err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
        assert (info.converged or info.at_floor) and err < 1e-9, (err, info.outer)


def test_poisson_stream_within_limits(server):
    reqs = make_requests(40, seed=11, sizes=(8, 16), orders=(2, 3), k2_range=(0, 50))
    poisson_arrivals(reqs, rate=400.0, seed=12)
    server.prepare(reqs)
    fp = max(server.problem_class(r.n, r.N).footprint for r in reqs)
    server.cfg.keep_solutions = False
    budget, server.cfg.hbm_budget = server.cfg.hbm_budget, 2 * fp
    try:
        rep = server.run(reqs)
    finally:
        server.cfg.hbm_budget = budget
        server.cfg.keep_solutions = True
    st = rep.stats()
    _say("serve stream:", st)
    assert st["requests"] == 40 and st["all_solved"]
    assert rep.peak_bytes <= 2 * fp and rep.peak_slots <= 4
    for res in rep.results:
        assert res.start >= res.arrival and res.end >= res.start


def test_pcg32_method_alone_equals_mix_and_matches_oracle(dev):
    """The "pcg32" method (float64 CG with the float32 V-cycle): bitwise the
    same alone and in a mix, and within 1e-9 of the oracle's dense solve."""
    srv = SolveServer(ServeConfig(method="pcg32", slots=4, keep_solutions=True, tol=1e-10))
    try:
        reqs = make_requests(12, seed=15, sizes=(8, 16), orders=(2, 3, 4), k2_range=(0, 50))
        srv.prepare(reqs)
        alone = {}
        for r in reqs:
            x, info = srv.solve_one(r, srv.streams[0])
            torch.cuda.synchronize()
            assert info.converged, (r, info.residual2, info.atol2)
            alone[r.rid] = x.clone()
        _say("serve pcg32: mix")
        poisson_arrivals(reqs, None)
        rep = srv.run(reqs)
        assert len(rep.results) == len(reqs) and rep.peak_slots > 1
        # requests of one class that arrived together ran as batches
        assert rep.stats()["max_batch"] > 1, rep.batch_sizes
        for res in rep.results:
            assert res.converged
            assert torch.equal(res.x, alone[res.rid]), f"request {res.rid} differs in the mix"
        r = reqs[0]
        xo = _oracle_solution(srv, r)
        err = np.linalg.norm(alone[r.rid].cpu().numpy() - xo) / np.linalg.norm(xo)
        assert err < 1e-9, err
        # orders 5 and 6 (two matrix-core row blocks, float32 V-cycle) at n=8
        rng = np.random.default_rng(16)
        for i, N in enumerate((5, 6)):
            r = Request(rid=100 + i, n=8, N=N, k2=float(rng.uniform(0, 50)),
                        seed=int(rng.integers(2 ** 31)))
            _say(f"serve pcg32 oracle n=8 N={N} k2={r.k2:.2f}")
            srv.prepare([r])
            x, info = srv.solve_one(r, srv.streams[1])
            torch.cuda.synchronize()
            xo = _oracle_solution(srv, r)
            err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
            assert info.converged and err < 1e-9, (N, err)
    finally:
        srv.close()
