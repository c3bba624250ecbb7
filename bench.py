"""Benchmark: SIPG matvec GDOF/s (% of HBM roofline) and CG solves/s with
p50/p99 latency, on one MI355X.

Workload (one GPU): the 2x256x256 triangle mesh of [0,1]^2 with interior
vertices jittered by a seeded U(-0.1h, 0.1h), polynomial order N=4
(Np=15, K=131072, 1,966,080 DOFs), Dirichlet on x = 0, 1 and Neumann on
y = 0, 1, penalty tau_scale = 200, float64.  A "step" is one SIPG apply over
the whole mesh with its input already resident in HBM.

Also reported, on the same mesh:
* float64 CG solves with the hp-multigrid preconditioner (p-levels 4, 2, 1
  on the mesh, then h-levels at degree 1 down to 2x16x16) of
  -Delta u = 8 e^x cos 3y, u = e^x cos 3y on the Dirichlet walls, dn u on
  the Neumann walls, to tol 1e-10: solves/s, p50/p99 latency;
* the same solves by mixed-precision iterative refinement (float32
  operator, float32 V-cycle, float64 residual and update), and by float64
  CG preconditioned with the float32 V-cycle (mixed-precision
  preconditioning: every float64 quantity of the CG stays float64);
* the first of those solves by float64 CG without a preconditioner (the
  plain CG) and with block-Jacobi, once each;
* both solvers on seeded random Fourier right-hand sides (64 modes,
  amplitude |k|^-2) on this mesh and on 2x128x128 at N=6, with the error
  against a float64 solve to tol 1e-13;
* the reference's CPU-runnable case (2x16x16, N=3, block-Jacobi CG to
  1e-9) through the drop-in, beside the oracle's same solve on the host;
* the CPU baseline: the float64 numpy restatement of the reference
  algorithm (oracle/) timed on the host cores.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--solves 20]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def workload_mesh(n=256, N=4, seed=0, order="forest"):
    """2 n^2 triangles of [0,1]^2, interior vertices jittered by U(-0.1h, 0.1h).
    order="forest": the structured mesh generated by uniform red refinement of
    a 2x2 cell mesh (quadtree element order, element 4k+c = child c of k);
    order="rowmajor": the reference's structured builder order."""
    from ddg_amd import mesh as pm
    if order == "forest":
        lv = int(round(np.log2(n))) - 1
        if 2 ** (lv + 1) != n:
            raise ValueError("forest order needs n a power of two")
        VX, VY, EToV = pm.forest_rectangle_hierarchy(2, lv)[-1]
    else:
        VX, VY, EToV = pm.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, n, n)
    VX, VY = pm.jitter_vertices(VX, VY, 1.0 / n, 0.1, seed)
    m = pm.build_mesh_2d(VX, VY, EToV, N)
    return m, pm.dirichlet_x_neumann_y(m)


def problem_data(m):
    """f, u_dir (Np,K) and q_neu (3Nfp,K) for u = e^x cos 3y."""
    ue = np.exp(m.x) * np.cos(3 * m.y)
    xf = m.x.reshape(-1, order="F")[m.vmapM]
    yf = m.y.reshape(-1, order="F")[m.vmapM]
    gx, gy = np.exp(xf) * np.cos(3 * yf), -3 * np.exp(xf) * np.sin(3 * yf)
    q = (m.nx.reshape(-1, order="F") * gx + m.ny.reshape(-1, order="F") * gy)
    return 8.0 * ue, ue, q.reshape((3 * m.Nfp, m.K), order="F")


def fourier_rhs(m, seed=0, modes=64):
    """Random Fourier series at the nodes: sum_m a_m |k_m|^-2 cos(2 pi k_m.x + phi_m),
    a ~ N(0,1), phi ~ U(0, 2 pi), integer wave vectors k_m != 0 with |k_i| <= 8."""
    rng = np.random.default_rng(seed)
    f = np.zeros_like(m.x)
    for _ in range(modes):
        k = np.zeros(2)
        while not k.any():
            k = rng.integers(-8, 9, size=2).astype(np.float64)
        a, phi = rng.standard_normal(), rng.uniform(0, 2 * np.pi)
        f += a / (k @ k) * np.cos(2 * np.pi * (k[0] * m.x + k[1] * m.y) + phi)
    return f


def hpmg_for(m, N, dtype, n_base=16, degrees=None):
    """hp-multigrid of the forest mesh m (2 n^2 elements): p-levels on m,
    then h-levels at degree 1 down to the 2 n_base^2 base mesh."""
    import torch
    from ddg_amd import mesh as pm
    from ddg_amd import multigrid as MG
    n = int(round(np.sqrt(m.K / 2)))
    lb = int(round(np.log2(n_base))) - 1
    lv = int(round(np.log2(n))) - 1
    VX, VY, EToV = pm.forest_rectangle_hierarchy(2, lb)[-1]
    base = pm.build_mesh_2d(VX, VY, EToV, N)
    if degrees is None:
        degrees = [N, N // 2, 1] if N >= 4 else list(range(N, 0, -1))
    mg, fin = MG.make_sipg_hpmg_preconditioner(base, lv - lb, bc_fn=pm.dirichlet_x_neumann_y,
                                               degrees=degrees, p_coarse=degrees[-1],
                                               fine_vertices=(m.VX, m.VY), dtype=dtype)
    assert np.array_equal(np.asarray(fin.EToV), np.asarray(m.EToV))
    return mg, fin


def apply_bytes(K, Np, s):
    """Algorithmic bytes per apply: read u, write y, 5 element + 9 face
    geometry words and 3 int32 face codes per element."""
    return K * (s * (2 * Np + 14) + 12)


def curved_apply_bytes(K, Np, Nfp, cub, s):
    """Algorithmic bytes of one curved apply: read u, write y, the per-node
    metrics and Jacobian (5 Np), the face-node normals (6 Nfp), the per-side
    weight, penalty and sJ (9), with cubature the per-side face masses
    (3 Nfp^2) and the element's symmetric volume stiffness (Np(Np+1)/2), and
    3 int32 face codes."""
    words = 2 * Np + 5 * Np + 6 * Nfp + 9 + (3 * Nfp * Nfp + Np * (Np + 1) // 2 if cub else 0)
    return K * (s * words + 12)


def curved_block(n, N, steps, warmup):
    """The curved apply on the reference's annulus ([-2, 2]^2 without
    [-0.5, 0.5]^2, inner boundary bent onto r = 0.5; test_curved_boundary_2d.py
    43-59) at 2 (n^2 - (n/4)^2) elements, order N, float64: nodal quadrature
    and cubature."""
    import torch
    from ddg_amd import cubature as CB
    from ddg_amd import curved as CV
    from ddg_amd import mesh as pm
    from ddg_amd.operator import SipgOperator
    VX, VY, E = pm.rectangle_with_square_hole_mesh_2d(-2, 2, -2, 2, n, n, -0.5, 0.5, -0.5, 0.5)
    m = CV.apply_circular_hole_2d(pm.build_mesh_2d(VX, VY, E, N), cx=0.0, cy=0.0, radius=0.5,
                                  detection_tol=0.5)
    bc = pm.all_dirichlet(m)
    out = {"workload": f"annulus 2x({n}^2-{n // 4}^2) curved triangles, N={N}, float64",
           "K": int(m.K), "Np": int(m.Np)}
    u = torch.randn(m.K * m.Np, dtype=torch.float64, device="cuda")
    for name, cub in (("nodal", None), ("cubature", CB.attach_cubature(m))):
        op = SipgOperator(m, bc, cub=cub)
        t, _ = time_apply(op, u, torch.empty_like(u), steps, warmup)
        ncub = 0 if cub is None else int(np.asarray(cub.w).size)
        B = curved_apply_bytes(m.K, m.Np, m.Nfp, cub is not None, 8)
        out[name] = {"us": round(t * 1e6, 2), "GDOF/s": round(m.K * m.Np / t / 1e9, 2),
                     "bytes": int(B), "frac": round(B / t / 1e9 / HBM_PEAK_GBS, 4),
                     "ncub": ncub}
        del op
    print(f"[bench] curved: {out}", file=sys.stderr, flush=True)
    return out


def copy_bandwidth(steps, warmup, mib=256):
    """Measured HBM bandwidth of a plain device-to-device copy (read + write
    bytes per second): what the box's memory sustains, beside the nominal
    8 TB/s the roofline fractions use."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.randn(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)

    class _Copy:
        def apply(self, x, y):
            y.copy_(x)
    t, _ = time_apply(_Copy(), a, b, max(20, steps // 4), warmup)
    del a, b
    return 2 * n * 8 / t / 1e9


def time_apply(op, u, y, steps, warmup, **kw):
    """Average device time of one apply.  The applies are captured once into
    a HIP graph and replayed, so host launch overhead is not in the timed
    region; HIP events on the launch stream bracket the replays."""
    import torch
    for _ in range(warmup):
        op.apply(u, y, **kw)
    torch.cuda.synchronize()
    reps = 10
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            for _ in range(reps):
                op.apply(u, y, **kw)
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    n_rep = max(1, steps // reps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with torch.cuda.stream(s):
        e0.record(s)
        for _ in range(n_rep):
            graph.replay()
        e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1e3 / (n_rep * reps), wall


def shard_block(n, N, steps, warmup, parts=8):
    """The float64 and float32 applies on one GPU's share of the 2 n^2 mesh
    split into ``parts`` contiguous element ranges of the forest order: the
    first range, a compact patch (the subtree of one root triangle).  Its cut
    faces are tagged Dirichlet: in a sharded apply they would read the
    neighbour GPU's traces from the same LDS halo, at the same cost."""
    import torch
    from ddg_amd import mesh as pm
    from ddg_amd.operator import SipgOperator
    lv = int(round(np.log2(n))) - 1
    VX, VY, EToV = pm.forest_rectangle_hierarchy(2, lv)[-1]
    VX, VY = pm.jitter_vertices(VX, VY, 1.0 / n, 0.1, 0)
    sub = EToV[:len(EToV) // parts]
    used, inv = np.unique(sub, return_inverse=True)
    m = pm.build_mesh_2d(VX[used], VY[used], inv.reshape(sub.shape).astype(EToV.dtype), N)
    bc = pm.all_dirichlet(m)
    out = {"workload": f"1/{parts} of the 2x{n}x{n} N={N} forest mesh (first element range), "
                       "cut faces Dirichlet", "K": m.K, "tiles": -(-m.K // 64)}
    for dt, s, name in ((torch.float64, 8, "f64"), (torch.float32, 4, "f32")):
        op = SipgOperator(m, bc, dtype=dt)
        u = torch.randn(m.K * m.Np, dtype=torch.float64, device="cuda").to(dt)
        t, _ = time_apply(op, u, torch.empty_like(u), steps, warmup)
        out[name] = {"us": round(t * 1e6, 3), "GDOF/s": round(m.K * m.Np / t / 1e9, 2),
                     "frac": round(apply_bytes(m.K, m.Np, s) / t / 1e9 / HBM_PEAK_GBS, 4)}
        print(f"[bench] shard apply {name}: K={m.K} {t * 1e6:.2f} us", file=sys.stderr,
              flush=True)
    return out


def cpu_baseline(m, bc, budget_s=12.0):
    """Oracle apply (numpy float64) on the same mesh, timed on host cores."""
    from oracle import sipg as S
    rng = np.random.default_rng(0)
    u = rng.standard_normal((m.Np, m.K))
    M = S.local_mass(m)
    m1d = S.face_1d_mass_matrices(m)
    t0 = time.perf_counter()
    S.sipg_apply_2d(m, u, bc_types=bc, M_ref=M, m1d_per_face=m1d)
    n, t = 1, time.perf_counter() - t0
    while t < budget_s and n < 50:
        S.sipg_apply_2d(m, u, bc_types=bc, M_ref=M, m1d_per_face=m1d)
        n += 1
        t = time.perf_counter() - t0
    return t / n, n


def _flat(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a.T).reshape(-1), device=dev)


def timed_solves(fn, n):
    import torch
    lat, out = [], None
    for _ in range(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    return np.array(lat), out


def lat_stats(lat):
    return {"solves": int(lat.size), "solves_per_s": round(lat.size / float(lat.sum()), 3),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}


FIXTURES = ROOT / "tests" / "golden" / "config_solutions.npz"


def fixture_error(x, name):
    """Error of the device solution x against the oracle's float64 solution
    of the same system (tests/golden/config_solutions.npz, made by
    tests/golden/make_solution_fixtures.py): relative L2 and max norms over
    the fixture's 4096 seeded nodes, and the largest relative error of its
    eight seeded projections <w_j, x>.  None when the fixture lacks the case."""
    import torch
    if not FIXTURES.exists():
        return None
    fx = np.load(FIXTURES)
    if f"{name}/idx" not in fx:
        return None
    idx = torch.as_tensor(fx[f"{name}/idx"], device=x.device)
    val = fx[f"{name}/val"]
    xs = x[idx].double().cpu().numpy()
    xh = x.double().cpu().numpy()
    proj = np.array([np.dot(np.random.default_rng(1000 + j).standard_normal(xh.size), xh)
                     for j in range(len(fx[f"{name}/proj"]))])
    pref = fx[f"{name}/proj"]
    return {"rel_l2_sampled": float(np.linalg.norm(xs - val) / np.linalg.norm(val)),
            "rel_max_sampled": float(np.abs(xs - val).max() / np.abs(val).max()),
            "rel_proj": float(np.max(np.abs(proj - pref) / np.abs(pref))),
            "oracle_iterations": int(fx[f"{name}/iterations"]),
            "oracle_true_rel_residual": float(fx[f"{name}/true_rel_residual"])}


def solver_block(m, bc, b, mg64, mg32, n_solves, tol, inner_tol, fixture=None, ue=None, log=""):
    """FP64 MG-PCG, mixed-precision refinement and FP64 CG with the FP32
    V-cycle on one right-hand side; errors against the oracle's solution."""
    import torch
    from ddg_amd import sipg as D
    from ddg_amd.operator import refine_solve
    op64 = D.get_operator(m, bc)
    op32 = D.get_operator(m, bc, dtype=torch.float32)
    bn = float(torch.linalg.norm(b))
    res = {}
    for name, fn in (
            ("fp64", lambda: op64.pcg(b, pc=mg64, tol=tol, maxiter=5000)),
            ("mixed", lambda: refine_solve(op64, op32, b, pc32=mg32, tol=tol,
                                           inner_tol=inner_tol, inner_maxiter=2000)),
            ("mixed_pc", lambda: op64.pcg(b, pc=mg32, tol=tol, maxiter=5000))):
        fn()  # warm-up
        lat, (x, info) = timed_solves(fn, n_solves)
        d = lat_stats(lat)
        d["iterations"] = info["iterations"]
        if "outer" in info:
            d["outer"] = info["outer"]
        d["converged"] = info["converged"]
        if "at_floor" in info:
            d["at_residual_floor"] = info["at_floor"]
        # the float64 residual of the returned x itself (CG stops on its
        # recursively updated residual, refinement on this one)
        d["true_rel_residual"] = float(torch.linalg.norm(b - op64.apply(x))) / bn
        d["err_vs_oracle"] = fixture_error(x, fixture) if fixture else None
        if ue is not None:
            xs = x.cpu().numpy().reshape(m.K, m.Np).T
            d["rel_L2_err_vs_exact"] = float(np.sqrt(np.sum(m.J * (xs - ue) ** 2) /
                                                     np.sum(m.J * ue ** 2)))
        res[name] = d
        print(f"[bench] {log} {name}: {d}", file=sys.stderr, flush=True)
    res["speedup_mixed_vs_fp64"] = round(res["fp64"]["p50_ms"] / res["mixed"]["p50_ms"], 3)
    res["speedup_mixed_pc_vs_fp64"] = round(res["fp64"]["p50_ms"] / res["mixed_pc"]["p50_ms"], 3)
    return res


def plain_block(m, bc, b, tol, mixed_pc_ms):
    """The same solve by float64 CG without a preconditioner (the plain CG)
    and with block-Jacobi, one timed solve each, beside the float64 CG with
    the float32 V-cycle (mixed_pc_ms, its p50)."""
    from ddg_amd import sipg as D
    op = D.get_operator(m, bc)
    res = {}
    for name, pc in (("fp64_plain_cg", None), ("fp64_block_jacobi_cg", op.block_jacobi())):
        lat, (x, info) = timed_solves(lambda: op.pcg(b, pc=pc, tol=tol, maxiter=400000), 1)
        res[name] = {"ms": round(float(lat[0]) * 1e3, 1), "iterations": info["iterations"],
                     "converged": info["converged"],
                     "err_vs_oracle": fixture_error(x, "config1")}
        print(f"[bench] config1 {name}: {res[name]}", file=sys.stderr, flush=True)
    res["speedup_mixed_pc_vs_fp64_plain_cg"] = round(res["fp64_plain_cg"]["ms"] / mixed_pc_ms, 1)
    res["speedup_mixed_pc_vs_fp64_block_jacobi_cg"] = round(
        res["fp64_block_jacobi_cg"]["ms"] / mixed_pc_ms, 1)
    return res


def serve_block(n_requests, slots, method="pcg32", calib=200, load=0.8, sample=8, seed=0,
                max_batch=8, batch_dofs=1 << 19, ab_requests=1000):
    """The request stream: Helmholtz solves on 2 n^2 meshes, n in {16, 32, 64,
    128}, N in {2..6}, k^2 in [0, 50], Fourier right-hand sides, mixed
    precision to 1e-10 (``method``: "refine", iterative refinement; "pcg32",
    float64 CG with the float32 V-cycle, requests of one class batched up to
    ``max_batch`` / ``batch_dofs``).  A closed-loop burst of ``calib``
    requests measures the capacity C; then ``n_requests`` arrive as a Poisson
    stream at ``load * C`` per second.  With ``ab_requests`` (method "pcg32"):
    the capacity C1 without batching (one request per slot), and two streams
    of that many requests at ``load * C1``, without and with batching (the
    same arrivals).  A sample of results is checked against a float64
    block-Jacobi PCG solve to tol 1e-13."""
    import torch
    from ddg_amd import serve as SV
    srv = SV.SolveServer(SV.ServeConfig(slots=slots, method=method, max_batch=max_batch,
                                        batch_dofs=batch_dofs))
    reqs = SV.make_requests(n_requests, seed=seed)
    t0 = time.perf_counter()
    srv.prepare(reqs)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    print(f"[bench] serve: {len(srv.classes)} resident classes built in {setup:.1f} s",
          file=sys.stderr, flush=True)
    burst = SV.poisson_arrivals(SV.make_requests(calib, seed=seed + 1), None)
    cap = srv.run(burst).stats()
    print(f"[bench] serve burst: {cap}", file=sys.stderr, flush=True)
    rate = load * cap["solves_per_s"]
    SV.poisson_arrivals(reqs, rate, seed=seed + 2)
    rep = srv.run(reqs, progress_every=500)
    st = rep.stats()
    print(f"[bench] serve stream at {rate:.1f}/s: {st}", file=sys.stderr, flush=True)
    ab = None
    if ab_requests and method == "pcg32":
        keep = ["solves_per_s", "p50_ms", "p99_ms", "service_p50_ms", "service_p99_ms",
                "converged", "mean_batch"]
        srv.cfg.max_batch = 1
        cap1 = srv.run(SV.poisson_arrivals(SV.make_requests(calib, seed=seed + 1), None)).stats()
        rate1 = load * cap1["solves_per_s"]
        ab = {"capacity_unbatched_per_s": cap1["solves_per_s"], "rate_per_s": round(rate1, 2),
              "requests": ab_requests}
        for label, mb in (("unbatched", 1), ("batched", max_batch)):
            srv.cfg.max_batch = mb
            rs = SV.poisson_arrivals(SV.make_requests(ab_requests, seed=seed + 3), rate1,
                                     seed=seed + 4)
            sab = srv.run(rs).stats()
            print(f"[bench] serve {label} at {rate1:.1f}/s: {sab}", file=sys.stderr, flush=True)
            ab[label] = {k: sab[k] for k in keep}
        srv.cfg.max_batch = max_batch
    errs = []
    for r in reqs[:sample]:
        pcl = srv.problem_class(r.n, r.N)
        x, info = srv.solve_one(r, srv.streams[0])
        f = SV.fourier_field(pcl.x, pcl.y, r.seed)
        b = pcl.op64.rhs(f)
        xr, _ = pcl.op64.pcg(b, alpha=r.k2, pc=pcl.op64.block_jacobi(r.k2), tol=1e-13,
                             maxiter=50000)
        errs.append(float(torch.linalg.norm(x - xr) / torch.linalg.norm(xr)))
    srv.close()
    out = {"method": method, "requests": st["requests"], "offered_rate_per_s": round(rate, 2),
           "solves_per_s": st["solves_per_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
           "service_p50_ms": st["service_p50_ms"], "service_p99_ms": st["service_p99_ms"],
           "converged": st["converged"], "at_residual_floor": st["at_residual_floor"],
           "max_rel_residual": st["max_rel_residual"], "slots": slots,
           "peak_slots_busy": st["peak_slots_busy"], "max_batch_allowed": max_batch,
           "batch_dofs": batch_dofs,
           "mean_batch": st["mean_batch"], "max_batch": st["max_batch"],
           "burst_mean_batch": cap["mean_batch"],
           "closed_loop_capacity_per_s": cap["solves_per_s"], "setup_s": round(setup, 2),
           "same_rate_ab": ab,
           "tol": 1e-10, "sample_max_rel_err_vs_fp64": max(errs) if errs else None,
           "sample": f"first {len(errs)} requests re-solved by float64 block-Jacobi PCG to 1e-13"}
    return out


def config3_block(n_grad=3, n_fd=4):
    """Differentiation (config[3]): -div(kappa grad u) = f on 2x64x64, N=3,
    per-element kappa log-normal (sigma 0.5) and per-element tau_scale,
    float64, block-Jacobi PCG to 1e-14 (the test's settings); loss = half the
    M-norm of the misfit to the solution of a perturbed kappa.  Times one
    loss + gradient (forward solve, adjoint solve, parameter-gradient kernel)
    and checks the gradient against centred differences (eps 1e-6) on
    ``n_fd`` seeded directions."""
    import torch
    from ddg_amd import mesh as pm
    from ddg_amd.adjoint import SolveSpec, mass_apply, sipg_solve
    from ddg_amd.operator import SipgOperator
    dev = torch.device("cuda")
    VX, VY, E = pm.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 64, 64)
    VX, VY = pm.jitter_vertices(VX, VY, 1.0 / 64, 0.1, 3)
    m = pm.build_mesh_2d(VX, VY, E, 3)
    bc = pm.dirichlet_x_neumann_y(m)
    rng = np.random.default_rng(33)
    kap = np.exp(0.5 * rng.standard_normal(m.K))
    tau = 200.0 * rng.uniform(0.8, 1.2, m.K)
    ud = np.exp(m.x) * np.cos(3 * m.y)
    spec = SolveSpec(m, bc, u_dir=_flat(ud, dev), preconditioner="bj", tol=1e-14,
                     maxiter=30000)
    ft = _flat(8.0 * np.exp(m.x) * np.cos(3 * m.y), dev)
    mop = SipgOperator(m, bc)
    with torch.no_grad():
        target = sipg_solve(torch.tensor(kap * np.exp(0.3 * rng.standard_normal(m.K)),
                                         device=dev), torch.tensor(tau, device=dev), ft, spec)

    def loss(k, t):
        e = sipg_solve(k, t, ft, spec) - target
        return 0.5 * torch.dot(e, mass_apply(e, mop))

    times = []
    for _ in range(n_grad):
        kt = torch.tensor(kap, device=dev, requires_grad=True)
        tt = torch.tensor(tau, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = loss(kt, tt)
        L.backward()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    g = torch.cat([kt.grad, tt.grad]).cpu().numpy()
    gn = float(np.linalg.norm(g))
    errs = []
    with torch.no_grad():
        for _ in range(n_fd):
            d = rng.standard_normal(2 * m.K)
            d /= np.linalg.norm(d)
            dk, dt = torch.tensor(d[:m.K], device=dev), torch.tensor(d[m.K:], device=dev)
            k0, t0_ = kt.detach(), tt.detach()
            fd = (loss(k0 + 1e-6 * dk, t0_ + 1e-6 * dt)
                  - loss(k0 - 1e-6 * dk, t0_ - 1e-6 * dt)) / 2e-6
            errs.append(abs(fd.item() - float(g @ d)) / gn)
    out = {"workload": "2x64x64, N=3, log-normal kappa and per-element tau, block-Jacobi PCG "
                       "to 1e-14, float64",
           "K": int(m.K), "dofs": int(m.K * m.Np),
           "loss_and_gradient_ms": round(1e3 * float(np.median(times)), 2),
           "what": "forward solve + adjoint solve (same operator and preconditioner) + "
                   "ddgx_sipg_param_grad, median of repeats",
           "fd_directions": n_fd, "fd_max_rel_err": max(errs), "grad_norm": gn,
           "loss": float(L.item())}
    print(f"[bench] config3: {out}", file=sys.stderr, flush=True)
    return out


def config0_block(n_solves, cpu):
    """The reference's CPU-runnable case: 2x16x16, N=3, all-wall Dirichlet,
    f = 2 pi^2 sin(pi x) sin(pi y), tau_scale = 200, block-Jacobi CG to tol
    1e-9.  GPU: the numpy-in / numpy-out drop-in solve (preconditioner set-up
    excluded).  With ``cpu``: the oracle's same block-Jacobi CG timed on the
    host cores, and the GPU solution's error against the oracle's LU solve."""
    from ddg_amd import mesh as pm
    from ddg_amd import sipg as D
    VX, VY, EToV = pm.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 16, 16)
    m0 = pm.build_mesh_2d(VX, VY, EToV, 3)
    bc0 = pm.all_dirichlet(m0)
    f = 2 * np.pi ** 2 * np.sin(np.pi * m0.x) * np.sin(np.pi * m0.y)
    pc = D.make_sipg_block_jacobi_preconditioner(m0, bc_types=bc0, matfree=True)
    # maxiter 2000: the solve needs 1208 iterations (the reference's default
    # cap of 1000 stops it unconverged, 9e-10 from the LU solution)
    it = D.IterativeConfig(tol=1e-9, maxiter=2000)

    def gpu_solve():
        return D.solve_sipg_poisson_2d(m0, f, bc_types=bc0, preconditioner=pc, iterative=it,
                                       return_info=True)
    gpu_solve()
    lat, (u, info) = timed_solves(gpu_solve, n_solves)
    out = {"workload": "2x16x16, N=3, Dirichlet walls, tau_scale=200, block-Jacobi CG, tol 1e-9, "
                       "maxiter 2000",
           "gpu": dict(lat_stats(lat), iterations=info["iterations"],
                       converged=info["converged"],
                       note="drop-in call, numpy in and out (host transfers included)")}
    if cpu:
        from oracle import mesh as om
        from oracle import sipg as S
        mo = om.build_mesh_2d(VX, VY, EToV, 3)
        opc = S.make_block_jacobi_preconditioner(mo, bc_types=bc0)
        t0 = time.perf_counter()
        uo, io = S.solve_sipg_poisson_2d(mo, f, bc_types=bc0, preconditioner=opc,
                                         iterative=it, return_info=True)
        t_cpu = time.perf_counter() - t0
        ulu = S.solve_sipg_poisson_2d(mo, f, bc_types=bc0)
        out["cpu_oracle"] = {"seconds_per_solve": round(t_cpu, 4), "iterations": io["iterations"],
                             "cores": os.cpu_count(), "kind": "port"}
        out["gpu_rel_err_vs_oracle_lu"] = float(np.linalg.norm(u - ulu) / np.linalg.norm(ulu))
        out["speedup_gpu_vs_cpu_oracle"] = round(t_cpu / out["gpu"]["p50_ms"] * 1e3, 1)
    print(f"[bench] config0: {out}", file=sys.stderr, flush=True)
    return out


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the same counters on a copy of known size), kept in
    profiles/<name>; None when absent."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--solves", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-13,
                    help="CG relative residual tolerance (1e-13: every config[1] solve within "
                         "1e-10 relative L2 error of the oracle's solution)")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config2", action="store_true")
    ap.add_argument("--no-curved", action="store_true", help="skip the curved-element apply")
    ap.add_argument("--config3-only", action="store_true",
                    help="run only the differentiation leg (config[3])")
    ap.add_argument("--curved-only", action="store_true",
                    help="only time the curved apply (for rocprofv3)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the config-1 solves by float64 CG without a preconditioner "
                         "and with block-Jacobi (about 15 s)")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests in the serving stream (0: skip the serving leg)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-max-batch", type=int, default=8,
                    help="requests of one class run together in a slot, at most this many")
    ap.add_argument("--serve-batch-dofs", type=int, default=1 << 19,
                    help="... and at most this many unknowns per batch")
    ap.add_argument("--serve-method", default="pcg32", choices=["pcg32", "refine"])
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--order", default="forest", choices=["forest", "rowmajor"])
    ap.add_argument("--sweep", action="store_true",
                    help="time the apply for N=1..6 in float64 and float32 and exit")
    ap.add_argument("--sweep-n", type=int, nargs="*", default=[1, 2, 3, 4, 5, 6])
    ap.add_argument("--copy", action="store_true",
                    help="time a device-to-device copy at several sizes and exit")
    ap.add_argument("--profile-apply", choices=["f64", "f32"], default=None,
                    help="only run --steps eager applies (for rocprofv3 counter runs)")
    ap.add_argument("--mg-degrees", type=int, nargs="*", default=None,
                    help="p-levels of the config-1 hp-multigrid (default 4 2 1); the last is the "
                         "degree of the h-levels")
    ap.add_argument("--profile-solve", action="store_true",
                    help="only run a few config-1 solves (FP64 MG-PCG and mixed) for rocprofv3")
    ap.add_argument("--profile-copy", type=int, default=0, metavar="MB",
                    help="only run --steps elementwise copies y = 2 x of MB megabytes each "
                         "(calibrates the HBM byte counters)")
    ap.add_argument("--vcycle", action="store_true",
                    help="only time one application of the config-1 hp-multigrid V-cycle, "
                         "float32 and float64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the SIPG apply and "
              "its CG run on one GPU; sharding them (element ranges with a halo-trace "
              "exchange per apply, two scalar all-reduces per CG iteration) is not built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from ddg_amd.operator import SipgOperator

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    if args.copy:
        for mb in (8, 16, 24, 48, 96, 256, 1024):
            n = mb * (1 << 20) // 8
            a = torch.randn(n, dtype=torch.float64, device="cuda")
            b = torch.empty_like(a)

            class _C:
                def apply(self, x, y):
                    y.copy_(x)
            t, _ = time_apply(_C(), a, b, args.steps, args.warmup)
            print(f"[copy] {mb:5d} MB read + write: {t * 1e6:9.2f} us  {2 * n * 8 / t / 1e9:8.1f} GB/s",
                  flush=True)
        return
    if args.sweep:
        for N in args.sweep_n:
            m, bc = workload_mesh(args.n, N=N, order=args.order)
            for dt, s in ((torch.float64, 8), (torch.float32, 4)):
                op = SipgOperator(m, bc, dtype=dt)
                u = torch.randn(m.K * m.Np, dtype=torch.float64, device="cuda").to(dt)
                t, _ = time_apply(op, u, torch.empty_like(u), args.steps, args.warmup)
                frac = apply_bytes(m.K, m.Np, s) / t / 1e9 / HBM_PEAK_GBS
                print(f"[sweep] N={N} {str(dt)[6:]} K={m.K} {t * 1e6:8.2f} us "
                      f"{m.K * m.Np / t / 1e9:7.1f} GDOF/s  {frac * 100:5.1f}% HBM", flush=True)
        return
    if args.vcycle:
        m, _ = workload_mesh(args.n, order=args.order)
        out = {}
        for dt in (torch.float32, torch.float64):
            mg, _ = hpmg_for(m, 4, dt)
            r = torch.randn(m.K * m.Np, dtype=torch.float64, device="cuda").to(dt)

            class _V:
                def apply(self, x, y):
                    mg(x, y)
            t, _ = time_apply(_V(), r, torch.empty_like(r), args.steps, args.warmup)
            out[str(dt)[6:]] = round(t * 1e6, 1)
        print(json.dumps({"vcycle_us": out, "workload": "config-1 hp-multigrid, 2x256x256 N=4, "
                          "one preconditioner application (graph replays)"}), flush=True)
        return
    if args.profile_copy:
        n = args.profile_copy * (1 << 20) // 8
        a = torch.randn(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        print(f"[profile] {args.steps} copies of {args.profile_copy} MB", flush=True)
        return
    if args.config3_only:
        print(json.dumps({"config3": config3_block()}), flush=True)
        return
    if args.curved_only:
        print(json.dumps({"curved": curved_block(args.n, 4, args.steps, args.warmup)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serve": serve_block(args.serve_requests, args.serve_slots,
                                               args.serve_method,
                                               max_batch=args.serve_max_batch,
                                               batch_dofs=args.serve_batch_dofs)}),
              flush=True)
        return
    m, bc = workload_mesh(args.n, order=args.order)
    K, Np = m.K, m.Np
    if args.profile_apply:
        dt = torch.float64 if args.profile_apply == "f64" else torch.float32
        op = SipgOperator(m, bc, dtype=dt)
        u = torch.randn(K * Np, dtype=torch.float64, device="cuda").to(dt)
        y = torch.empty_like(u)
        for _ in range(args.steps):
            op.apply(u, y)
        torch.cuda.synchronize()
        print(f"[profile] {args.steps} applies {args.profile_apply} K={K} N={m.N}", flush=True)
        return

    if args.profile_solve:
        from ddg_amd import sipg as D
        from ddg_amd.operator import refine_solve
        mg64, fin = hpmg_for(m, 4, torch.float64, degrees=args.mg_degrees)
        mg32, _ = hpmg_for(fin, 4, torch.float32, degrees=args.mg_degrees)
        f, ud, q = problem_data(fin)
        op64 = D.get_operator(fin, bc)
        op32 = D.get_operator(fin, bc, dtype=torch.float32)
        b = op64.rhs(_flat(f, dev), u_dir=_flat(ud, dev), q_neu=_flat(q, dev))
        for _ in range(max(1, args.solves)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, i64 = op64.pcg(b, pc=mg64, tol=args.tol, maxiter=5000)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, imx = refine_solve(op64, op32, b, pc32=mg32, tol=args.tol, inner_tol=args.inner_tol,
                                  inner_maxiter=2000)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            _, ipc = op64.pcg(b, pc=mg32, tol=args.tol, maxiter=5000)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            print(f"[profile] fp64 {1e3 * (t1 - t0):.1f} ms {i64['iterations']} it; mixed "
                  f"{1e3 * (t2 - t1):.1f} ms {imx['iterations']} it {imx['outer']} outer; "
                  f"fp64 CG + fp32 V-cycle {1e3 * (t3 - t2):.1f} ms {ipc['iterations']} it",
                  flush=True)
        return

    op = SipgOperator(m, bc)
    print(f"[bench] mesh K={K} N={m.N} Np={Np} dofs={K * Np}", file=sys.stderr, flush=True)
    g = torch.Generator(device="cuda").manual_seed(0)
    u = torch.randn(K * Np, dtype=torch.float64, device="cuda", generator=g)
    y = torch.empty_like(u)
    t_apply, wall = time_apply(op, u, y, args.steps, args.warmup)
    gdof = K * Np / t_apply / 1e9
    B = apply_bytes(K, Np, 8)
    achieved = B / t_apply / 1e9
    print(f"[bench] apply f64: {t_apply * 1e6:.2f} us, {gdof:.1f} GDOF/s, {achieved:.0f} GB/s",
          file=sys.stderr, flush=True)
    op32 = SipgOperator(m, bc, dtype=torch.float32)
    u32 = u.float()
    t32, _ = time_apply(op32, u32, torch.empty_like(u32), args.steps, args.warmup)
    print(f"[bench] apply f32: {t32 * 1e6:.2f} us", file=sys.stderr, flush=True)
    del op, op32
    shard = shard_block(args.n, 4, args.steps, args.warmup) if args.order == "forest" else None
    copy = copy_bandwidth(args.steps, args.warmup)
    print(f"[bench] device copy: {copy:.0f} GB/s", file=sys.stderr, flush=True)

    # --- solves on the same mesh (hp-multigrid, tol 1e-10) ---
    from ddg_amd import sipg as D
    t0 = time.perf_counter()
    mg64, fin = hpmg_for(m, 4, torch.float64)
    mg32, _ = hpmg_for(fin, 4, torch.float32)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    print(f"[bench] hp-MG setup (f64 + f32): {t_setup:.2f} s, levels K={mg64.sizes} "
          f"N={mg64.degrees}", file=sys.stderr, flush=True)
    f, ud, q = problem_data(fin)
    opb = D.get_operator(fin, bc)
    b = opb.rhs(_flat(f, dev), u_dir=_flat(ud, dev), q_neu=_flat(q, dev))
    fx = args.n == 256 and args.order == "forest"
    cg1 = solver_block(fin, bc, b, mg64, mg32, args.solves, args.tol, args.inner_tol,
                       fixture="config1" if fx else None, ue=ud, log="config1 e^x cos3y")
    cg1["setup_s"] = round(t_setup, 3)
    if not args.no_plain:
        cg1["plain"] = plain_block(fin, bc, b, args.tol, cg1["mixed_pc"]["p50_ms"])
    fr = fourier_rhs(fin, seed=1)
    b2 = opb.rhs(_flat(fr, dev))
    cg2 = {"N4_2x%dx%d" % (args.n, args.n):
           solver_block(fin, bc, b2, mg64, mg32, max(3, args.solves // 4), args.tol,
                        args.inner_tol, fixture="config2_n4" if fx else None,
                        log="config2 fourier N=4")}
    if not args.no_config2:
        m6, bc6 = workload_mesh(args.n // 2, N=6)
        mg6_64, _ = hpmg_for(m6, 6, torch.float64)
        mg6_32, _ = hpmg_for(m6, 6, torch.float32)
        op6 = D.get_operator(m6, bc6)
        b6 = op6.rhs(_flat(fourier_rhs(m6, seed=2), dev))
        cg2["N6_2x%dx%d" % (args.n // 2, args.n // 2)] = solver_block(
            m6, bc6, b6, mg6_64, mg6_32, max(3, args.solves // 4), args.tol, args.inner_tol,
            fixture="config2_n6" if fx else None, log="config2 fourier N=6")

    c0 = config0_block(10, cpu=not args.no_cpu)
    c3 = config3_block()
    curved = None if args.no_curved else curved_block(args.n, 4, args.steps, args.warmup)

    serve = None
    if args.serve_requests > 0:
        serve = serve_block(args.serve_requests, args.serve_slots, args.serve_method,
                            max_batch=args.serve_max_batch, batch_dofs=args.serve_batch_dofs)

    cpu = None
    if not args.no_cpu:
        t_cpu, n_cpu = cpu_baseline(m, bc)
        cpu = {"value": round(K * Np / t_cpu / 1e9, 5), "unit": "GDOF/s",
               "cores": os.cpu_count(), "kind": "port",
               "sample": f"{n_cpu} float64 applies of the numpy restatement (oracle/sipg.py) "
                         f"on the same 2x{args.n}x{args.n} N=4 mesh, {t_cpu:.3f} s each; "
                         "numpy/BLAS threads = host cores"}
        print(f"[bench] cpu oracle apply: {t_cpu:.3f} s", file=sys.stderr, flush=True)

    out = {
        "metric": "SIPG matvec GDOF/s (% HBM roofline); CG solves/s & p99 at 1e-10 rel err",
        "value": round(gdof, 3),
        "unit": "GDOF/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_apply * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"SIPG apply, 2x{args.n}x{args.n} jittered triangles "
                               f"({args.order} order), N=4, Dirichlet x / Neumann y, "
                               "tau_scale=200, float64",
                   "K": K, "Np": Np, "dofs": K * Np},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("sipg_apply_f64_n4_k131072_traffic.txt")},
        "cpu_baseline": cpu,
        "hbm_copy": {"GB/s": round(copy, 1), "frac_of_peak": round(copy / HBM_PEAK_GBS, 4),
                     "apply_f64_frac_of_copy": round(achieved / copy, 4),
                     "what": "device-to-device copy of 256 MiB (float64), read + write bytes "
                             "over the average time of graph-replayed copies"},
        "apply_f32": {"us": round(t32 * 1e6, 3), "GDOF/s": round(K * Np / t32 / 1e9, 2),
                      "frac": round(apply_bytes(K, Np, 4) / t32 / 1e9 / HBM_PEAK_GBS, 4)},
        "cg": {"preconditioner": f"hp-multigrid levels K={mg64.sizes} N={mg64.degrees}",
               "tol": args.tol, "inner_tol": args.inner_tol, "config1": cg1, "config2": cg2,
               "config0": c0, "config3": c3},
        "serve": serve,
        "curved": curved,
        "shard": shard,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
