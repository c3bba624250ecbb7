"""Benchmark: SIPG matvec GDOF/s (% of HBM roofline) and CG solves/s with
p50/p99 latency, on one MI355X.

Workload (one GPU): the 2x256x256 triangle mesh of [0,1]^2 with interior
vertices jittered by a seeded U(-0.1h, 0.1h), polynomial order N=4
(Np=15, K=131072, 1,966,080 DOFs), Dirichlet on x = 0, 1 and Neumann on
y = 0, 1, penalty tau_scale = 200, float64.  A "step" is one SIPG apply over
the whole mesh with its input already resident in HBM.

Also reported: block-Jacobi preconditioned CG solves of the same problem
(f = 8 e^x cos 3y, u = e^x cos 3y on the Dirichlet walls, dn u on the
Neumann walls) to tol 1e-10, and the CPU baseline: the float64 numpy
restatement of the reference algorithm (oracle/) timed on the host cores.

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--solves 3]
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


def apply_bytes(K, Np, s):
    """Algorithmic bytes per apply: read u, write y, 5 element + 9 face
    geometry words and 3 int32 face codes per element."""
    return K * (s * (2 * Np + 14) + 12)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--solves", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--order", default="forest", choices=["forest", "rowmajor"])
    ap.add_argument("--sweep", action="store_true",
                    help="time the apply for N=1..6 in float64 and float32 and exit")
    ap.add_argument("--sweep-n", type=int, nargs="*", default=[1, 2, 3, 4, 5, 6])
    ap.add_argument("--copy", action="store_true",
                    help="time a device-to-device copy at several sizes and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the SIPG path is "
              "single-GPU in this build, see DESIGN.md for how it would shard.", file=sys.stderr)
        sys.exit(2)

    import torch
    from ddg_amd import sipg as D
    from ddg_amd.operator import SipgOperator

    torch.cuda.set_device(0)
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
    m, bc = workload_mesh(args.n, order=args.order)
    K, Np = m.K, m.Np
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

    # float32 apply of the same operator (reported alongside)
    op32 = SipgOperator(m, bc, dtype=torch.float32)
    u32 = u.float()
    y32 = torch.empty_like(u32)
    t32, _ = time_apply(op32, u32, y32, args.steps, args.warmup)

    # block-Jacobi CG solves to 1e-10
    f, ud, q = problem_data(m)
    pc = D.make_sipg_block_jacobi_preconditioner(m, bc_types=bc)
    cfg = D.IterativeConfig(tol=1e-10, maxiter=20000)
    lat, iters = [], []
    ue = ud
    us = np.zeros_like(ue)
    for i in range(args.solves):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        us, info = D.solve_sipg_poisson_2d(m, f, bc_types=bc, u_dir=ud, q_neu=q, iterative=cfg,
                                           preconditioner=pc, return_info=True)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        iters.append(info["iterations"])
        print(f"[bench] BJ-PCG solve {i}: {lat[-1]:.3f} s, {info['iterations']} it, "
              f"converged={info['converged']}", file=sys.stderr, flush=True)
    err = float(np.sqrt(np.sum(m.J * (us - ue) ** 2) / np.sum(m.J * ue ** 2)))

    cpu = None
    if not args.no_cpu:
        t_cpu, n_cpu = cpu_baseline(m, bc)
        cpu = {"value": K * Np / t_cpu / 1e9, "unit": "GDOF/s", "cores": os.cpu_count(),
               "kind": "port",
               "sample": f"{n_cpu} float64 applies of the numpy restatement (oracle/sipg.py) on "
                         f"the same 2x{args.n}x{args.n} N=4 mesh; numpy/BLAS threads = host cores"}
        print(f"[bench] cpu oracle apply: {t_cpu:.3f} s", file=sys.stderr, flush=True)

    lat_s = np.array(lat) if lat else np.array([np.nan])
    out = {
        "metric": "SIPG matvec GDOF/s (% HBM roofline); CG solves/s & p99 at 1e-10 rel err",
        "value": round(gdof, 3),
        "unit": "GDOF/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_apply * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2x{args.n}x{args.n} jittered triangles ({args.order} order), N=4, "
                               "Dirichlet x / Neumann y, tau_scale=200, float64 SIPG apply",
                   "K": K, "Np": Np, "dofs": K * Np},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "apply_f32": {"us": round(t32 * 1e6, 3), "GDOF/s": round(K * Np / t32 / 1e9, 2),
                      "frac": round(apply_bytes(K, Np, 4) / t32 / 1e9 / HBM_PEAK_GBS, 4)},
        "cg": {"preconditioner": "block-Jacobi", "tol": 1e-10, "solves": len(lat),
               "solves_per_s": round(len(lat) / float(lat_s.sum()), 4) if lat else None,
               "p50_s": round(float(np.percentile(lat_s, 50)), 4),
               "p99_s": round(float(np.percentile(lat_s, 99)), 4),
               "iterations": iters, "rel_L2_err_vs_exact": err},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
