"""Write tests/golden/config_solutions.npz: the oracle's float64 solutions of
the benchmark solves, as seeded samples.

    python tests/golden/make_solution_fixtures.py [--cases config1 config2_n4 config2_n6]

Runs the oracle (NumPy float64 restatement of src/ddg/dg2d/sipg.py and
multigrid.py) on the CPU, about 30-40 minutes for all three cases:

* config1    -- 2x256x256 triangles of [0,1]^2 (uniform red refinement of
  2x2 cells: element 4k+c is child c of element k), interior vertices moved
  by U(-0.1h, 0.1h) from numpy.random.default_rng(0), N=4, Dirichlet on
  x = 0, 1 and Neumann on y = 0, 1, tau_scale = 200;
  -Delta u = 8 e^x cos 3y with u = e^x cos 3y on the Dirichlet walls and
  du/dn on the Neumann walls (sipg.py:812-864);
* config2_n4 -- the same mesh and operator, right-hand side J M f for the
  64-mode Fourier series f of seed 1 (amplitudes a_m |k_m|^-2,
  a ~ N(0,1), phi ~ U(0, 2 pi), integer k with |k_i| <= 8), no boundary data;
* config2_n6 -- 2x128x128 (same construction and jitter seed), N=6, the
  Fourier series of seed 2.

Each system is solved by the oracle's CG (jax.scipy.sparse.linalg.cg
semantics) preconditioned by its hp-multigrid V-cycle (p-levels N, N/2, 1 on
the mesh, h-levels at degree 1 down to 2x16x16; multigrid.py:301-367) to a
relative residual of 1e-14 from x0 = 0.  The fixture keeps, per case, 4096
seeded node indices of the element-major solution with their values, eight
projections <w_j, x> on seeded N(0,1) vectors w_j (seed 1000+j), the norms,
the iteration count and the true relative residual |b - A x| / |b|.
"""

from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import mesh as om          # noqa: E402
from oracle import multigrid as omg    # noqa: E402
from oracle import sipg as S           # noqa: E402
from oracle.cg import cg as jax_cg     # noqa: E402

OUT = Path(__file__).resolve().parent / "config_solutions.npz"
N_SAMPLE = 4096
N_PROJ = 8


def forest_vertices(n):
    """2 n^2 triangles by uniform refinement of the 2x2-cell mesh."""
    VX, VY, E = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 2, 2)
    levels = int(round(np.log2(n))) - 1
    return omg.forest_hierarchy(VX, VY, E, levels)


def jitter(VX, VY, h, frac=0.1, seed=0):
    """Interior vertices moved by U(-frac h, frac h) per coordinate; one
    (2, nv) draw from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    VX, VY = np.array(VX, dtype=np.float64), np.array(VY, dtype=np.float64)
    tol = 1e-12
    inner = (np.abs(VX) > tol) & (np.abs(VX - 1) > tol) & (np.abs(VY) > tol) & (np.abs(VY - 1) > tol)
    d = rng.uniform(-frac * h, frac * h, size=(2, VX.size))
    VX[inner] += d[0, inner]
    VY[inner] += d[1, inner]
    return VX, VY


def bc_dirichlet_x_neumann_y(m):
    """Boundary faces on x = 0, 1 Dirichlet, on y = 0, 1 Neumann."""
    E = np.asarray(m.EToE)
    bdy = E == np.arange(m.K)[:, None]
    nx = m.nx.reshape((m.Nfp, 3, m.K), order="F")[0].T
    return np.where(bdy, np.where(np.abs(nx) > 0.5, 1, 2), 0).astype(np.int32)


def fourier(m, seed, modes=64):
    rng = np.random.default_rng(seed)
    f = np.zeros_like(m.x)
    for _ in range(modes):
        k = np.zeros(2)
        while not k.any():
            k = rng.integers(-8, 9, size=2).astype(np.float64)
        a, phi = rng.standard_normal(), rng.uniform(0, 2 * np.pi)
        f += a / (k @ k) * np.cos(2 * np.pi * (k[0] * m.x + k[1] * m.y) + phi)
    return f


def exact_data(m):
    ue = np.exp(m.x) * np.cos(3 * m.y)
    xf = m.x.reshape(-1, order="F")[m.vmapM]
    yf = m.y.reshape(-1, order="F")[m.vmapM]
    q = (m.nx.reshape(-1, order="F") * np.exp(xf) * np.cos(3 * yf)
         - m.ny.reshape(-1, order="F") * 3 * np.exp(xf) * np.sin(3 * yf))
    return 8.0 * ue, ue, q.reshape((3 * m.Nfp, m.K), order="F")


def solve_case(n, N, rhs_kind, log):
    t0 = time.time()
    hier = forest_vertices(n)
    VX, VY, E = hier[-1]
    VX, VY = jitter(VX, VY, 1.0 / n)
    m = om.build_mesh_2d(VX, VY, E, N)
    bc = bc_dirichlet_x_neumann_y(m)
    log(f"mesh K={m.K} N={N} built in {time.time() - t0:.1f} s")
    if rhs_kind == "exact":
        f, ud, q = exact_data(m)
        b = S._rhs(m, f, bc, ud, q, 200.0)
    else:
        b = S._rhs(m, fourier(m, rhs_kind), bc, None, None, 200.0)
    t0 = time.time()
    bVX, bVY, bE = om.uniform_rectangle_mesh_2d(0.0, 1.0, 0.0, 1.0, 2, 2)
    base = omg.forest_hierarchy(bVX, bVY, bE, 3)[-1]          # 2x16x16
    n_h = int(round(np.log2(n))) - 4
    degrees = [N, N // 2, 1]
    mg, fin = omg.make_sipg_hpmg_preconditioner(*base, N, n_h, bc_fn=bc_dirichlet_x_neumann_y,
                                                degrees=degrees, fine_vertices=(VX, VY))
    assert np.array_equal(np.asarray(fin.EToV), np.asarray(E))
    log(f"hp-multigrid {degrees} + {n_h} h-levels built in {time.time() - t0:.1f} s")
    A = S.flat_apply(m, bc, 200.0)
    t0 = time.time()
    x, info = jax_cg(A, b, np.zeros_like(b), tol=1e-14, maxiter=2000, M=mg)
    res = float(np.linalg.norm(b - A(x)) / np.linalg.norm(b))
    log(f"PCG {info['iterations']} iterations in {time.time() - t0:.1f} s, "
        f"true relative residual {res:.2e}")
    rng = np.random.default_rng(7 + N)
    idx = np.sort(rng.choice(x.size, N_SAMPLE, replace=False)).astype(np.int64)
    proj = np.array([np.dot(np.random.default_rng(1000 + j).standard_normal(x.size), x)
                     for j in range(N_PROJ)])
    return {"n": n, "N": N, "K": m.K, "idx": idx, "val": x[idx], "proj": proj,
            "norm2": float(np.linalg.norm(x)), "norminf": float(np.abs(x).max()),
            "bnorm2": float(np.linalg.norm(b)), "iterations": int(info["iterations"]),
            "true_rel_residual": res}


CASES = {"config1": (256, 4, "exact"), "config2_n4": (256, 4, 1), "config2_n6": (128, 6, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", nargs="*", default=list(CASES))
    args = ap.parse_args()
    data = dict(np.load(OUT)) if OUT.exists() else {}
    for name in args.cases:
        n, N, kind = CASES[name]
        out = solve_case(n, N, kind, lambda s: print(f"[{name}] {s}", flush=True))
        for k, v in out.items():
            data[f"{name}/{k}"] = np.asarray(v)
        np.savez_compressed(OUT, **data)
        print(f"[{name}] written to {OUT}", flush=True)


if __name__ == "__main__":
    main()
