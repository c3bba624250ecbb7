"""Oracle: nodal reference element on [-1,1] and on the triangle (float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``src/ddg/dg1d/reference.py:20-143`` and
``src/ddg/dg2d/reference.py:27-255`` (Hesthaven & Warburton, *Nodal
Discontinuous Galerkin Methods*, Codes1D/Codes2D).  Pinned by the reference's
known answers (``test/test_reference.py:18-51``,
``test/test_reference_2d.py:20-59``) in ``tests/test_oracle_reference.py``.
"""

from __future__ import annotations

import math

import numpy as np


# --------------------------------------------------------------------------
# 1D (src/ddg/dg1d/reference.py)
# --------------------------------------------------------------------------

def jacobi_p(x, alpha: float, beta: float, n: int) -> np.ndarray:
    """Orthonormal Jacobi polynomial P_n^(alpha,beta) at ``x``.

    Three-term recurrence in normalised form (dg1d/reference.py:20-68).
    """
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    g0 = (2.0 ** (alpha + beta + 1) / (alpha + beta + 1)
          * math.gamma(alpha + 1) * math.gamma(beta + 1)
          / math.gamma(alpha + beta + 1))
    p0 = np.full_like(x, 1.0 / math.sqrt(g0))
    if n == 0:
        return p0
    g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0
    p1 = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(g1)
    if n == 1:
        return p1
    a_prev = 2.0 / (2 + alpha + beta) * math.sqrt(
        (alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, n):
        h1 = 2 * i + alpha + beta
        a_next = 2.0 / (h1 + 2) * math.sqrt(
            (i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta)
            / (h1 + 1) / (h1 + 3))
        b_next = -(alpha ** 2 - beta ** 2) / h1 / (h1 + 2)
        p0, p1 = p1, (-a_prev * p0 + (x - b_next) * p1) / a_next
        a_prev = a_next
    return p1


def grad_jacobi_p(x, alpha: float, beta: float, n: int) -> np.ndarray:
    """d/dx of the orthonormal Jacobi polynomial (dg1d/reference.py:71-76)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    if n == 0:
        return np.zeros_like(x)
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def jacobi_gauss_quadrature(alpha: float, beta: float, n: int):
    """(n+1)-point Gauss rule for the Jacobi weight via Golub-Welsch
    (dg1d/reference.py:79-124)."""
    if n == 0:
        return (np.array([-(alpha - beta) / (alpha + beta + 2.0)]), np.array([2.0]))
    k = np.arange(n + 1, dtype=np.float64)
    h1 = 2.0 * k + alpha + beta
    with np.errstate(divide="ignore", invalid="ignore"):
        diag = -0.5 * (alpha ** 2 - beta ** 2) / (h1 + 2.0) / h1
    if abs(alpha + beta) < 10 * np.finfo(np.float64).eps:
        diag[0] = 0.0
    i = np.arange(1, n + 1, dtype=np.float64)
    hl = h1[:n]
    off = 2.0 / (hl + 2.0) * np.sqrt(
        i * (i + alpha + beta) * (i + alpha) * (i + beta) / (hl + 1.0) / (hl + 3.0))
    T = np.diag(diag) + np.diag(off, 1)
    T = T + T.T
    lam, vec = np.linalg.eigh(T)
    w = (vec[0, :] ** 2 * (2.0 ** (alpha + beta + 1) / (alpha + beta + 1))
         * math.gamma(alpha + 1) * math.gamma(beta + 1) / math.gamma(alpha + beta + 1))
    return lam, w


def jacobi_gauss_lobatto_nodes(alpha: float, beta: float, n: int) -> np.ndarray:
    """Gauss-Lobatto nodes: endpoints plus interior Gauss nodes of the
    (alpha+1, beta+1) rule of degree n-2 (dg1d/reference.py:127-136)."""
    if n == 1:
        return np.array([-1.0, 1.0])
    xi, _ = jacobi_gauss_quadrature(alpha + 1, beta + 1, n - 2)
    return np.concatenate([[-1.0], xi, [1.0]])


def vandermonde_1d(n: int, r) -> np.ndarray:
    """V[i, j] = P_j(r_i), Legendre (dg1d/reference.py:139-143)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    return np.stack([jacobi_p(r, 0.0, 0.0, j) for j in range(n + 1)], axis=1)


def grad_vandermonde_1d(n: int, r) -> np.ndarray:
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    return np.stack([grad_jacobi_p(r, 0.0, 0.0, j) for j in range(n + 1)], axis=1)


def d_matrix_1d(n: int, r, V) -> np.ndarray:
    """Dr = Vr V^-1 (dg1d/reference.py:153-156)."""
    Vr = grad_vandermonde_1d(n, r)
    return np.linalg.solve(V.T, Vr.T).T


# --------------------------------------------------------------------------
# 2D triangle (src/ddg/dg2d/reference.py)
# --------------------------------------------------------------------------

# Optimised blending parameters, Hesthaven & Warburton Table 6.1
# (dg2d/reference.py:27-30).
ALPHA_OPT = np.array([0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999,
                      1.2832, 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223,
                      1.6258])


def _warp(n: int, rout: np.ndarray) -> np.ndarray:
    """Edge warp (dg2d/reference.py:39-59): interpolant of LGL - equispaced
    displacement, divided by 1 - r^2 away from the endpoints."""
    lgl = jacobi_gauss_lobatto_nodes(0.0, 0.0, n)
    req = np.linspace(-1.0, 1.0, n + 1)
    Veq = vandermonde_1d(n, req)
    P = np.stack([jacobi_p(rout, 0.0, 0.0, i) for i in range(n + 1)], axis=0)
    L = np.linalg.solve(Veq.T, P)
    warp = L.T @ (lgl - req)
    inner = np.abs(rout) < 1.0 - 1e-10
    sf = 1.0 - np.where(inner, rout, 0.0) ** 2
    return warp / sf + warp * (inner.astype(np.float64) - 1.0)


def equilateral_nodes_2d(n: int):
    """Warp-and-blend nodes on the equilateral triangle
    (dg2d/reference.py:62-101).  Node order: rows n=1..N+1, m=1..N+2-n."""
    Np = (n + 1) * (n + 2) // 2
    L1 = np.zeros(Np)
    L3 = np.zeros(Np)
    sk = 0
    for a in range(1, n + 2):
        for b in range(1, n + 3 - a):
            L1[sk] = (a - 1) / n if n > 0 else 0.0
            L3[sk] = (b - 1) / n if n > 0 else 0.0
            sk += 1
    L2 = 1.0 - L1 - L3
    x = -L2 + L3
    y = (-L2 - L3 + 2.0 * L1) / math.sqrt(3.0)
    if n == 0:
        return x, y
    alpha = float(ALPHA_OPT[n - 1]) if 1 <= n <= 15 else 5.0 / 3.0
    w1 = 4.0 * L2 * L3 * _warp(n, L3 - L2) * (1.0 + (alpha * L1) ** 2)
    w2 = 4.0 * L1 * L3 * _warp(n, L1 - L3) * (1.0 + (alpha * L2) ** 2)
    w3 = 4.0 * L1 * L2 * _warp(n, L2 - L1) * (1.0 + (alpha * L3) ** 2)
    x = x + w1 + math.cos(2 * math.pi / 3) * w2 + math.cos(4 * math.pi / 3) * w3
    y = y + math.sin(2 * math.pi / 3) * w2 + math.sin(4 * math.pi / 3) * w3
    return x, y


def xy_to_rs(x, y):
    """Equilateral -> right reference triangle (dg2d/reference.py:104-111)."""
    L1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    L2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    L3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    return -L2 + L3 - L1, -L2 - L3 + L1


def rs_to_ab(r, s):
    """Collapsed coordinates (dg2d/reference.py:114-121)."""
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(np.abs(s - 1.0) > 1e-12, 2.0 * (1.0 + r) / (1.0 - s) - 1.0, -1.0)
    return a, s


def simplex_2d_p(a, b, i: int, j: int) -> np.ndarray:
    """Orthonormal Dubiner mode (i, j) (dg2d/reference.py:124-131)."""
    h1 = jacobi_p(a, 0.0, 0.0, i)
    h2 = jacobi_p(b, 2 * i + 1.0, 0.0, j)
    return math.sqrt(2.0) * h1 * h2 * ((1.0 - b) ** i if i > 0 else 1.0)


def grad_simplex_2d_p(a, b, i: int, j: int):
    """(d/dr, d/ds) of mode (i, j) (dg2d/reference.py:134-161)."""
    fa = jacobi_p(a, 0.0, 0.0, i)
    dfa = grad_jacobi_p(a, 0.0, 0.0, i)
    gb = jacobi_p(b, 2 * i + 1.0, 0.0, j)
    dgb = grad_jacobi_p(b, 2 * i + 1.0, 0.0, j)
    half = 0.5 * (1.0 - b)
    dr = dfa * gb
    if i > 0:
        dr = dr * half ** (i - 1)
    ds = dfa * (gb * (0.5 * (1.0 + a)))
    if i > 0:
        ds = ds * half ** (i - 1)
    t = dgb * half ** i
    if i > 0:
        t = t - 0.5 * i * gb * half ** (i - 1)
    ds = ds + fa * t
    scale = 2.0 ** (i + 0.5)
    return scale * dr, scale * ds


def vandermonde_2d(n: int, r, s) -> np.ndarray:
    """(dg2d/reference.py:164-171); mode order i outer, j inner."""
    a, b = rs_to_ab(r, s)
    return np.stack([simplex_2d_p(a, b, i, j)
                     for i in range(n + 1) for j in range(n - i + 1)], axis=1)


def grad_vandermonde_2d(n: int, r, s):
    a, b = rs_to_ab(r, s)
    cr, cs = [], []
    for i in range(n + 1):
        for j in range(n - i + 1):
            dr, ds = grad_simplex_2d_p(a, b, i, j)
            cr.append(dr)
            cs.append(ds)
    return np.stack(cr, axis=1), np.stack(cs, axis=1)


def d_matrices_2d(n: int, r, s, V):
    """Dr, Ds = Vr V^-1, Vs V^-1 (dg2d/reference.py:188-195)."""
    Vr, Vs = grad_vandermonde_2d(n, r, s)
    return np.linalg.solve(V.T, Vr.T).T, np.linalg.solve(V.T, Vs.T).T


def face_masks_2d(n: int, r, s) -> np.ndarray:
    """(Nfp, 3) node indices on faces s=-1, r+s=0, r=-1, ascending node index
    (dg2d/reference.py:198-210)."""
    tol = 1e-12
    f0 = np.where(np.abs(s + 1.0) < tol)[0]
    f1 = np.where(np.abs(r + s) < tol)[0]
    f2 = np.where(np.abs(r + 1.0) < tol)[0]
    return np.stack([f0, f1, f2], axis=1)


def face_mass_1d(n: int, param) -> np.ndarray:
    """inv(V1D V1D^T) on a face-trace parameter (sipg.py:98-110)."""
    V1 = vandermonde_1d(n, param)
    return np.linalg.inv(V1 @ V1.T)


def lift_2d(n: int, r, s, V, Fmask) -> np.ndarray:
    """LIFT = V V^T E (dg2d/reference.py:213-255)."""
    Np = (n + 1) * (n + 2) // 2
    Nfp = n + 1
    E = np.zeros((Np, 3 * Nfp))
    params = [r[Fmask[:, 0]], r[Fmask[:, 1]], s[Fmask[:, 2]]]
    for f in range(3):
        E[np.ix_(Fmask[:, f], np.arange(f * Nfp, (f + 1) * Nfp))] = face_mass_1d(n, params[f])
    return V @ (V.T @ E)
