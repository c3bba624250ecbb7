"""Nodal reference element on the triangle: nodes, Vandermonde, Dr/Ds,
face masks, mass matrices, and the constant tables the HIP kernels read.

Host-side setup data for the SIPG operator (float64, numpy).  Follows the
reference's ``src/ddg/dg1d/reference.py:20-143`` and
``src/ddg/dg2d/reference.py:27-255`` (Hesthaven & Warburton Codes1D/Codes2D:
JacobiP, JacobiGQ, JacobiGL, Nodes2D, Warpfactor, xytors, rstoab, Simplex2DP,
GradSimplex2DP, Vandermonde2D, Dmatrices2D) and ``src/ddg/dg2d/sipg.py:78-110``
for the volume and face mass matrices.
"""

from __future__ import annotations

import functools
import math

import numpy as np


def jacobi_p(x, alpha: float, beta: float, n: int) -> np.ndarray:
    """Orthonormal Jacobi polynomial of degree ``n`` (dg1d/reference.py:20-68)."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    lg = math.lgamma
    g0 = math.exp((alpha + beta + 1) * math.log(2.0) - math.log(alpha + beta + 1)
                  + lg(alpha + 1) + lg(beta + 1) - lg(alpha + beta + 1))
    pm = np.full(x.shape, 1.0 / math.sqrt(g0))
    if n == 0:
        return pm
    g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0
    pc = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(g1)
    aold = 2.0 / (2 + alpha + beta) * math.sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, n):
        h = 2 * i + alpha + beta
        anew = 2.0 / (h + 2) * math.sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha)
                                         * (i + 1 + beta) / (h + 1) / (h + 3))
        bnew = -(alpha * alpha - beta * beta) / h / (h + 2)
        pm, pc = pc, ((x - bnew) * pc - aold * pm) / anew
        aold = anew
    return pc


def grad_jacobi_p(x, alpha: float, beta: float, n: int) -> np.ndarray:
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    if n == 0:
        return np.zeros_like(x)
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def gauss_lobatto(n: int) -> np.ndarray:
    """Legendre-Gauss-Lobatto nodes: +-1 and the Gauss nodes of the (1,1)
    Jacobi weight, from the eigenvalues of its symmetric recurrence matrix
    (dg1d/reference.py:79-136)."""
    if n == 1:
        return np.array([-1.0, 1.0])
    m = n - 2                       # Gauss rule with m+1 points, alpha=beta=1
    if m == 0:
        return np.array([-1.0, 0.0, 1.0])
    k = np.arange(m + 1, dtype=np.float64)
    h = 2.0 * k + 2.0
    i = np.arange(1, m + 1, dtype=np.float64)
    hl = h[:m]
    off = 2.0 / (hl + 2.0) * np.sqrt(i * (i + 2.0) * (i + 1.0) * (i + 1.0) / (hl + 1.0) / (hl + 3.0))
    T = np.diag(off, 1) + np.diag(off, -1)      # zero diagonal for alpha == beta
    xi = np.linalg.eigvalsh(T)
    return np.concatenate([[-1.0], xi, [1.0]])


def vandermonde_1d(n: int, r) -> np.ndarray:
    r = np.atleast_1d(np.asarray(r, dtype=np.float64))
    return np.stack([jacobi_p(r, 0.0, 0.0, j) for j in range(n + 1)], axis=1)


_ALPHA_OPT = (0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999,
              1.2832, 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258)


def _edge_warp(n: int, rout: np.ndarray) -> np.ndarray:
    lgl = gauss_lobatto(n)
    req = np.linspace(-1.0, 1.0, n + 1)
    Lint = np.linalg.solve(vandermonde_1d(n, req).T,
                           np.stack([jacobi_p(rout, 0.0, 0.0, i) for i in range(n + 1)]))
    warp = Lint.T @ (lgl - req)
    inside = np.abs(rout) < 1.0 - 1e-10
    # divided by 1 - r^2 inside the edge, zero at the two endpoints
    return np.where(inside, warp / (1.0 - np.where(inside, rout, 0.0) ** 2), 0.0)


def nodes_rs(n: int):
    """Warp-and-blend nodes mapped to the right triangle (r, s)
    (dg2d/reference.py:62-111).  Node order: s-rows bottom to top, r
    increasing within a row."""
    Np = (n + 1) * (n + 2) // 2
    L1 = np.empty(Np)
    L3 = np.empty(Np)
    idx = 0
    for a in range(n + 1):
        for b in range(n + 1 - a):
            L1[idx] = a / n
            L3[idx] = b / n
            idx += 1
    L2 = 1.0 - L1 - L3
    x = L3 - L2
    y = (2.0 * L1 - L2 - L3) / math.sqrt(3.0)
    al = _ALPHA_OPT[n - 1] if 1 <= n <= 15 else 5.0 / 3.0
    w1 = 4.0 * L2 * L3 * _edge_warp(n, L3 - L2) * (1.0 + (al * L1) ** 2)
    w2 = 4.0 * L1 * L3 * _edge_warp(n, L1 - L3) * (1.0 + (al * L2) ** 2)
    w3 = 4.0 * L1 * L2 * _edge_warp(n, L2 - L1) * (1.0 + (al * L3) ** 2)
    c2, s2 = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
    c4, s4 = math.cos(4 * math.pi / 3), math.sin(4 * math.pi / 3)
    x = x + w1 + c2 * w2 + c4 * w3
    y = y + s2 * w2 + s4 * w3
    B1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    B2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    B3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    return B3 - B2 - B1, B1 - B2 - B3


def _modes(n: int, r, s):
    """Orthonormal Dubiner modes and their r/s derivatives, mode order
    (i outer, j inner) (dg2d/reference.py:124-185)."""
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(np.abs(s - 1.0) > 1e-12, 2.0 * (1.0 + r) / (1.0 - s) - 1.0, -1.0)
    b = s
    V, Vr, Vs = [], [], []
    half = 0.5 * (1.0 - b)
    for i in range(n + 1):
        fa, dfa = jacobi_p(a, 0, 0, i), grad_jacobi_p(a, 0, 0, i)
        for j in range(n - i + 1):
            gb = jacobi_p(b, 2 * i + 1.0, 0, j)
            dgb = grad_jacobi_p(b, 2 * i + 1.0, 0, j)
            V.append(math.sqrt(2.0) * fa * gb * (1.0 - b) ** i)
            hm1 = half ** (i - 1) if i > 0 else 1.0
            dr = dfa * gb * hm1
            ds = dfa * gb * 0.5 * (1.0 + a) * hm1
            t = dgb * half ** i - (0.5 * i * gb * hm1 if i > 0 else 0.0)
            ds = ds + fa * t
            sc = 2.0 ** (i + 0.5)
            Vr.append(sc * dr)
            Vs.append(sc * ds)
    return np.stack(V, 1), np.stack(Vr, 1), np.stack(Vs, 1)


def vandermonde_2d(n: int, r, s) -> np.ndarray:
    return _modes(n, r, s)[0]


def fmask_canonical(n: int) -> np.ndarray:
    """(Nfp, 3) face-node indices implied by the node order: face 0 is the
    first row, face 1 the last node of each row, face 2 the first node of
    each row.  Equals the coordinate test of dg2d/reference.py:198-210."""
    starts = [a * (n + 1) - a * (a - 1) // 2 for a in range(n + 1)]
    f0 = list(range(n + 1))
    f1 = [starts[a] + (n - a) for a in range(n + 1)]
    f2 = starts
    return np.array([f0, f1, f2], dtype=np.int64).T


class RefElement:
    """All degree-``N`` reference data (float64)."""

    def __init__(self, n: int):
        if n < 1:
            raise ValueError(f"polynomial order must be >= 1, got {n}")
        self.N = n
        self.Np = (n + 1) * (n + 2) // 2
        self.Nfp = n + 1
        self.r, self.s = nodes_rs(n)
        self.V, Vr, Vs = _modes(n, self.r, self.s)
        self.invV = np.linalg.inv(self.V)
        self.Dr = np.linalg.solve(self.V.T, Vr.T).T
        self.Ds = np.linalg.solve(self.V.T, Vs.T).T
        self.Fmask = fmask_canonical(n)
        # volume mass (sipg.py:78-80) and per-face 1D mass (sipg.py:83-110)
        self.M = np.linalg.inv(self.V @ self.V.T)
        params = [self.r[self.Fmask[:, 0]], self.r[self.Fmask[:, 1]], self.s[self.Fmask[:, 2]]]
        self.m1d = np.stack([np.linalg.inv(vandermonde_1d(n, p) @ vandermonde_1d(n, p).T)
                             for p in params])

    def kernel_tables(self) -> np.ndarray:
        """Flat float64 table in the layout the HIP kernels read.

        Matrices, each stored by columns (column j contiguous), every column
        zero-padded to a multiple of 8 entries (64-byte aligned):
          Dr, Ds, ArT = Dr^T M^T, AsT = Ds^T M^T, M      (Np x Np)
          L1[f][q][p] = m1d_f[p][q]                       (3 x Nfp x Nfp)
          LR[f][i][p] = sum_q Dr[Fm_f[q], i] m1d_f[p][q]  (3 x Np x Nfp)
          LS[f][i][p] = sum_q Ds[Fm_f[q], i] m1d_f[p][q]
        """
        Fm = self.Fmask
        ArT = self.Dr.T @ self.M.T
        AsT = self.Ds.T @ self.M.T
        L1 = np.transpose(self.m1d, (0, 2, 1))
        LR = np.stack([self.Dr[Fm[:, f], :].T @ self.m1d[f].T for f in range(3)])
        LS = np.stack([self.Ds[Fm[:, f], :].T @ self.m1d[f].T for f in range(3)])

        def by_columns(a):
            c = np.swapaxes(a, -1, -2)                    # (..., columns, rows)
            w = (c.shape[-1] + 7) // 8 * 8
            out = np.zeros(c.shape[:-1] + (w,))
            out[..., :c.shape[-1]] = c
            return out.reshape(-1)
        parts = [self.Dr, self.Ds, ArT, AsT, self.M, L1, LR, LS]
        return np.concatenate([by_columns(np.ascontiguousarray(p)) for p in parts])


def mfma_crow(dtype: str, g, r):
    """Node row held by lane group g, accumulator register r, in the output
    layout of v_mfma_f64_16x16x4f64 (rows g + 4r) / v_mfma_f32_16x16x4f32
    (rows 4g + r)."""
    return g + 4 * r if dtype == "f64" else 4 * g + r


def mfma_node_map(ref: "RefElement", dtype: str) -> np.ndarray:
    """(NB, 4, 4) node held by block b, lane group g, accumulator register r
    (Np marks padding); the kernels' RowMap.

    One block (N <= 4): face-major -- register f < 3 holds face f's nodes
    p = g (faces 0, 1) or p = g + 1 (face 2) for g < N, register 3 the
    interior nodes in ascending order.  Two blocks: node 16 b + crow(g, r).
    """
    Np, N, Fm = ref.Np, ref.N, ref.Fmask
    NB = (Np + 15) // 16
    out = np.full((NB, 4, 4), Np, dtype=np.int64)
    if NB == 1:
        interior = [n for n in range(Np) if n not in set(Fm.reshape(-1).tolist())]
        for g in range(4):
            for f in range(3):
                if g < N:
                    out[0, g, f] = Fm[g if f < 2 else g + 1, f]
            if g < len(interior):
                out[0, g, 3] = interior[g]
    else:
        for b in range(NB):
            for g in range(4):
                for r in range(4):
                    out[b, g, r] = 16 * b + mfma_crow(dtype, g, r)
    return out


def mfma_face_ksteps(Np: int, Nfp: int) -> int:
    """k-steps of the face lift: 12 with the face rows formed in registers
    (one 16-row block), else the 9 Nfp rows four per k-step."""
    return 12 if (Np + 15) // 16 == 1 else (9 * Nfp + 3) // 4


def mfma_face_row(N: int, Np: int, t, g):
    """Face row (index into [P | aX | bX] x 3 Nfp face nodes, 9 Nfp meaning
    padding) that lane group g supplies at k-step t of the face lift; the
    kernels' face_rows_inreg for one block, 4 t + g for two.

    One block: slot s = t // 3, quantity q = t % 3.  Slots 0-2 are face s
    at node p = g (p = g + 1 on face 2) for g < N; slot 3 is one vertex's
    second face: g = 0 face 2 p = 0, g = 1 face 0 p = N, g = 2 face 1 p = N.
    """
    Nfp = N + 1
    NF = 9 * Nfp
    t = np.asarray(t)
    g = np.asarray(g)
    if (Np + 15) // 16 != 1:
        r = 4 * t + g
        return np.where(r < NF, r, NF)
    s, q = t // 3, t % 3
    f = np.where(s < 3, s, np.choose(np.minimum(g, 3), [2, 0, 1, 0]))
    p = np.where(s < 3, np.where(s < 2, g, g + 1), np.where((g == 1) | (g == 2), N, 0))
    on = np.where(s < 3, g < N, g < 3)
    return np.where(on, q * 3 * Nfp + f * Nfp + p, NF)


def mfma_fragments(ref: "RefElement", dtype: str) -> np.ndarray:
    """A-operand fragments of the matrix-core apply, one 64-lane fragment per
    (output block, k-step), in the order GEMM1 | GEMM2 volume | face | mass.

    Lane l = 16 g + m supplies A[row(b, m)][col(t, g)]: output row m of
    block b lands in the accumulator position (g', r') with crow(g', r') = m,
    so row(b, m) is the node mfma_node_map puts there; for products whose B
    operand is a nodal vector held in accumulator layout, col(t, g) is the
    node of block t // 4, lane group g, register t % 4; for the face rows it
    is mfma_face_row(t, g).  Entries outside the matrix (padding) are zero.
    """
    Np, Nfp, Fm = ref.Np, ref.Nfp, ref.Fmask
    NB = (Np + 15) // 16
    KB = 4 * NB
    NF = 9 * Nfp
    FK = mfma_face_ksteps(Np, Nfp)
    lanes = np.arange(64)
    m, g = lanes % 16, lanes // 16
    nmap = mfma_node_map(ref, dtype)
    # output row m of a block -> (lane group, register) of the instruction
    gi = m % 4 if dtype == "f64" else m // 4
    ri = m // 4 if dtype == "f64" else m % 4

    def row_node(b):
        return nmap[b, gi, ri]

    def frag_nodal(Mat, b, t):
        row = row_node(b)
        col = nmap[t // 4, g, t % 4]
        ok = (row < Mat.shape[0]) & (col < Mat.shape[1])
        return np.where(ok, Mat[np.minimum(row, Mat.shape[0] - 1), np.minimum(col, Mat.shape[1] - 1)], 0.0)

    ArT = ref.Dr.T @ ref.M.T
    AsT = ref.Ds.T @ ref.M.T
    G = np.zeros((Np, NF))
    for f in range(3):
        for q in range(Nfp):
            G[Fm[q, f], f * Nfp:(f + 1) * Nfp] = ref.m1d[f][:, q]
        LR = ref.Dr[Fm[:, f], :].T @ ref.m1d[f].T
        LS = ref.Ds[Fm[:, f], :].T @ ref.m1d[f].T
        G[:, 3 * Nfp + f * Nfp:3 * Nfp + (f + 1) * Nfp] = LR
        G[:, 6 * Nfp + f * Nfp:6 * Nfp + (f + 1) * Nfp] = LS
    out = []
    for b in range(2 * NB):                              # GEMM1: Dr blocks then Ds blocks
        R = ref.Dr if b < NB else ref.Ds
        for t in range(KB):
            out.append(frag_nodal(R, b % NB, t))
    for b in range(NB):                                  # GEMM2 volume: ArT then AsT
        for t in range(2 * KB):
            out.append(frag_nodal(ArT if t < KB else AsT, b, t % KB))
    for b in range(NB):                                  # face lift
        for t in range(FK):
            row, col = row_node(b), mfma_face_row(ref.N, Np, t, g)
            ok = (row < Np) & (col < NF)
            out.append(np.where(ok, G[np.minimum(row, Np - 1), np.minimum(col, NF - 1)], 0.0))
    for b in range(NB):                                  # mass
        for t in range(KB):
            out.append(frag_nodal(ref.M, b, t))
    return np.concatenate(out)


@functools.lru_cache(maxsize=None)
def ref_element(n: int) -> RefElement:
    return RefElement(n)


def p_interp_matrix(n_lo: int, n_hi: int) -> np.ndarray:
    """(Np_hi, Np_lo) nodal interpolation between nested degrees
    (multigrid.py:106-114)."""
    lo, hi = ref_element(n_lo), ref_element(n_hi)
    return vandermonde_2d(n_lo, hi.r, hi.s) @ lo.invV


def h_child_interp_matrices(n: int) -> np.ndarray:
    """(4, Np, Np): a parent element's nodal values -> the nodal values of its
    four red children (multigrid.py:137-163).  Child c has vertices
    (A, mAB, mCA), (mAB, B, mBC), (mCA, mBC, C), (mAB, mBC, mCA) of the
    reference triangle A=(-1,-1), B=(1,-1), C=(-1,1) -- the child order of the
    refinement forest (refinement.py:359-382)."""
    ref = ref_element(n)
    A, B, Cv = np.array([-1.0, -1.0]), np.array([1.0, -1.0]), np.array([-1.0, 1.0])
    mAB, mBC, mCA = (A + B) / 2, (B + Cv) / 2, (Cv + A) / 2
    kids = ((A, mAB, mCA), (mAB, B, mBC), (mCA, mBC, Cv), (mAB, mBC, mCA))
    # barycentric weights of the child's own nodes
    l0, l1, l2 = -(ref.r + ref.s) / 2, (1 + ref.r) / 2, (1 + ref.s) / 2
    out = np.empty((4, ref.Np, ref.Np))
    for c, (P0, P1, P2) in enumerate(kids):
        pr = l0 * P0[0] + l1 * P1[0] + l2 * P2[0]
        ps = l0 * P0[1] + l1 * P1[1] + l2 * P2[1]
        out[c] = vandermonde_2d(n, pr, ps) @ ref.invV
    return out
