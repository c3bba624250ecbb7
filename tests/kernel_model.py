"""Numpy model of the arithmetic of the HIP apply kernel (test helper).

It consumes exactly what the kernel consumes -- the packed per-element
records, face codes and reference tables from ``ddg_amd.operator.pack`` and
``RefElement.kernel_tables`` -- and evaluates the same closed form, so the
packing and the formula can be checked against the oracle on a machine
without a GPU.  It is not used by the product.
"""

from __future__ import annotations

import numpy as np

from ddg_amd.reference import mfma_face_ksteps, mfma_face_row


def unpack_tables(tab, N):
    """Inverse of RefElement.kernel_tables (matrices stored by columns,
    columns padded to multiples of 8)."""
    Np, Nfp = (N + 1) * (N + 2) // 2, N + 1
    o = 0

    def take(shape):
        nonlocal o
        rows, cols = shape[-2], shape[-1]
        w = (rows + 7) // 8 * 8
        lead = shape[:-2]
        n = int(np.prod(lead, dtype=np.int64)) * cols * w
        c = tab[o:o + n].reshape(lead + (cols, w))[..., :rows]
        o += n
        return np.swapaxes(c, -1, -2)
    Dr = take((Np, Np))
    Ds = take((Np, Np))
    ArT = take((Np, Np))
    AsT = take((Np, Np))
    M = take((Np, Np))
    L1 = take((3, Nfp, Nfp))
    LR = take((3, Np, Nfp))
    LS = take((3, Np, Nfp))
    assert o == tab.size
    return Dr, Ds, ArT, AsT, M, L1, LR, LS


def fmask(N):
    starts = [a * (N + 1) - a * (a - 1) // 2 for a in range(N + 1)]
    return np.array([list(range(N + 1)), [starts[a] + N - a for a in range(N + 1)], starts])


def model_apply(N, tab, rec, nbr, u, *, udir=None, alpha=0.0, shift=0.0, probe=False):
    """u: (K, Np) element-major.  Returns (K, Np)."""
    Dr, Ds, ArT, AsT, M, L1, LR, LS = unpack_tables(tab, N)
    K, Np = u.shape
    Nfp = N + 1
    Fm = fmask(N)
    ur = u @ Dr.T
    us = u @ Ds.T
    wr = rec[:, [0]] * ur + rec[:, [1]] * us
    ws = rec[:, [1]] * ur + rec[:, [2]] * us
    y = wr @ ArT.T + ws @ AsT.T
    for f in range(3):
        uM = u[:, Fm[f]]
        dnM = rec[:, [4 + f]] * ur[:, Fm[f]] + rec[:, [7 + f]] * us[:, Fm[f]]
        code = nbr[:, f]
        uP = np.zeros((K, Nfp))
        dnP = np.zeros((K, Nfp))
        inter = code >= 0
        if not probe:
            kn = code[inter] >> 3
            fn = (code[inter] >> 1) & 3
            rv = code[inter] & 1
            p = np.arange(Nfp)
            pn = np.where(rv[:, None] == 1, Nfp - 1 - p[None, :], p[None, :])
            nn = Fm[fn[:, None], pn]
            uP[inter] = u[kn[:, None], nn]
            urn = ur[kn[:, None], nn]
            usn = us[kn[:, None], nn]
            dnP[inter] = -(rec[kn, 4 + fn][:, None] * urn + rec[kn, 7 + fn][:, None] * usn)
        dirf = code == -1
        if udir is not None:
            uP[dirf] = udir[dirf][:, Fm[f]]
        none = code == -2
        uP[none] = uM[none]
        ju = uM - uP
        P = rec[:, [13 + f]] * ju - rec[:, [10 + f]] * (dnM + dnP)
        X = -rec[:, [10 + f]] * ju
        y += (rec[:, [4 + f]] * (X @ LR[f].T) + rec[:, [7 + f]] * (X @ LS[f].T))
        y[:, Fm[f]] += P @ L1[f].T
    if alpha != 0.0 or shift != 0.0:
        y += alpha * rec[:, [3]] * (u @ M.T) + shift * u
    return y


# ---------------------------------------------------------------------------
# emulation of the matrix-core (MFMA 16x16x4) apply, lane by lane
# ---------------------------------------------------------------------------
def _crow(dtype, g, r):
    return g + 4 * r if dtype == "f64" else 4 * g + r


def mma(a, b, c, dtype):
    """One 16x16x4 product: a, b are (64,) lane values, c is (64, 4)."""
    lanes = np.arange(64)
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    A[lanes % 16, lanes // 16] = a
    B[lanes // 16, lanes % 16] = b
    D = A @ B
    out = c.copy()
    for r in range(4):
        out[:, r] += D[_crow(dtype, lanes // 16, r), lanes % 16]
    return out


def model_apply_mfma(N, tab, frag, rec, nbr, u, dtype="f64", node_map=None):
    """Apply with the MFMA data flow for u (K, Np); faces via the closed form
    (as model_apply), contractions via mma() on the fragments."""
    K, Np = u.shape
    Nfp = N + 1
    NB = (Np + 15) // 16
    KB = 4 * NB
    NF = 9 * Nfp
    FK = mfma_face_ksteps(Np, Nfp)
    nfr = [2 * NB * KB, NB * 2 * KB, NB * FK, NB * KB]
    offs = np.cumsum([0] + nfr)
    F = frag.reshape(-1, 64)
    Dr, Ds, ArT, AsT, M, L1, LR, LS = unpack_tables(tab, N)
    Fm = fmask(N)
    lanes = np.arange(64)
    g, e = lanes // 16, lanes % 16
    y = np.zeros_like(u)
    # face rows from the exact model (P, aX, bX per face node)
    ur_all, us_all = u @ Dr.T, u @ Ds.T
    Fr = np.zeros((K, NF + 1))  # the last column is the padding row (zero)
    for f in range(3):
        dnM = rec[:, [4 + f]] * ur_all[:, Fm[f]] + rec[:, [7 + f]] * us_all[:, Fm[f]]
        code = nbr[:, f]
        uM = u[:, Fm[f]]
        uP = uM.copy()
        dnP = np.zeros_like(uM)
        inter = code >= 0
        kn, fn, rv = code[inter] >> 3, (code[inter] >> 1) & 3, code[inter] & 1
        p = np.arange(Nfp)
        pn = np.where(rv[:, None] == 1, Nfp - 1 - p[None, :], p[None, :])
        nn = Fm[fn[:, None], pn]
        uP[inter] = u[kn[:, None], nn]
        dnP[inter] = -(rec[kn, 4 + fn][:, None] * ur_all[kn[:, None], nn]
                       + rec[kn, 7 + fn][:, None] * us_all[kn[:, None], nn])
        uP[code == -1] = 0.0
        ju = uM - uP
        X = -rec[:, [10 + f]] * ju
        Fr[:, f * Nfp:(f + 1) * Nfp] = rec[:, [13 + f]] * ju - rec[:, [10 + f]] * (dnM + dnP)
        Fr[:, 3 * Nfp + f * Nfp:3 * Nfp + (f + 1) * Nfp] = rec[:, [4 + f]] * X
        Fr[:, 6 * Nfp + f * Nfp:6 * Nfp + (f + 1) * Nfp] = rec[:, [7 + f]] * X
    for s0 in range(0, K, 16):
        ek = np.minimum(s0 + e, K - 1)
        valid = (s0 + e) < K

        def node(b, r):
            return node_map[b, g, r] if node_map is not None else 16 * b + _crow(dtype, g, r)

        def nodal(vec_rows, b, r):
            n = node(b, r)
            return np.where(valid & (n < Np), vec_rows[ek, np.minimum(n, Np - 1)], 0.0)
        U = [[nodal(u, b, r) for r in range(4)] for b in range(NB)]
        URS = []
        for b in range(2 * NB):
            acc = np.zeros((64, 4))
            for t in range(KB):
                acc = mma(F[offs[0] + b * KB + t], U[t // 4][t % 4], acc, dtype)
            URS.append(acc)
        UR = [[URS[b][:, r] for r in range(4)] for b in range(NB)]
        US = [[URS[NB + b][:, r] for r in range(4)] for b in range(NB)]
        vrr, vrs, vss = rec[ek, 0], rec[ek, 1], rec[ek, 2]
        WR = [[vrr * UR[b][r] + vrs * US[b][r] for r in range(4)] for b in range(NB)]
        WS = [[vrs * UR[b][r] + vss * US[b][r] for r in range(4)] for b in range(NB)]
        for bo in range(NB):
            acc = np.zeros((64, 4))
            for t in range(KB):
                acc = mma(F[offs[1] + bo * 2 * KB + t], WR[t // 4][t % 4], acc, dtype)
            for t in range(KB):
                acc = mma(F[offs[1] + bo * 2 * KB + KB + t], WS[t // 4][t % 4], acc, dtype)
            for t in range(FK):
                acc = mma(F[offs[2] + bo * FK + t], Fr[ek, mfma_face_row(N, Np, t, g)], acc, dtype)
            for r in range(4):
                n = node(bo, r)
                ok = valid & (n < Np)
                y[ek[ok], n[ok]] = acc[ok, r]
    return y
