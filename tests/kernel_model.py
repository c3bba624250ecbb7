"""Numpy model of the arithmetic of the HIP apply kernel (test helper).

It consumes exactly what the kernel consumes -- the packed per-element
records, face codes and reference tables from ``ddg_amd.operator.pack`` and
``RefElement.kernel_tables`` -- and evaluates the same closed form, so the
packing and the formula can be checked against the oracle on a machine
without a GPU.  It is not used by the product.
"""

from __future__ import annotations

import numpy as np


def unpack_tables(tab, N):
    Np, Nfp = (N + 1) * (N + 2) // 2, N + 1
    o = 0

    def take(n, shape):
        nonlocal o
        a = tab[o:o + n].reshape(shape)
        o += n
        return a
    Dr = take(Np * Np, (Np, Np))
    Ds = take(Np * Np, (Np, Np))
    ArT = take(Np * Np, (Np, Np))
    AsT = take(Np * Np, (Np, Np))
    M = take(Np * Np, (Np, Np))
    L1 = take(3 * Nfp * Nfp, (3, Nfp, Nfp))
    LR = take(3 * Np * Nfp, (3, Np, Nfp))
    LS = take(3 * Np * Nfp, (3, Np, Nfp))
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
