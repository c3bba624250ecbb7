// SIPG apply at order 1 on the vector ALUs, one thread per element:
//   y = (A_sipg + alpha*J*M + shift*I) u          (or y = b - that, residual mode)
// with the same closed form as the matrix-core kernel (src/ddg/dg2d/sipg.py
// :142-337):
//
//   volume : y_k = ArT (vrr ur + vrs us) + AsT (vrs ur + vss us)
//   face f : P = pen_f [u] - sJw_f (dn^- + dn^+)    -> y[Fmask_f] += m1d_f^T P
//            X = -sJw_f [u]                          -> y += (a_f LR_f + b_f LS_f) X
//
// At N = 1 a nodal vector has 3 entries, so the 16-row matrix-core tiles
// would be mostly padding; here every element's contractions are
// unrolled scalar FMAs with the tables read through the scalar cache
// (uniform addresses).  A workgroup is one wave and one 64-element tile
// (same tiles, XCD mapping and dot partials as the matrix-core kernel).
// A thread reads its element's values, record and face codes, then its three
// neighbours' values (cache hits mostly: tiles are compact patches of the
// mesh), each batch issued before its first wait; the results go through LDS
// so the tile's rows are stored contiguously.

#pragma once

#include "ddgx_internal.h"

namespace ddgx {

// Orders the apply runs on the vector ALUs instead of the matrix cores: the
// faster of the two kernels at 2x256x256 on MI355X.  N=1 takes 10 us against
// about 20 us in float64; at N=2 the matrix cores win (float64 21 against
// 25 us; float32 residual mode 13.3 against 15.0 us).
template <typename T, int N>
struct UseValuApply {
  static constexpr bool value = N == 1;
};

template <typename T, int N, int MODE>
__global__ void __launch_bounds__(64)
sipg_apply_valu_kernel(const T* __restrict__ tab, const T* __restrict__ rec,
                       const int* __restrict__ nbrp, const int* __restrict__ perm,
                       const T* __restrict__ u, T* __restrict__ y, const T* __restrict__ udir,
                       T alpha, T shift, int K, double* __restrict__ partial,
                       const int* __restrict__ skip, const T* __restrict__ bvec,
                       const T* __restrict__ binv, T omega, const T* __restrict__ pxc,
                       const T* __restrict__ pmat, T* __restrict__ xco, int npt) {
  using S = Shape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, RS = S::RS, RSF = S::RSF;
  constexpr bool HELM = MODE & M_HELM, UDIR = MODE & M_UDIR, DOT = MODE & M_DOT;
  constexpr bool RESID = MODE & M_RESID;
  // smoothing sweep: y = u + omega Binv (bvec - A u), element by element
  constexpr bool SMOOTH = MODE & M_SMOOTH;
  // the input is u + the prolongation of the coarser level's pxc (element
  // 4k+c is child c of coarse element k; pmat holds the 4 child matrices,
  // rows c*Np + i), formed on the fly instead of stored
  constexpr bool PROLONG = MODE & M_PROLONG;
  // with RESID: the residual is restricted to the coarser h-level instead of
  // stored -- y receives the coarse right-hand side P^T (b - A u) of every
  // coarse element whose four children 4k..4k+3 the tile holds (pmat: the
  // 4 Np x Np child matrices), and xco, when given, the coarse level's first
  // sweep omega Binv_c (that) with binv the coarse level's packed blocks.
  // Identity tile order only (the children of a coarse element are then
  // consecutive tile positions of one tile).
  constexpr bool RESTRICT = MODE & M_RESTRICT;
  static_assert(!RESTRICT || (RESID && !SMOOTH && !DOT && !PROLONG), "restriction of a residual");
  // with SMOOTH: the sweep's result is also prolongated into the next-finer
  // p-level's iterate, xco[k] += pmat x_new[k] (pmat: npt x Np)
  constexpr bool PROLONG_OUT = MODE & M_PROLONG_OUT;
  static_assert(!PROLONG_OUT || (SMOOTH && !PROLONG), "prolongation of a sweep's result");
  auto uval = [&](size_t kid, int j) -> T {
    T v = u[kid * Np + j];
    if constexpr (PROLONG) {
      const size_t kc = kid >> 2;
      const int c = (int)(kid & 3);
      T acc = 0;
#pragma unroll
      for (int l = 0; l < Np; ++l) acc += pmat[(c * Np + j) * Np + l] * pxc[kc * Np + l];
      v += acc;
    }
    return v;
  };
  const T* __restrict__ Dr = tab + S::oDr;
  const T* __restrict__ Ds = tab + S::oDs;
  const T* __restrict__ Ar = tab + S::oAr;
  const T* __restrict__ As = tab + S::oAs;
  const T* __restrict__ Mm = tab + S::oM;
  const T* __restrict__ L1 = tab + S::oL1;
  const T* __restrict__ LR = tab + S::oLR;
  const T* __restrict__ LS = tab + S::oLS;
  __shared__ T s_u[TILE * Np];

  if (skip != nullptr && *skip) return;

  const int el = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int k0 = tile * TILE;
  const int ne = min(TILE, K - k0);
  const int k = k0 + el;
  const bool valid = el < ne;

  // Loads in two batches, each issued before its first wait: this element's
  // values, record and face codes (and the smoothing or residual data), then,
  // once the codes are in, the three neighbours' values and face
  // coefficients.  Addresses are clamped (a lane past the tile reads tile
  // position k0, a boundary face its own element), so no load sits in a
  // lane-divergent branch; a neighbour in the tile is read from memory like
  // any other (a cache hit).
  const int kc = valid ? k : k0;  // clamped tile position
  const size_t kid = (size_t)(perm ? perm[kc] : kc);
  T uk[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) uk[j] = uval(kid, j);
  T rc[DDGX_REC_WORDS];
#pragma unroll
  for (int w = 0; w < DDGX_REC_WORDS; ++w) rc[w] = rec[(size_t)w * K + kc];
  int cf[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) cf[f] = nbrp[(size_t)f * K + kc];
  constexpr int NPK = SymPack<N>::NPK;
  T bvv[Np], bkk[NPK];  // used (and kept) by the residual and smoothing modes only
  if constexpr (SMOOTH || RESID) {
#pragma unroll
    for (int i = 0; i < Np; ++i) bvv[i] = bvec[kid * Np + i];
  }
  if constexpr (SMOOTH) {
#pragma unroll
    for (int q = 0; q < NPK; ++q) bkk[q] = binv[kid * NPK + q];
  }
  // the neighbours: tile positions, element ids (all three before any of
  // their loads, so no load waits on another's branch), then values and
  // face coefficients
  T unb[3][Np], aPf[3], bPf[3];
  int npf[3], fnf[3];
  size_t ngf[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    npf[f] = cf[f] >= 0 ? nbr_elem(cf[f]) : kc;
    fnf[f] = cf[f] >= 0 ? nbr_face(cf[f]) : 0;
    ngf[f] = (size_t)npf[f];
  }
  if (perm != nullptr) {  // uniform branch
#pragma unroll
    for (int f = 0; f < 3; ++f) ngf[f] = (size_t)perm[npf[f]];
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
#pragma unroll
    for (int j = 0; j < Np; ++j) unb[f][j] = uval(ngf[f], j);
    aPf[f] = rec[(size_t)(R_A + fnf[f]) * K + npf[f]];
    bPf[f] = rec[(size_t)(R_B + fnf[f]) * K + npf[f]];
  }

  double dacc = 0.0;
  T v[Np];
  if (valid) {
    T ur[Np], us[Np];
#pragma unroll
    for (int n = 0; n < Np; ++n) {
      T a = 0, b = 0;
#pragma unroll
      for (int j = 0; j < Np; ++j) {
        a += Dr[j * RS + n] * uk[j];
        b += Ds[j * RS + n] * uk[j];
      }
      ur[n] = a;
      us[n] = b;
    }
    // volume
    {
      T wr[Np], ws[Np];
#pragma unroll
      for (int c = 0; c < Np; ++c) {
        wr[c] = rc[R_VRR] * ur[c] + rc[R_VRS] * us[c];
        ws[c] = rc[R_VRS] * ur[c] + rc[R_VSS] * us[c];
      }
#pragma unroll
      for (int i = 0; i < Np; ++i) {
        T acc = 0;
#pragma unroll
        for (int c = 0; c < Np; ++c) acc += Ar[c * RS + i] * wr[c] + As[c * RS + i] * ws[c];
        v[i] = acc;
      }
    }
    if constexpr (HELM) {
      const T aj = alpha * rc[R_J];
#pragma unroll
      for (int i = 0; i < Np; ++i) {
        T m = 0;
#pragma unroll
        for (int j = 0; j < Np; ++j) m += Mm[j * RS + i] * uk[j];
        v[i] += aj * m + shift * uk[i];
      }
    }
    // faces
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const int c = cf[f];
      const T aM = rc[R_A + f], bM = rc[R_B + f];
      const T sjw = rc[R_SJW + f], pen = rc[R_PEN + f];
      T uP[Nfp], dnP[Nfp];
      if (c >= 0) {
        // interior: the neighbour's trace and normal derivative (its own
        // face coefficients, so the sign flips to the own normal)
        const int fn = nbr_face(c);
        const bool rv = nbr_rev(c);
        const T* un = unb[f];
        const T aP = aPf[f], bP = bPf[f];
        T urn[Np], usn[Np];
#pragma unroll
        for (int n = 0; n < Np; ++n) {
          T a = 0, b = 0;
#pragma unroll
          for (int j = 0; j < Np; ++j) {
            a += Dr[j * RS + n] * un[j];
            b += Ds[j * RS + n] * un[j];
          }
          urn[n] = a;
          usn[n] = b;
        }
        // the neighbour's face node is a run-time index: pick it with a
        // compare chain over the registers (no indexed register access)
#pragma unroll
        for (int p = 0; p < Nfp; ++p) {
          const int idx = fmask<N>(fn, rv ? Nfp - 1 - p : p);
          T uu = un[0], dr = urn[0], ds = usn[0];
#pragma unroll
          for (int m = 1; m < Np; ++m)
            if (idx == m) {
              uu = un[m];
              dr = urn[m];
              ds = usn[m];
            }
          uP[p] = uu;
          dnP[p] = -(aP * dr + bP * ds);
        }
      } else if (c == CODE_DIRICHLET) {
#pragma unroll
        for (int p = 0; p < Nfp; ++p) {
          if constexpr (UDIR)
            uP[p] = udir[kid * Np + fmask<N>(f, p)];
          else
            uP[p] = T(0);
          dnP[p] = T(0);
        }
      } else {
#pragma unroll
        for (int p = 0; p < Nfp; ++p) {
          uP[p] = uk[fmask<N>(f, p)];
          dnP[p] = T(0);
        }
      }
      T P[Nfp], Xa[Nfp], Xb[Nfp];
#pragma unroll
      for (int p = 0; p < Nfp; ++p) {
        const int n = fmask<N>(f, p);
        const T ju = uk[n] - uP[p];
        const T dnM = aM * ur[n] + bM * us[n];
        const T X = -sjw * ju;
        P[p] = pen * ju - sjw * (dnM + dnP[p]);
        Xa[p] = aM * X;
        Xb[p] = bM * X;
      }
#pragma unroll
      for (int q = 0; q < Nfp; ++q) {
        T acc = 0;
#pragma unroll
        for (int p = 0; p < Nfp; ++p) acc += L1[(f * Nfp + p) * RSF + q] * P[p];
        v[fmask<N>(f, q)] += acc;
      }
#pragma unroll
      for (int i = 0; i < Np; ++i) {
        T acc = 0;
#pragma unroll
        for (int p = 0; p < Nfp; ++p)
          acc += LR[(f * Nfp + p) * RS + i] * Xa[p] + LS[(f * Nfp + p) * RS + i] * Xb[p];
        v[i] += acc;
      }
    }
    if constexpr (DOT) {
#pragma unroll
      for (int i = 0; i < Np; ++i) dacc += (double)uk[i] * (double)v[i];
    }
    if constexpr (SMOOTH) {
      const size_t kg = kid;
      T r[Np];
#pragma unroll
      for (int i = 0; i < Np; ++i) r[i] = bvv[i] - v[i];
#pragma unroll
      for (int i = 0; i < Np; ++i) {
        T acc = 0;
#pragma unroll
        for (int j = 0; j < Np; ++j) acc += bkk[diag_pk(Np, i, j)] * r[j];
        v[i] = uk[i] + omega * acc;
      }
      if constexpr (PROLONG_OUT) {
        for (int i = 0; i < npt; ++i) {
          T acc = 0;
#pragma unroll
          for (int l = 0; l < Np; ++l) acc += pmat[i * Np + l] * v[l];
          xco[kg * npt + i] += acc;
        }
      }
    }
  }
  // results (the residual b - A u in the residual modes) through LDS, so the
  // tile's rows are stored contiguously
  if (valid) {
#pragma unroll
    for (int i = 0; i < Np; ++i) s_u[el * Np + i] = (RESID && !SMOOTH) ? bvv[i] - v[i] : v[i];
  }
  __syncthreads();
  if constexpr (RESTRICT) {
    // one thread per (coarse element, node) restricts the tile's residual
    const int cl = el / Np, j = el % Np;
    const bool on = cl < ne / 4;
    const size_t kcs = (size_t)(k0 / 4 + cl);
    T lo = 0;
    if (on) {
#pragma unroll
      for (int h = 0; h < 4 * Np; ++h) lo += pmat[h * Np + j] * s_u[4 * cl * Np + h];
      y[kcs * Np + j] = lo;
    }
    if (xco != nullptr) {
      __syncthreads();  // every thread has read the residual rows
      if (on) s_u[el] = lo;
      __syncthreads();
      if (on) {
        T a = 0;
#pragma unroll
        for (int i = 0; i < Np; ++i) a += binv[kcs * NPK + diag_pk(Np, j, i)] * s_u[cl * Np + i];
        xco[kcs * Np + j] = omega * a;
      }
    }
    return;
  }
  if (perm == nullptr) {
    T* dst = y + (size_t)k0 * Np;
    for (int i = el; i < ne * Np; i += TILE) dst[i] = s_u[i];
  } else {
    for (int i = el; i < ne * Np; i += TILE) y[(size_t)perm[k0 + i / Np] * Np + i % Np] = s_u[i];
  }
  if constexpr (DOT) {
    for (int off = 32; off > 0; off >>= 1) dacc += __shfl_down(dacc, off, 64);
    if (el == 0) partial[tile] = dacc;  // one partial per tile, at the tile's index
  }
}

}  // namespace ddgx
