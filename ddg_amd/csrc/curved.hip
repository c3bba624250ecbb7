// SIPG apply on curved (non-affine) elements, for gfx950:
//   y = (A_sipg + alpha M + shift I) u       (or b - that: residual mode)
// with the reference's two quadratures (src/ddg/dg2d/sipg.py:142-337):
//   nodal    (cub=None): vol = sum J (ux (M vx) + uy (M vy)) with per-node J
//                        and metrics, faces  sum_side w sJ_0 X^T m1d Y;
//   cubature (cub):      vol = sum_q w_q J_q (ux_q vx_q + uy_q vy_q), faces
//                        sum_side w X^T (L1d^T diag(w_q sJ_q) L1d) Y
//                        (sipg.py:190-201, 281-299).
// In both, the face traces are nodal: u and grad u at the face nodes with
// the per-node metrics, normal derivatives with the side's own nodal
// normal; the penalty is tau_scale (N+1)^2 Fscale at face node 0.
//
// The apply is the reverse-mode gradient d a(u, v) / dv, in three launches:
//   1. curved_trace_kernel: per face node, u, du/dx, du/dy       -> trace
//   2. curved_side_kernel:  per element side, jump and average with the
//      side's normal and face mass, then the cotangents of v, dv/dx, dv/dy
//      at its own face nodes (cown) and at the matched neighbour nodes (cnbr)
//   3. curved_elem_kernel:  per element, gathers its cown and the cnbr of
//      the sides facing it, pulls dv/dx, dv/dy back through the per-node
//      metrics and Dr^T, Ds^T, adds the nodal or cubature volume term and
//      the mass.
// Every element owns its output rows, so the result is deterministic.
// Layouts: element-major, element k's data contiguous (see struct Curved).

#include <string>
#include <vector>

#include "ddgx_internal.h"

namespace ddgx {

template <int N>
struct CShape {
  static constexpr int Np = (N + 1) * (N + 2) / 2;
  static constexpr int Nfp = N + 1;
};

template <typename T, int N>
__global__ void __launch_bounds__(256)
curved_trace_kernel(const T* __restrict__ tab, const T* __restrict__ geo,
                    const T* __restrict__ u, T* __restrict__ tr, int64_t K,
                    const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  constexpr int Np = CShape<N>::Np, Nfp = CShape<N>::Nfp;
  // Dr and Ds in LDS for the workgroup's lifetime (each thread reads a row)
  __shared__ T s_d[2 * Np * Np];
  for (int i = threadIdx.x; i < 2 * Np * Np; i += 256) s_d[i] = tab[i];
  __syncthreads();
  const int64_t nf = K * 3 * Nfp;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nf; t += stride) {
    const int64_t k = t / (3 * Nfp);
    const int fp = (int)(t - k * 3 * Nfp);
    const int i = fmask<N>(fp / Nfp, fp % Nfp);
    T ui = 0, ur = 0, us = 0;
    if (u != nullptr) {
      const T* uk = u + k * Np;
      const T* dr = s_d + i * Np;
      const T* ds = s_d + Np * Np + i * Np;
#pragma unroll
      for (int j = 0; j < Np; ++j) {
        const T uj = uk[j];
        ur += dr[j] * uj;
        us += ds[j] * uj;
      }
      ui = uk[i];
    }
    const T* g = geo + k * 5 * Np;
    tr[t] = ui;
    tr[nf + t] = g[i] * ur + g[Np + i] * us;
    tr[2 * nf + t] = g[2 * Np + i] * ur + g[3 * Np + i] * us;
  }
}

// One thread per (element side, face node): the side's jump and normal-
// derivative average at its node, exchanged through LDS, then row p of the
// side's face-mass products.  SPB sides per 256-thread block.
template <typename T, int N>
__global__ void __launch_bounds__(256)
curved_side_kernel(const T* __restrict__ tab, const T* __restrict__ fnorm,
                   const T* __restrict__ fpar, const T* __restrict__ fmass,
                   const int* __restrict__ codes, const T* __restrict__ tr,
                   const T* __restrict__ udir, T* __restrict__ cown, T* __restrict__ cnbr,
                   int64_t K, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  constexpr int Np = CShape<N>::Np, Nfp = CShape<N>::Nfp;
  constexpr int SPB = 256 / Nfp;
  __shared__ T s_ju[SPB * Nfp], s_av[SPB * Nfp];
  const int64_t nf = K * 3 * Nfp;
  const int64_t nside = K * 3;
  const T* m1d = tab + 3 * Np * Np;
  const int ls = threadIdx.x / Nfp, p = threadIdx.x - (threadIdx.x / Nfp) * Nfp;
  for (int64_t s0 = (int64_t)blockIdx.x * SPB; s0 < nside; s0 += (int64_t)gridDim.x * SPB) {
    const int64_t t0 = s0 + ls;
    const bool act = ls < SPB && t0 < nside;
    // every load of the side from clamped indices, in two batches (the
    // side's own data, then the neighbour's trace at the node its code
    // names), each issued before its first use
    const int64_t t = act ? t0 : 0;
    const int64_t k = t / 3;
    const int f = (int)(t - k * 3);
    const int64_t at = t * Nfp + p;
    const int c = codes[t];
    const T w = fpar[t * 3], wtau = fpar[t * 3 + 1], sj0 = fpar[t * 3 + 2];
    const T nx = fnorm[at * 2], ny = fnorm[at * 2 + 1];
    const T uM = tr[at];
    const T dM = nx * tr[nf + at] + ny * tr[2 * nf + at];
    T m[Nfp];  // the side's face-mass column p
#pragma unroll
    for (int q = 0; q < Nfp; ++q)
      m[q] = fmass != nullptr ? fmass[t * Nfp * Nfp + q * Nfp + p]
                              : sj0 * m1d[f * Nfp * Nfp + q * Nfp + p];
    const int64_t o = c >= 0 ? ((int64_t)nbr_elem(c) * 3 + nbr_face(c)) * Nfp +
                                   (nbr_rev(c) ? Nfp - 1 - p : p)
                             : at;
    const T uN = tr[o];
    const T dN = nx * tr[nf + o] + ny * tr[2 * nf + o];
    // Dirichlet: ghost u_dir (or 0), ghost gradient = own
    const T uD = udir != nullptr ? udir[k * Np + fmask<N>(f, p)] : T(0);
    const bool on = act && c != CODE_NONE && w != T(0);
    const T uP = c >= 0 ? uN : uD;
    const T dP = c >= 0 ? dN : dM;
    const T ju = on ? uM - uP : T(0);
    const T av = on ? T(0.5) * (dM + dP) : T(0);
    __syncthreads();  // the previous group's LDS reads are done
    if (act) {
      s_ju[ls * Nfp + p] = ju;
      s_av[ls * Nfp + p] = av;
    }
    __syncthreads();
    if (!act) continue;
    if (c == CODE_NONE || w == T(0)) {
      for (int q = 0; q < 3; ++q) {
        cown[q * nf + at] = 0;
        cnbr[q * nf + at] = 0;
      }
      continue;
    }
    T mj = 0, ma = 0;
#pragma unroll
    for (int q = 0; q < Nfp; ++q) {
      mj += m[q] * s_ju[ls * Nfp + q];
      ma += m[q] * s_av[ls * Nfp + q];
    }
    const bool inner = c >= 0;
    const T gjv = wtau * mj - w * ma;  // d/d[v]
    const T gav = -w * mj;             // d/d{dv/dn}
    const T gdM = inner ? T(0.5) * gav : gav, gdP = inner ? T(0.5) * gav : T(0);
    cown[at] = gjv;
    cown[nf + at] = nx * gdM;
    cown[2 * nf + at] = ny * gdM;
    cnbr[at] = inner ? -gjv : T(0);
    cnbr[nf + at] = nx * gdP;
    cnbr[2 * nf + at] = ny * gdP;
  }
}

__device__ inline double block_sum_256c(double v, double* s_red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
  __syncthreads();
  return t;
}

// LDS per element of a group (curved_elem_words): u, the face cotangents
// gv, gx, gy, the pulled-back cotangents vr, vs, the per-node metrics (rx,
// sx, ry, sy, and J for the nodal form), and with the nodal form its volume
// products a, b, with cubature the element's stiffness (and, for a
// Helmholtz apply, mass) matrix as a packed upper triangle of NPK words;
// then Dr, Ds, M for the workgroup's lifetime.
__host__ __device__ constexpr int curved_elem_words(int Np, bool cub, bool mass) {
  return cub ? 10 * Np + (Np * (Np + 1) / 2) * (mass ? 2 : 1) : 13 * Np;
}

// Elements per group of the element kernel: its LDS plan within 4096 words.
__host__ __device__ constexpr int curved_epb(int Np, bool cub, bool mass) {
  const int e = 4096 / curved_elem_words(Np, cub, mass);
  return e < 1 ? 1 : (e > 64 ? 64 : e);
}

template <typename T, int N, bool CUB>
__global__ void __launch_bounds__(256)
curved_elem_kernel(const T* __restrict__ tab, const T* __restrict__ geo,
                   const T* __restrict__ estiff, const T* __restrict__ emass,
                   const int* __restrict__ codes, const T* __restrict__ cown,
                   const T* __restrict__ cnbr, const T* __restrict__ u, T* __restrict__ y,
                   T alpha, T shift, int epb, int64_t K, double* __restrict__ partial,
                   const int* __restrict__ skip, const T* __restrict__ bvec) {
  if (skip != nullptr && *skip) return;
  constexpr int Np = CShape<N>::Np, Nfp = CShape<N>::Nfp, NN = SymPack<N>::NPK;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double s_red[4];
  T* sh = reinterpret_cast<T*>(smem_raw);
  const int L = epb * Np;
  const bool mass = alpha != T(0);
  constexpr int GW = CUB ? 4 : 5;  // metric planes staged (J: nodal form only)
  T *s_u = sh, *s_gv = sh + L, *s_gx = sh + 2 * L, *s_gy = sh + 3 * L;
  T* nxt = sh + 4 * L;
  T *s_a = nxt, *s_b = nxt + L;  // nodal form only
  if constexpr (!CUB) nxt += 2 * L;
  T *s_vr = nxt, *s_vs = nxt + L;
  T* s_geo = nxt + 2 * L;
  T* s_S = s_geo + GW * L;
  T* s_Q = s_S + (CUB ? epb * NN : 0);
  T* s_tab = s_Q + (CUB && mass ? epb * NN : 0);
  for (int i = threadIdx.x; i < 3 * Np * Np; i += 256) s_tab[i] = tab[i];
  const T* Dr = s_tab;
  const T* Ds = s_tab + Np * Np;
  const T* M = s_tab + 2 * Np * Np;
  const int64_t nf = K * 3 * Nfp;
  const int tid = threadIdx.x;
  const int64_t ngroups = (K + epb - 1) / epb;
  // The group's staging loads are issued (from clamped indices) before the
  // first LDS store, in compile-time-sized batches for the largest group
  // this kernel is launched with, and the three faces' cotangents in one
  // batch after their codes: a group pays a few memory round trips, not
  // one per staging pass and face.
  constexpr int EMAX = curved_epb(Np, CUB, false);
  constexpr int IU = (EMAX * Np + 255) / 256, IG = (EMAX * GW * Np + 255) / 256;
  constexpr int IS = CUB ? (EMAX * NN + 255) / 256 : 1;
  static_assert(EMAX * Nfp <= 256, "one face node per thread and face");
  double dacc = 0.0;
  for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const int64_t k0 = g * epb;
    const int ne = (int)((K - k0) < epb ? (K - k0) : epb);
    T ru[IU], rg[IG], rs[IS], rq[IS];
#pragma unroll
    for (int j = 0; j < IU; ++j) {
      const int idx = tid + 256 * j;
      ru[j] = u != nullptr ? u[k0 * Np + (idx < ne * Np ? idx : ne * Np - 1)] : T(0);
    }
    // the group's metrics (the first GW planes of each element's five):
    // contiguous runs, coalesced loads
#pragma unroll
    for (int j = 0; j < IG; ++j) {
      const int idx = tid + 256 * j;
      const int ic = idx < ne * GW * Np ? idx : ne * GW * Np - 1;
      const int e = ic / (GW * Np), r = ic - e * (GW * Np);
      rg[j] = geo[(k0 + e) * 5 * Np + r];
    }
    if constexpr (CUB) {
      // the group's element matrices are contiguous: coalesced loads
#pragma unroll
      for (int j = 0; j < IS; ++j) {
        const int idx = tid + 256 * j, ic = idx < ne * NN ? idx : ne * NN - 1;
        rs[j] = estiff[k0 * NN + ic];
        rq[j] = mass ? emass[k0 * NN + ic] : T(0);
      }
    }
    // face cotangents of this thread's face node (e, p) on each face: its
    // own side's, plus the neighbour side's for an interior face
    const bool fon = tid < ne * Nfp;
    const int fe = fon ? tid / Nfp : 0, fpn = fon ? tid - (tid / Nfp) * Nfp : 0;
    T cv[3][3];
    int64_t co[3];
    int cc[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const int64_t side = (k0 + fe) * 3 + f;
      const int64_t t = side * Nfp + fpn;
      cv[f][0] = cown[t];
      cv[f][1] = cown[nf + t];
      cv[f][2] = cown[2 * nf + t];
      cc[f] = codes[side];
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const int c = cc[f];
      co[f] = c >= 0 ? ((int64_t)nbr_elem(c) * 3 + nbr_face(c)) * Nfp +
                           (nbr_rev(c) ? Nfp - 1 - fpn : fpn)
                     : 0;
    }
    T cn[3][3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cn[f][0] = cnbr[co[f]];
      cn[f][1] = cnbr[nf + co[f]];
      cn[f][2] = cnbr[2 * nf + co[f]];
    }
    __syncthreads();  // the previous group's LDS reads are done
#pragma unroll
    for (int j = 0; j < IU; ++j) {
      const int idx = tid + 256 * j;
      if (idx < ne * Np) {
        s_u[idx] = ru[j];
        s_gv[idx] = 0;
        s_gx[idx] = 0;
        s_gy[idx] = 0;
      }
    }
#pragma unroll
    for (int j = 0; j < IG; ++j) {
      const int idx = tid + 256 * j;
      if (idx < ne * GW * Np) s_geo[idx] = rg[j];
    }
    if constexpr (CUB) {
#pragma unroll
      for (int j = 0; j < IS; ++j) {
        const int idx = tid + 256 * j;
        if (idx < ne * NN) {
          s_S[idx] = rs[j];
          if (mass) s_Q[idx] = rq[j];
        }
      }
    }
    __syncthreads();
    // face cotangents into the element rows, one face at a time (a vertex
    // node lies on two faces)
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (fon) {
        const bool in = cc[f] >= 0;
        const int i = fe * Np + fmask<N>(f, fpn);
        s_gv[i] += cv[f][0] + (in ? cn[f][0] : T(0));
        s_gx[i] += cv[f][1] + (in ? cn[f][1] : T(0));
        s_gy[i] += cv[f][2] + (in ? cn[f][2] : T(0));
      }
      __syncthreads();
    }
    if constexpr (!CUB) {
      // nodal volume: d/dv of sum J (ux (M vx) + uy (M vy)), per-node metrics
      for (int idx = tid; idx < ne * Np; idx += 256) {
        const int e = idx / Np, i = idx - e * Np;
        const T* ue = s_u + e * Np;
        T ur = 0, us = 0;
#pragma unroll
        for (int j = 0; j < Np; ++j) {
          ur += Dr[i * Np + j] * ue[j];
          us += Ds[i * Np + j] * ue[j];
        }
        const T* gk = s_geo + e * GW * Np;
        const T J = gk[4 * Np + i];
        s_a[idx] = J * (gk[i] * ur + gk[Np + i] * us);
        s_b[idx] = J * (gk[2 * Np + i] * ur + gk[3 * Np + i] * us);
      }
      __syncthreads();
    }
    for (int idx = tid; idx < ne * Np; idx += 256) {
      const int e = idx / Np, i = idx - e * Np;
      T wx = s_gx[idx], wy = s_gy[idx];
      if constexpr (!CUB) {
        const T* ae = s_a + e * Np;
        const T* be = s_b + e * Np;
#pragma unroll
        for (int j = 0; j < Np; ++j) {
          wx += M[i * Np + j] * ae[j];
          wy += M[i * Np + j] * be[j];
        }
      }
      const T* gk = s_geo + e * GW * Np;
      s_vr[idx] = gk[i] * wx + gk[2 * Np + i] * wy;
      s_vs[idx] = gk[Np + i] * wx + gk[3 * Np + i] * wy;
    }
    __syncthreads();
    for (int idx = tid; idx < ne * Np; idx += 256) {
      const int e = idx / Np, i = idx - e * Np;
      const T* vre = s_vr + e * Np;
      const T* vse = s_vs + e * Np;
      const T* ue = s_u + e * Np;
      T acc = s_gv[idx];
#pragma unroll
      for (int j = 0; j < Np; ++j) acc += Dr[j * Np + i] * vre[j] + Ds[j * Np + i] * vse[j];
      if constexpr (CUB) {
        // volume cubature: the element's stiffness row, and its mass row
        const T* Se = s_S + e * NN;
        T sv = 0;
#pragma unroll
        for (int j = 0; j < Np; ++j) sv += Se[SymPack<N>::pk(i, j)] * ue[j];
        acc += sv;
        if (mass) {
          const T* Qe = s_Q + e * NN;
          T mv = 0;
#pragma unroll
          for (int j = 0; j < Np; ++j) mv += Qe[SymPack<N>::pk(i, j)] * ue[j];
          acc += alpha * mv;
        }
      } else if (mass) {
        T mu = 0;
#pragma unroll
        for (int j = 0; j < Np; ++j) mu += M[i * Np + j] * ue[j];
        acc += alpha * s_geo[e * GW * Np + 4 * Np + i] * mu;
      }
      acc += shift * s_u[idx];
      const int64_t gi = k0 * Np + idx;
      y[gi] = bvec != nullptr ? bvec[gi] - acc : acc;
      if (partial != nullptr) dacc += (double)s_u[idx] * (double)acc;
    }
  }
  if (partial != nullptr) {
    dacc = block_sum_256c(dacc, s_red);
    if (tid == 0) partial[blockIdx.x] = dacc;
  }
}

// rhs = mass(f) + L_N(q): the nodal (J M f) or cubature (element matrix)
// mass and the Neumann surface load sJ_0 m1d q on Neumann sides
// (sipg.py:420-453, 849-864; cubature.py:242-245).  One thread per node.
template <typename T, int N, bool CUB>
__global__ void __launch_bounds__(256)
curved_mass_kernel(const T* __restrict__ tab, const T* __restrict__ geo,
                   const T* __restrict__ emass, const int* __restrict__ codes,
                   const T* __restrict__ fpar, const T* __restrict__ f, const T* __restrict__ q,
                   T* __restrict__ out, int64_t K) {
  constexpr int Np = CShape<N>::Np, Nfp = CShape<N>::Nfp;
  const T* M = tab + 2 * Np * Np;
  const T* m1d = tab + 3 * Np * Np;
  const int64_t n = K * Np;
  for (int64_t gi = (int64_t)blockIdx.x * 256 + threadIdx.x; gi < n;
       gi += (int64_t)gridDim.x * 256) {
    const int64_t k = gi / Np;
    const int i = (int)(gi - k * Np);
    T acc = 0;
    if (f != nullptr) {
      const T* fk = f + k * Np;
      T mf = 0;
      if constexpr (CUB) {
        const T* Qk = emass + k * SymPack<N>::NPK;
#pragma unroll
        for (int j = 0; j < Np; ++j) mf += Qk[SymPack<N>::pk(i, j)] * fk[j];
        acc = mf;
      } else {
#pragma unroll
        for (int j = 0; j < Np; ++j) mf += M[i * Np + j] * fk[j];
        acc = geo[k * 5 * Np + 4 * Np + i] * mf;
      }
    }
    if (q != nullptr) {
#pragma unroll
      for (int fc = 0; fc < 3; ++fc) {
        const int p = facepos<N>(fc, i);
        if (p < 0 || codes[k * 3 + fc] != CODE_NONE) continue;
        const T* qf = q + (k * 3 + fc) * Nfp;
        T v = 0;
#pragma unroll
        for (int r = 0; r < Nfp; ++r) v += m1d[fc * Nfp * Nfp + p * Nfp + r] * qf[r];
        acc += fpar[(k * 3 + fc) * 3 + 2] * v;
      }
    }
    out[gi] = acc;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
static int upload(const double* h, size_t n, void** d) {
  std::vector<T> v(h, h + n);
  DDGX_HIP(hipMalloc(d, n * sizeof(T)));
  DDGX_HIP(hipMemcpy(*d, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return DDGX_OK;
}

void curved_free(Curved* cv) {
  if (cv == nullptr) return;
  void* ptrs[] = {cv->geo,    cv->fnorm,  cv->fpar,  cv->fmass, cv->estiff, cv->emass,
                  cv->reftab, cv->codes,  cv->trace, cv->cown,  cv->cnbr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete cv;
}

int curved_create(Op& op, int use_cub, const double* reftab, const double* geo,
                  const double* fnorm, const double* fpar, const int32_t* codes,
                  const double* fmass, const double* estiff, const double* emass) {
  const int Np = op.Np, Nfp = op.Nfp;
  const int64_t K = op.K;
  if (use_cub && (fmass == nullptr || estiff == nullptr || emass == nullptr)) {
    set_error("cubature operator needs face masses and element stiffness and mass matrices");
    return DDGX_ERR_INVALID;
  }
  for (int64_t t = 0; t < 3 * K; ++t) {
    const int c = codes[t];
    if (c >= 0 && (nbr_elem(c) >= K || nbr_face(c) > 2)) {
      set_error("face code names a neighbour outside the mesh");
      return DDGX_ERR_INVALID;
    }
    if (c < CODE_NONE) {
      set_error("bad face code");
      return DDGX_ERR_INVALID;
    }
  }
  Curved* cv = new Curved();
  cv->use_cub = use_cub ? 1 : 0;
  const bool f64 = op.dtype == DDGX_F64;
  const size_t es = f64 ? 8 : 4;
  auto up = [&](const double* h, size_t n, void** d) {
    return f64 ? upload<double>(h, n, d) : upload<float>(h, n, d);
  };
  int rc = DDGX_OK;
  const size_t nfn = (size_t)K * 3 * Nfp;
  const size_t nn = (size_t)K * (Np * (Np + 1) / 2);
  do {
    if ((rc = up(reftab, 3 * (size_t)Np * Np + 3 * (size_t)Nfp * Nfp, &cv->reftab))) break;
    if ((rc = up(geo, (size_t)K * 5 * Np, &cv->geo))) break;
    if ((rc = up(fnorm, nfn * 2, &cv->fnorm))) break;
    if ((rc = up(fpar, (size_t)K * 9, &cv->fpar))) break;
    if (use_cub) {
      if ((rc = up(fmass, nfn * Nfp, &cv->fmass))) break;
      if ((rc = up(estiff, nn, &cv->estiff))) break;
      if ((rc = up(emass, nn, &cv->emass))) break;
    }
    hipError_t e = hipMalloc((void**)&cv->codes, 3 * (size_t)K * sizeof(int));
    if (e == hipSuccess)
      e = hipMemcpy(cv->codes, codes, 3 * (size_t)K * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&cv->trace, 3 * nfn * es);
    if (e == hipSuccess) e = hipMalloc(&cv->cown, 3 * nfn * es);
    if (e == hipSuccess) e = hipMalloc(&cv->cnbr, 3 * nfn * es);
    if (e != hipSuccess) rc = hip_fail(e, "curved operator buffers");
  } while (0);
  if (rc != DDGX_OK) {
    curved_free(cv);
    return rc;
  }
  // elements per workgroup of the element kernel: its LDS (elem_words per
  // element) within 4096 words
  // elements per group of the element kernel: its LDS plan within 4096 words
  // (the largest plan here; each launch sizes its groups to its own plan)
  const int epb = curved_epb(Np, use_cub != 0, use_cub != 0);
  cv->epb = epb;
  const int64_t cap = 4 * (int64_t)device_cu_count();
  auto grid = [&](int64_t items) {
    const int64_t g = (items + 255) / 256;
    return (int)(g < 1 ? 1 : (g < 8 * cap ? g : 8 * cap));
  };
  cv->grid_face = grid((int64_t)nfn);
  cv->grid_side = grid((int64_t)nfn);   // one thread per side node
  cv->grid_node = grid((int64_t)K * Np);
  const int64_t ngroups = (K + epb - 1) / epb;
  cv->grid_elem = (int)(ngroups < 2 * cap ? ngroups : 2 * cap);
  op.cv = cv;
  op.grid_apply = cv->grid_elem;
  return DDGX_OK;
}

template <typename T, int N>
static int curved_t(const Op& op, const void* u, void* y, const void* udir, double alpha,
                    double shift, double* partial, const int* skip, hipStream_t s,
                    const void* bvec, bool face_nodal) {
  const Curved& c = *op.cv;
  const int64_t K = op.K;
  hipLaunchKernelGGL((curved_trace_kernel<T, N>), dim3(c.grid_face), dim3(256), 0, s,
                     (const T*)c.reftab, (const T*)c.geo, (const T*)u, (T*)c.trace, K, skip);
  DDGX_HIP(hipGetLastError());
  const bool cub = c.use_cub && !face_nodal;
  const T* fm = cub ? (const T*)c.fmass : nullptr;
  hipLaunchKernelGGL((curved_side_kernel<T, N>), dim3(c.grid_side), dim3(256), 0, s,
                     (const T*)c.reftab, (const T*)c.fnorm, (const T*)c.fpar, fm, c.codes,
                     (const T*)c.trace, (const T*)udir, (T*)c.cown, (T*)c.cnbr, K, skip);
  DDGX_HIP(hipGetLastError());
  // the element matrices are staged whenever the operator has them (CUB
  // kernel), the tables always
  // group size for this launch's LDS plan (the grid, and so the number of dot
  // partials, is the operator's: idle workgroups write a zero partial)
  const int words = curved_elem_words(op.Np, cub, alpha != 0.0);
  const int epb = curved_epb(op.Np, cub, alpha != 0.0);
  const size_t lds = ((size_t)epb * words + 3 * (size_t)op.Np * op.Np) * sizeof(T);
  if (cub)
    hipLaunchKernelGGL((curved_elem_kernel<T, N, true>), dim3(c.grid_elem), dim3(256), lds, s,
                       (const T*)c.reftab, (const T*)c.geo, (const T*)c.estiff,
                       (const T*)c.emass, c.codes, (const T*)c.cown, (const T*)c.cnbr,
                       (const T*)u, (T*)y, (T)alpha, (T)shift, epb, K, partial, skip,
                       (const T*)bvec);
  else
    hipLaunchKernelGGL((curved_elem_kernel<T, N, false>), dim3(c.grid_elem), dim3(256), lds, s,
                       (const T*)c.reftab, (const T*)c.geo, (const T*)nullptr,
                       (const T*)nullptr, c.codes, (const T*)c.cown, (const T*)c.cnbr,
                       (const T*)u, (T*)y, (T)alpha, (T)shift, epb, K, partial, skip,
                       (const T*)bvec);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

#define CURVED_DISPATCH(FN, ...)                                       \
  switch (op.N) {                                                      \
    case 1: return op.dtype == DDGX_F64 ? FN<double, 1>(__VA_ARGS__) : FN<float, 1>(__VA_ARGS__); \
    case 2: return op.dtype == DDGX_F64 ? FN<double, 2>(__VA_ARGS__) : FN<float, 2>(__VA_ARGS__); \
    case 3: return op.dtype == DDGX_F64 ? FN<double, 3>(__VA_ARGS__) : FN<float, 3>(__VA_ARGS__); \
    case 4: return op.dtype == DDGX_F64 ? FN<double, 4>(__VA_ARGS__) : FN<float, 4>(__VA_ARGS__); \
    case 5: return op.dtype == DDGX_F64 ? FN<double, 5>(__VA_ARGS__) : FN<float, 5>(__VA_ARGS__); \
    case 6: return op.dtype == DDGX_F64 ? FN<double, 6>(__VA_ARGS__) : FN<float, 6>(__VA_ARGS__); \
    default: break;                                                    \
  }

int launch_curved(const Op& op, const void* u, void* y, const void* udir, double alpha,
                  double shift, int mode, double* partial, const int* skip, hipStream_t s,
                  const void* bvec, bool face_nodal) {
  if (mode & (M_SMOOTH | M_PROLONG)) {
    set_error("the multigrid smoother does not run on curved operators");
    return DDGX_ERR_UNSUPPORTED;
  }
  if ((mode & M_RESID) && bvec == nullptr) {
    set_error("residual mode needs the right-hand side");
    return DDGX_ERR_INVALID;
  }
  const void* b = (mode & M_RESID) ? bvec : nullptr;
  double* part = (mode & M_DOT) ? partial : nullptr;
  CURVED_DISPATCH(curved_t, op, u, y, udir, alpha, shift, part, skip, s, b, face_nodal)
  set_error("polynomial order " + std::to_string(op.N) + " not built (1..6)");
  return DDGX_ERR_UNSUPPORTED;
}

template <typename T, int N>
static int curved_rhs_t(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
                        hipStream_t s) {
  const Curved& c = *op.cv;
  if (c.use_cub)
    hipLaunchKernelGGL((curved_mass_kernel<T, N, true>), dim3(c.grid_node), dim3(256), 0, s,
                       (const T*)c.reftab, (const T*)c.geo, (const T*)c.emass, c.codes,
                       (const T*)c.fpar, (const T*)f, (const T*)q, (T*)rhs, op.K);
  else
    hipLaunchKernelGGL((curved_mass_kernel<T, N, false>), dim3(c.grid_node), dim3(256), 0, s,
                       (const T*)c.reftab, (const T*)c.geo, (const T*)nullptr, c.codes,
                       (const T*)c.fpar, (const T*)f, (const T*)q, (T*)rhs, op.K);
  DDGX_HIP(hipGetLastError());
  // - L_D: the nodal-quadrature apply of the zero field with the Dirichlet
  // ghost values (sipg_dirichlet_load_2d has no cubature, sipg.py:398-417),
  // subtracted in place through the residual mode
  if (udir != nullptr)
    return curved_t<T, N>(op, nullptr, rhs, udir, 0.0, 0.0, nullptr, nullptr, s, rhs, true);
  return DDGX_OK;
}

int launch_curved_rhs(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
                      hipStream_t s) {
  CURVED_DISPATCH(curved_rhs_t, op, f, udir, q, rhs, s)
  set_error("polynomial order " + std::to_string(op.N) + " not built (1..6)");
  return DDGX_ERR_UNSUPPORTED;
}

#undef CURVED_DISPATCH

}  // namespace ddgx
