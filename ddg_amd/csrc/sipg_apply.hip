// Fused matrix-free SIPG apply for gfx950:
//   y = (A_sipg + alpha*J*M + shift*I) u
// with A_sipg the mass-weighted SIPG Laplacian of src/ddg/dg2d/sipg.py:142-337
// (mixed Dirichlet/Neumann, per-side penalty tau*(N+1)^2*sJ/J), evaluated in
// the closed form of its reverse-mode gradient:
//
//   volume : y_k = ArT (vrr ur + vrs us) + AsT (vrs ur + vss us),
//            ur = Dr u_k, us = Ds u_k, ArT = Dr^T M^T, AsT = Ds^T M^T
//   face f : P = pen_f [u] - sJw_f (dn^- + dn^+)    -> y[Fmask_f] += m1d_f^T P
//            X = -sJw_f [u]                          -> y += (a_f LR_f + b_f LS_f) X
//
// [u] = u^- - u^+, dn^- = a_f ur + b_f us at the own face nodes, dn^+ the
// neighbour's normal derivative taken with the own normal.  The neighbour
// traces come from the tile in LDS when the neighbour is in the same tile and
// are recomputed from its nodal values (cached in L2) otherwise, so the
// kernel reads u, the element records and the face codes once from HBM and
// writes y once.  No atomics: every element owns its output.
//
// Layout: one 256-thread workgroup processes tiles of EPB elements with NL
// lanes per element (NL = next power of two >= Np); the workgroups are
// persistent (grid-stride over tiles) so the reference-element tables are
// staged into LDS once per workgroup.

#include "ddgx_internal.h"

namespace ddgx {

template <typename T, int N, int MODE>
__global__ void __launch_bounds__(256)
sipg_apply_kernel(const T* __restrict__ tab, const T* __restrict__ rec,
                  const int* __restrict__ nbr, const T* __restrict__ u,
                  T* __restrict__ y, const T* __restrict__ udir, T alpha, T shift,
                  int K, int probe_j, double* __restrict__ partial,
                  const int* __restrict__ skip, const T* __restrict__ bvec) {
  using S = Shape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, NL = S::NL, EPB = S::EPB;
  constexpr bool HELM = MODE & M_HELM, UDIR = MODE & M_UDIR;
  constexpr bool PROBE = MODE & M_PROBE, DOT = MODE & M_DOT, RESID = MODE & M_RESID;

  if (skip != nullptr && *skip) return;  // CG already converged

  __shared__ T s_tab[S::size];
  __shared__ T s_u[EPB * Np];
  __shared__ T s_ur[EPB * Np];
  __shared__ T s_us[EPB * Np];
  __shared__ T s_wr[EPB * Np];
  __shared__ T s_ws[EPB * Np];
  __shared__ T s_P[EPB * 3 * Nfp];
  __shared__ T s_X[EPB * 3 * Nfp];
  __shared__ T s_rec[EPB * DDGX_REC_WORDS];
  __shared__ int s_nbr[EPB * 4];
  __shared__ double s_red[4];

  const int tid = threadIdx.x;
  for (int i = tid; i < S::size; i += 256) s_tab[i] = tab[i];

  const int e = tid / NL;
  const int li = tid % NL;
  double dot_acc = 0.0;
  const int ntiles = (K + EPB - 1) / EPB;

  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int k0 = tile * EPB;
    const int ne = min(EPB, K - k0);
    __syncthreads();  // tables loaded / previous tile done with LDS

    // ---- phase 1: stage the tile ------------------------------------------
    for (int i = tid; i < ne * Np; i += 256) {
      if constexpr (PROBE)
        s_u[i] = (i % Np) == probe_j ? T(1) : T(0);
      else
        s_u[i] = u[(size_t)k0 * Np + i];
    }
    for (int i = tid; i < ne * DDGX_REC_WORDS; i += 256)
      s_rec[i] = rec[(size_t)k0 * DDGX_REC_WORDS + i];
    for (int i = tid; i < ne * 4; i += 256) s_nbr[i] = nbr[(size_t)k0 * 4 + i];
    __syncthreads();

    // ---- phase 2: reference derivatives and metric products ----------------
    if (li < Np && e < ne) {
      T ur = 0, us = 0;
#pragma unroll
      for (int j = 0; j < Np; ++j) {
        const T v = s_u[e * Np + j];
        ur += s_tab[S::oDr + li * Np + j] * v;
        us += s_tab[S::oDs + li * Np + j] * v;
      }
      const T* r = s_rec + e * DDGX_REC_WORDS;
      s_ur[e * Np + li] = ur;
      s_us[e * Np + li] = us;
      s_wr[e * Np + li] = r[R_VRR] * ur + r[R_VRS] * us;
      s_ws[e * Np + li] = r[R_VRS] * ur + r[R_VSS] * us;
    }
    __syncthreads();

    // ---- phase 3: face traces, jumps, averages, penalty ---------------------
    for (int fi = tid; fi < ne * 3 * Nfp; fi += 256) {
      const int ee = fi / (3 * Nfp);
      const int f = (fi / Nfp) % 3;
      const int p = fi % Nfp;
      const int n = fmask<N>(f, p);
      const T* r = s_rec + ee * DDGX_REC_WORDS;
      const T uM = s_u[ee * Np + n];
      const T dnM = r[R_A + f] * s_ur[ee * Np + n] + r[R_B + f] * s_us[ee * Np + n];
      const int code = s_nbr[ee * 4 + f];
      T uP, dnP;
      if (code >= 0) {
        if constexpr (PROBE) {
          uP = T(0);  // neighbours carry no probe: diagonal block only
          dnP = T(0);
        } else {
          const int kn = nbr_elem(code), fn = nbr_face(code);
          const int pn = nbr_rev(code) ? (Nfp - 1 - p) : p;
          const int nn = fmask<N>(fn, pn);
          const int el = kn - k0;
          if (el >= 0 && el < ne) {
            uP = s_u[el * Np + nn];
            const T* rn = s_rec + el * DDGX_REC_WORDS;
            dnP = -(rn[R_A + fn] * s_ur[el * Np + nn] + rn[R_B + fn] * s_us[el * Np + nn]);
          } else {
            const T* un = u + (size_t)kn * Np;
            T ur = 0, us = 0;
#pragma unroll
            for (int j = 0; j < Np; ++j) {
              const T v = un[j];
              ur += s_tab[S::oDr + nn * Np + j] * v;
              us += s_tab[S::oDs + nn * Np + j] * v;
            }
            uP = un[nn];
            const T* rn = rec + (size_t)kn * DDGX_REC_WORDS;
            dnP = -(rn[R_A + fn] * ur + rn[R_B + fn] * us);
          }
        }
      } else if (code == CODE_DIRICHLET) {
        if constexpr (UDIR)
          uP = udir[(size_t)(k0 + ee) * Np + n];
        else
          uP = T(0);
        dnP = T(0);
      } else {
        uP = uM;  // no coupling: weights are zero
        dnP = T(0);
      }
      const T ju = uM - uP;
      s_P[fi] = r[R_PEN + f] * ju - r[R_SJW + f] * (dnM + dnP);
      s_X[fi] = -r[R_SJW + f] * ju;
    }
    __syncthreads();

    // ---- phase 4: volume contraction + face lifts, write y ----------------
    if (li < Np && e < ne) {
      const T* r = s_rec + e * DDGX_REC_WORDS;
      T acc = 0;
#pragma unroll
      for (int j = 0; j < Np; ++j)
        acc += s_tab[S::oAr + li * Np + j] * s_wr[e * Np + j] +
               s_tab[S::oAs + li * Np + j] * s_ws[e * Np + j];
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const T af = r[R_A + f], bf = r[R_B + f];
#pragma unroll
        for (int p = 0; p < Nfp; ++p)
          acc += (af * s_tab[S::oLR + (f * Np + li) * Nfp + p] +
                  bf * s_tab[S::oLS + (f * Np + li) * Nfp + p]) *
                 s_X[(e * 3 + f) * Nfp + p];
      }
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        int q = -1;
#pragma unroll
        for (int p = 0; p < Nfp; ++p)
          if (fmask<N>(f, p) == li) q = p;
        if (q >= 0) {
#pragma unroll
          for (int p = 0; p < Nfp; ++p)
            acc += s_tab[S::oL1 + (f * Nfp + q) * Nfp + p] * s_P[(e * 3 + f) * Nfp + p];
        }
      }
      if constexpr (HELM) {
        T m = 0;
#pragma unroll
        for (int j = 0; j < Np; ++j) m += s_tab[S::oM + li * Np + j] * s_u[e * Np + j];
        acc += alpha * r[R_J] * m + shift * s_u[e * Np + li];
      }
      if constexpr (PROBE)
        y[((size_t)(k0 + e) * Np + li) * Np + probe_j] = acc;
      else if constexpr (RESID)
        y[(size_t)(k0 + e) * Np + li] = bvec[(size_t)(k0 + e) * Np + li] - acc;
      else
        y[(size_t)(k0 + e) * Np + li] = acc;
      if constexpr (DOT) dot_acc += (double)s_u[e * Np + li] * (double)acc;
    }
  }

  if constexpr (DOT) {
    // fixed-order wave then block reduction -> one partial per workgroup
    for (int off = 32; off > 0; off >>= 1) dot_acc += __shfl_down(dot_acc, off, 64);
    if ((tid & 63) == 0) s_red[tid >> 6] = dot_acc;
    __syncthreads();
    if (tid == 0) partial[blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
  }
}

template <typename T, int N, int MODE>
static int launch_t(const Op& op, const void* u, void* y, const void* udir, double alpha,
                    double shift, int probe_j, double* partial, const int* skip,
                    hipStream_t s, const void* bvec) {
  const int K = (int)op.K;
  const int ntiles = (K + Shape<N>::EPB - 1) / Shape<N>::EPB;
  int grid = op.grid_apply > 0 ? op.grid_apply : 1;
  if (!(MODE & M_DOT)) grid = ntiles < grid ? ntiles : grid;
  hipLaunchKernelGGL((sipg_apply_kernel<T, N, MODE>), dim3(grid), dim3(256), 0, s,
                     (const T*)op.tab, (const T*)op.rec, op.nbr, (const T*)u, (T*)y,
                     (const T*)udir, (T)alpha, (T)shift, K, probe_j, partial, skip,
                     (const T*)bvec);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

template <typename T, int N>
static int launch_mode(const Op& op, const void* u, void* y, const void* udir, double alpha,
                       double shift, int mode, int probe_j, double* partial, const int* skip,
                       hipStream_t s, const void* bvec) {
  switch (mode) {
#define CASE(m) \
  case m: return launch_t<T, N, m>(op, u, y, udir, alpha, shift, probe_j, partial, skip, s, bvec);
    CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(8) CASE(9) CASE(16) CASE(17)
#undef CASE
    default:
      set_error("unsupported apply mode " + std::to_string(mode));
      return DDGX_ERR_UNSUPPORTED;
  }
}

int launch_apply(const Op& op, const void* u, void* y, const void* udir, double alpha,
                 double shift, int mode, int probe_j, double* partial, const int* skip,
                 hipStream_t s, const void* bvec) {
  if ((mode & M_RESID) && bvec == nullptr) {
    set_error("residual mode needs the right-hand side");
    return DDGX_ERR_INVALID;
  }
  if (alpha != 0.0 || shift != 0.0) mode |= M_HELM;
  if (udir != nullptr) mode |= M_UDIR;
#define DISPATCH_N(T)                                                                      \
  switch (op.N) {                                                                          \
    case 1: return launch_mode<T, 1>(op, u, y, udir, alpha, shift, mode, probe_j, partial, skip, s, bvec); \
    case 2: return launch_mode<T, 2>(op, u, y, udir, alpha, shift, mode, probe_j, partial, skip, s, bvec); \
    case 3: return launch_mode<T, 3>(op, u, y, udir, alpha, shift, mode, probe_j, partial, skip, s, bvec); \
    case 4: return launch_mode<T, 4>(op, u, y, udir, alpha, shift, mode, probe_j, partial, skip, s, bvec); \
    case 5: return launch_mode<T, 5>(op, u, y, udir, alpha, shift, mode, probe_j, partial, skip, s, bvec); \
    case 6: return launch_mode<T, 6>(op, u, y, udir, alpha, shift, mode, probe_j, partial, skip, s, bvec); \
    default: break;                                                                        \
  }
  if (op.dtype == DDGX_F64) {
    DISPATCH_N(double)
  } else {
    DISPATCH_N(float)
  }
#undef DISPATCH_N
  set_error("polynomial order " + std::to_string(op.N) + " not built (1..6)");
  return DDGX_ERR_UNSUPPORTED;
}

// Occupancy-derived persistent grid for the apply (same for every mode).
template <typename T, int N>
static int apply_grid_t() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sipg_apply_kernel<T, N, 9>, 256, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * device_cu_count();
}

int apply_grid(int N, int dtype) {
  if (dtype == DDGX_F64) {
    switch (N) {
      case 1: return apply_grid_t<double, 1>();
      case 2: return apply_grid_t<double, 2>();
      case 3: return apply_grid_t<double, 3>();
      case 4: return apply_grid_t<double, 4>();
      case 5: return apply_grid_t<double, 5>();
      case 6: return apply_grid_t<double, 6>();
    }
  } else {
    switch (N) {
      case 1: return apply_grid_t<float, 1>();
      case 2: return apply_grid_t<float, 2>();
      case 3: return apply_grid_t<float, 3>();
      case 4: return apply_grid_t<float, 4>();
      case 5: return apply_grid_t<float, 5>();
      case 6: return apply_grid_t<float, 6>();
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Right-hand side: rhs = J M f - L_D(u_dir) + L_N(q)   (sipg.py:849-864)
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__(256)
rhs_kernel(const T* __restrict__ tab, const T* __restrict__ rec, const int* __restrict__ nbr,
           const T* __restrict__ fsj, const T* __restrict__ f, const T* __restrict__ udir,
           const T* __restrict__ q, T* __restrict__ out, int K) {
  using S = Shape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, NL = S::NL, EPB = S::EPB;
  __shared__ T s_tab[S::size];
  __shared__ T s_f[EPB * Np];
  __shared__ T s_P[EPB * 3 * Nfp];
  __shared__ T s_X[EPB * 3 * Nfp];
  const int tid = threadIdx.x;
  for (int i = tid; i < S::size; i += 256) s_tab[i] = tab[i];
  const int e = tid / NL, li = tid % NL;
  const int ntiles = (K + EPB - 1) / EPB;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int k0 = tile * EPB;
    const int ne = min(EPB, K - k0);
    __syncthreads();
    for (int i = tid; i < ne * Np; i += 256) s_f[i] = f ? f[(size_t)k0 * Np + i] : T(0);
    for (int fi = tid; fi < ne * 3 * Nfp; fi += 256) {
      const int ee = fi / (3 * Nfp), fc = (fi / Nfp) % 3, p = fi % Nfp;
      const size_t k = (size_t)(k0 + ee);
      const int code = nbr[k * 4 + fc];
      const T* r = rec + k * DDGX_REC_WORDS;
      T P = 0, X = 0;
      if (code == CODE_DIRICHLET && udir != nullptr) {
        // minus the Dirichlet load: apply(0; u_dir) has [u] = -u_dir
        const T g = udir[k * Np + fmask<N>(fc, p)];
        P = r[R_PEN + fc] * g;
        X = -r[R_SJW + fc] * g;
      }
      s_P[fi] = P;
      s_X[fi] = X;
    }
    __syncthreads();
    if (li < Np && e < ne) {
      const size_t k = (size_t)(k0 + e);
      const T* r = rec + k * DDGX_REC_WORDS;
      T m = 0;
#pragma unroll
      for (int j = 0; j < Np; ++j) m += s_tab[S::oM + li * Np + j] * s_f[e * Np + j];
      T acc = r[R_J] * m;
#pragma unroll
      for (int fc = 0; fc < 3; ++fc) {
        const T af = r[R_A + fc], bf = r[R_B + fc];
#pragma unroll
        for (int p = 0; p < Nfp; ++p)
          acc += (af * s_tab[S::oLR + (fc * Np + li) * Nfp + p] +
                  bf * s_tab[S::oLS + (fc * Np + li) * Nfp + p]) *
                 s_X[(e * 3 + fc) * Nfp + p];
        int qp = -1;
#pragma unroll
        for (int p = 0; p < Nfp; ++p)
          if (fmask<N>(fc, p) == li) qp = p;
        if (qp >= 0) {
          T lift = 0;
#pragma unroll
          for (int p = 0; p < Nfp; ++p)
            lift += s_tab[S::oL1 + (fc * Nfp + qp) * Nfp + p] * s_P[(e * 3 + fc) * Nfp + p];
          acc += lift;
          if (q != nullptr && fsj[k * 4 + fc] != T(0)) {
            // Neumann load: sJ * (m1d_f q_f)[qp]  (sipg.py:442-453); L1[f][p][qp] = m1d_f[qp][p]
            T mq = 0;
#pragma unroll
            for (int p = 0; p < Nfp; ++p)
              mq += s_tab[S::oL1 + (fc * Nfp + p) * Nfp + qp] * q[k * 3 * Nfp + fc * Nfp + p];
            acc += fsj[k * 4 + fc] * mq;
          }
        }
      }
      out[k * Np + li] = acc;
    }
  }
}

template <typename T, int N>
static int launch_rhs_t(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
                        hipStream_t s) {
  const int K = (int)op.K;
  const int ntiles = (K + Shape<N>::EPB - 1) / Shape<N>::EPB;
  int grid = op.grid_apply > 0 ? op.grid_apply : 1;
  grid = ntiles < grid ? ntiles : grid;
  hipLaunchKernelGGL((rhs_kernel<T, N>), dim3(grid), dim3(256), 0, s, (const T*)op.tab,
                     (const T*)op.rec, op.nbr, (const T*)op.fsj, (const T*)f, (const T*)udir,
                     (const T*)q, (T*)rhs, K);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

int launch_rhs(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
               hipStream_t s) {
#define DISPATCH_N(T)                                                 \
  switch (op.N) {                                                     \
    case 1: return launch_rhs_t<T, 1>(op, f, udir, q, rhs, s);        \
    case 2: return launch_rhs_t<T, 2>(op, f, udir, q, rhs, s);        \
    case 3: return launch_rhs_t<T, 3>(op, f, udir, q, rhs, s);        \
    case 4: return launch_rhs_t<T, 4>(op, f, udir, q, rhs, s);        \
    case 5: return launch_rhs_t<T, 5>(op, f, udir, q, rhs, s);        \
    case 6: return launch_rhs_t<T, 6>(op, f, udir, q, rhs, s);        \
    default: break;                                                   \
  }
  if (op.dtype == DDGX_F64) {
    DISPATCH_N(double)
  } else {
    DISPATCH_N(float)
  }
#undef DISPATCH_N
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

}  // namespace ddgx
