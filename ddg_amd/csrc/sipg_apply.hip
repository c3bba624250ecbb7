// Fused matrix-free SIPG apply for gfx950:
//   y = (A_sipg + alpha*J*M + shift*I) u          (or y = b - that, residual mode)
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
// neighbour's normal derivative taken with the own normal.
//
// The kernels are in sipg_apply_mfma.h (matrix cores, 64-element tiles,
// halo traces through LDS) and sipg_apply_valu.h (vector ALUs, one thread
// per element: order 1); this file holds the launch code and the
// right-hand-side kernel.

#include "ddgx_internal.h"
#include "sipg_apply_mfma.h"
#include "sipg_apply_valu.h"

namespace ddgx {

template <typename T, int N, int MODE>
static int launch_t(const Op& op, const void* u, void* y, const void* udir, double alpha,
                    double shift, double* partial, const int* skip, hipStream_t s,
                    const void* bvec, const void* binv = nullptr, double omega = 0.0,
                    const void* pxc = nullptr, const void* pmat = nullptr, int npt = 0,
                    void* xco = nullptr, CGScalars* sc = nullptr, const double* rzp = nullptr,
                    int nrz = 0, int q = 0) {
  // one workgroup per 64-element tile; in dot mode each writes one partial
  const int* perm = op.perm_id ? nullptr : op.perm;
  if constexpr (UseValuApply<T, N>::value)
    hipLaunchKernelGGL((sipg_apply_valu_kernel<T, N, MODE>), dim3(op.grid_apply), dim3(TILE), 0,
                       s, (const T*)op.tab, (const T*)op.rec, op.nbrp, perm, (const T*)u, (T*)y,
                       (const T*)udir, (T)alpha, (T)shift, (int)op.K, partial, skip,
                       (const T*)bvec, (const T*)binv, (T)omega, (const T*)pxc, (const T*)pmat,
                       (T*)xco, npt);
  else if (op.halo_cap == HALO_SMALL)
    hipLaunchKernelGGL((sipg_apply_mfma_kernel<T, N, MODE, HALO_SMALL>), dim3(op.grid_apply),
                       dim3(256), 0, s, (const T*)op.frag, (const T*)op.rec, op.fcode, perm,
                       op.hpos, (const T*)op.hab, (const T*)u, (T*)y, (const T*)udir, (T)alpha,
                       (T)shift, (int)op.K, partial, skip, (const T*)bvec, (const T*)binv,
                       (T)omega, (const T*)pmat, npt, (T*)xco, sc, rzp, nrz, q);
  else
    hipLaunchKernelGGL((sipg_apply_mfma_kernel<T, N, MODE, HALO_MAX>), dim3(op.grid_apply),
                       dim3(256), 0, s, (const T*)op.frag, (const T*)op.rec, op.fcode, perm,
                       op.hpos, (const T*)op.hab, (const T*)u, (T*)y, (const T*)udir, (T)alpha,
                       (T)shift, (int)op.K, partial, skip, (const T*)bvec, (const T*)binv,
                       (T)omega, (const T*)pmat, npt, (T*)xco, sc, rzp, nrz, q);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

template <typename T, int N>
static int launch_mode(const Op& op, const void* u, void* y, const void* udir, double alpha,
                       double shift, int mode, double* partial, const int* skip, hipStream_t s,
                       const void* bvec) {
  switch (mode) {
#define CASE(m) \
  case m: return launch_t<T, N, m>(op, u, y, udir, alpha, shift, partial, skip, s, bvec);
    CASE(0) CASE(1) CASE(2) CASE(3) CASE(8) CASE(9) CASE(16) CASE(17)
#undef CASE
    default:
      set_error("unsupported apply mode " + std::to_string(mode));
      return DDGX_ERR_UNSUPPORTED;
  }
}

// Float64, orders 2-4: there the apply holds the direction update's extra
// values (z and p_old of the tile and the halo) without register spills at
// its four waves per SIMD; float32 and orders 5-6 would spill.
bool apply_pupdate_ok(const Op& op) {
  return op.cv == nullptr && op.dtype == DDGX_F64 && op.N >= 2 && op.N <= 4 && op.nseg == 1;
}

template <typename T, int N>
static int pupdate_n(const Op& op, const void* z, const void* p_old, void* p, void* Ap,
                     double alpha, double shift, double* partial, CGScalars* sc,
                     const double* rz, int nrz, int q, hipStream_t s) {
  if (alpha != 0.0 || shift != 0.0)
    return launch_t<T, N, M_DOT | M_PUPDATE | M_HELM>(op, z, Ap, nullptr, alpha, shift, partial,
                                                      nullptr, s, p_old, nullptr, 0.0, nullptr,
                                                      nullptr, 0, p, sc, rz, nrz, q);
  return launch_t<T, N, M_DOT | M_PUPDATE>(op, z, Ap, nullptr, alpha, shift, partial, nullptr, s,
                                           p_old, nullptr, 0.0, nullptr, nullptr, 0, p, sc, rz,
                                           nrz, q);
}

int launch_apply_pupdate(const Op& op, const void* z, const void* p_old, void* p, void* Ap,
                         double alpha, double shift, double* partial, CGScalars* sc,
                         const double* rz, int nrz, int q, hipStream_t s) {
  if (!apply_pupdate_ok(op)) {
    set_error("the direction update inside the apply needs an affine float64 operator of "
              "order 2..4");
    return DDGX_ERR_UNSUPPORTED;
  }
#define PCASE(T, n) \
  case n: return pupdate_n<T, n>(op, z, p_old, p, Ap, alpha, shift, partial, sc, rz, nrz, q, s);
  switch (op.N) { PCASE(double, 2) PCASE(double, 3) PCASE(double, 4) }
#undef PCASE
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

int launch_apply(const Op& op, const void* u, void* y, const void* udir, double alpha,
                 double shift, int mode, double* partial, const int* skip, hipStream_t s,
                 const void* bvec) {
  if (op.cv != nullptr)
    return launch_curved(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec, false);
  if ((mode & M_RESID) && bvec == nullptr) {
    set_error("residual mode needs the right-hand side");
    return DDGX_ERR_INVALID;
  }
  if (alpha != 0.0 || shift != 0.0) mode |= M_HELM;
  if (udir != nullptr) mode |= M_UDIR;
#define DISPATCH_N(T)                                                                      \
  switch (op.N) {                                                                          \
    case 1: return launch_mode<T, 1>(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec); \
    case 2: return launch_mode<T, 2>(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec); \
    case 3: return launch_mode<T, 3>(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec); \
    case 4: return launch_mode<T, 4>(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec); \
    case 5: return launch_mode<T, 5>(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec); \
    case 6: return launch_mode<T, 6>(op, u, y, udir, alpha, shift, mode, partial, skip, s, bvec); \
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

// Fused block-Jacobi sweep of the multigrid smoother: the apply's smoothing
// mode keeps the residual on chip (registers at order 1, LDS rows on the
// matrix cores), and one launch replaces the residual apply and the sweep.
// the fused sweeps are modes of the affine kernels; a general (curved)
// level smooths with the unfused block-Jacobi sweeps
bool smooth_fused(const Op& op) { return op.cv == nullptr && op.N >= 1 && op.N <= 6; }

bool smooth_prolongs(const Op& op) { return op.cv == nullptr && op.N == 1; }

template <typename T, int N>
static int smooth_n(const Op& op, const void* binv, double omega, double alpha, const void* x_in,
                    void* x_out, const void* b, const int* skip, hipStream_t s, const void* pxc,
                    const void* pmat, void* pto, const void* ptm, int np_to) {
  if (pto != nullptr) {
    constexpr int M = M_SMOOTH | M_PROLONG_OUT;
    if (alpha != 0.0)
      return launch_t<T, N, M | M_HELM>(op, x_in, x_out, nullptr, alpha, 0.0, nullptr, skip, s,
                                         b, binv, omega, nullptr, ptm, np_to, pto);
    return launch_t<T, N, M>(op, x_in, x_out, nullptr, alpha, 0.0, nullptr, skip, s, b, binv,
                             omega, nullptr, ptm, np_to, pto);
  }
  if constexpr (N == 1) {
    if (pxc != nullptr) {
      if (alpha != 0.0)
        return launch_t<T, N, M_SMOOTH | M_HELM | M_PROLONG>(op, x_in, x_out, nullptr, alpha, 0.0,
                                                             nullptr, skip, s, b, binv, omega,
                                                             pxc, pmat);
      return launch_t<T, N, M_SMOOTH | M_PROLONG>(op, x_in, x_out, nullptr, alpha, 0.0, nullptr,
                                                  skip, s, b, binv, omega, pxc, pmat);
    }
  }
  if (alpha != 0.0)
    return launch_t<T, N, M_SMOOTH | M_HELM>(op, x_in, x_out, nullptr, alpha, 0.0, nullptr, skip,
                                             s, b, binv, omega);
  return launch_t<T, N, M_SMOOTH>(op, x_in, x_out, nullptr, alpha, 0.0, nullptr, skip, s, b,
                                  binv, omega);
}

int launch_smooth(const Op& op, const void* binv, double omega, double alpha, const void* x_in,
                  void* x_out, const void* b, const int* skip, hipStream_t s,
                  const void* prolong_from, const void* prolong_mat, void* prolong_to,
                  const void* prolong_to_mat, int np_to) {
  if (prolong_from != nullptr && (!smooth_prolongs(op) || prolong_mat == nullptr)) {
    set_error("fused smoothing sweep: prolongation on the fly is built for order 1 only");
    return DDGX_ERR_UNSUPPORTED;
  }
  if (prolong_to != nullptr &&
      (prolong_from != nullptr || prolong_to_mat == nullptr || np_to < op.Np || np_to > 28)) {
    set_error("fused smoothing sweep: bad prolongation into the finer level");
    return DDGX_ERR_INVALID;
  }
  if (!smooth_fused(op)) {
    set_error("fused smoothing sweep: polynomial order not built");
    return DDGX_ERR_UNSUPPORTED;
  }
  if (x_in == x_out || binv == nullptr || b == nullptr) {
    set_error("fused smoothing sweep: x_out must not alias x_in; blocks and b are required");
    return DDGX_ERR_INVALID;
  }
#define SMOOTH_N(T)                                                                   \
  switch (op.N) {                                                                     \
    case 1: return smooth_n<T, 1>(op, binv, omega, alpha, x_in, x_out, b, skip, s, PF, PM, PT, PTM, NPT); \
    case 2: return smooth_n<T, 2>(op, binv, omega, alpha, x_in, x_out, b, skip, s, PF, PM, PT, PTM, NPT); \
    case 3: return smooth_n<T, 3>(op, binv, omega, alpha, x_in, x_out, b, skip, s, PF, PM, PT, PTM, NPT); \
    case 4: return smooth_n<T, 4>(op, binv, omega, alpha, x_in, x_out, b, skip, s, PF, PM, PT, PTM, NPT); \
    case 5: return smooth_n<T, 5>(op, binv, omega, alpha, x_in, x_out, b, skip, s, PF, PM, PT, PTM, NPT); \
    case 6: return smooth_n<T, 6>(op, binv, omega, alpha, x_in, x_out, b, skip, s, PF, PM, PT, PTM, NPT); \
  }
  const void* PF = prolong_from;
  const void* PM = prolong_mat;
  void* PT = prolong_to;
  const void* PTM = prolong_to_mat;
  const int NPT = np_to;
  if (op.dtype == DDGX_F64) {
    SMOOTH_N(double)
  } else {
    SMOOTH_N(float)
  }
#undef SMOOTH_N
  set_error("fused smoothing sweep: polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

// Residual + restriction (+ the coarser level's first sweep) in one launch:
// the order-1 kernel takes h-transfers (identity tile order, whole families
// in a tile), the matrix-core kernels p-transfers.
static bool resid_restrict_fused(const Op& op, int tkind, int np_lo) {
  if (op.cv != nullptr || op.N < 1 || op.N > 6) return false;
  if (op.N == 1) return tkind == 1 && op.perm_id && op.K % 4 == 0 && np_lo == op.Np;
  return tkind == 0 && np_lo >= 1 && np_lo <= op.Np;
}

template <typename T, int N>
static int resid_restrict_n(const Op& op, const void* P, int np_lo, const void* x, const void* b,
                            double alpha, void* b_c, const void* binv_c, double omega, void* x_c,
                            const int* skip, hipStream_t s) {
  constexpr int M = M_RESID | M_RESTRICT;
  if (alpha != 0.0)
    return launch_t<T, N, M | M_HELM>(op, x, b_c, nullptr, alpha, 0.0, nullptr, skip, s, b,
                                       binv_c, omega, nullptr, P, np_lo, x_c);
  return launch_t<T, N, M>(op, x, b_c, nullptr, alpha, 0.0, nullptr, skip, s, b, binv_c, omega,
                           nullptr, P, np_lo, x_c);
}

int launch_resid_restrict(const Op& op, int tkind, int np_lo, const void* P, const void* x,
                          const void* b, double alpha, void* b_c, const void* binv_c,
                          double omega, void* x_c, const int* skip, hipStream_t s) {
  if (!resid_restrict_fused(op, tkind, np_lo)) return DDGX_ERR_UNSUPPORTED;
#define RR_N(T)                                                                          \
  switch (op.N) {                                                                        \
    case 1: return resid_restrict_n<T, 1>(op, P, np_lo, x, b, alpha, b_c, binv_c, omega, x_c, skip, s); \
    case 2: return resid_restrict_n<T, 2>(op, P, np_lo, x, b, alpha, b_c, binv_c, omega, x_c, skip, s); \
    case 3: return resid_restrict_n<T, 3>(op, P, np_lo, x, b, alpha, b_c, binv_c, omega, x_c, skip, s); \
    case 4: return resid_restrict_n<T, 4>(op, P, np_lo, x, b, alpha, b_c, binv_c, omega, x_c, skip, s); \
    case 5: return resid_restrict_n<T, 5>(op, P, np_lo, x, b, alpha, b_c, binv_c, omega, x_c, skip, s); \
    case 6: return resid_restrict_n<T, 6>(op, P, np_lo, x, b, alpha, b_c, binv_c, omega, x_c, skip, s); \
  }
  if (op.dtype == DDGX_F64) {
    RR_N(double)
  } else {
    RR_N(float)
  }
#undef RR_N
  return DDGX_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
// Right-hand side: rhs = J M f - L_D(u_dir) + L_N(q)   (sipg.py:849-864)
// One thread per element; records and face codes in structure-of-arrays.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__(256)
rhs_kernel(const T* __restrict__ tab, const T* __restrict__ rec, const int* __restrict__ nbr,
           const int* __restrict__ perm, const T* __restrict__ fsj, const T* __restrict__ f,
           const T* __restrict__ udir, const T* __restrict__ q, T* __restrict__ out, int K) {
  using S = Shape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp;
  const T* __restrict__ Mm = tab + S::oM;
  const T* __restrict__ L1 = tab + S::oL1;
  const T* __restrict__ LR = tab + S::oLR;
  const T* __restrict__ LS = tab + S::oLS;
  // k: tile position (records, codes); kg: element id (vectors)
  for (int k = blockIdx.x * 256 + threadIdx.x; k < K; k += gridDim.x * 256) {
    const size_t kg = (size_t)perm[k];
    T ff[Np], yy[Np];
#pragma unroll
    for (int j = 0; j < Np; ++j) ff[j] = f ? f[kg * Np + j] : T(0);
    const T J = rec[(size_t)R_J * K + k];
#pragma unroll
    for (int i = 0; i < Np; ++i) {
      T m = 0;
#pragma unroll
      for (int j = 0; j < Np; ++j) m += Mm[j * S::RS + i] * ff[j];
      yy[i] = J * m;
    }
#pragma unroll
    for (int fc = 0; fc < 3; ++fc) {
      const int code = nbr[(size_t)fc * K + k];
      if (code == CODE_DIRICHLET && udir != nullptr) {
        // minus the Dirichlet load: apply(0; u_dir) has [u] = -u_dir
        const T pen = rec[(size_t)(R_PEN + fc) * K + k];
        const T sjw = rec[(size_t)(R_SJW + fc) * K + k];
        const T a = rec[(size_t)(R_A + fc) * K + k], b = rec[(size_t)(R_B + fc) * K + k];
        T P[Nfp], Xa[Nfp], Xb[Nfp];
#pragma unroll
        for (int p = 0; p < Nfp; ++p) {
          const T g = udir[kg * Np + fmask<N>(fc, p)];
          P[p] = pen * g;
          Xa[p] = -sjw * g * a;
          Xb[p] = -sjw * g * b;
        }
#pragma unroll
        for (int qq = 0; qq < Nfp; ++qq) {
          T acc = 0;
#pragma unroll
          for (int p = 0; p < Nfp; ++p) acc += L1[(fc * Nfp + p) * S::RSF + qq] * P[p];
          yy[fmask<N>(fc, qq)] += acc;
        }
#pragma unroll
        for (int i = 0; i < Np; ++i) {
          T acc = 0;
#pragma unroll
          for (int p = 0; p < Nfp; ++p)
            acc += LR[(fc * Nfp + p) * S::RS + i] * Xa[p] + LS[(fc * Nfp + p) * S::RS + i] * Xb[p];
          yy[i] += acc;
        }
      }
      const T sj = fsj[(size_t)k * 4 + fc];
      if (q != nullptr && sj != T(0)) {
        // Neumann load: sJ * (m1d_f q_f) at the face rows (sipg.py:442-453);
        // L1[f][p][qq] = m1d_f[qq][p]
#pragma unroll
        for (int qq = 0; qq < Nfp; ++qq) {
          T mq = 0;
#pragma unroll
          for (int p = 0; p < Nfp; ++p)
            mq += L1[(fc * Nfp + qq) * S::RSF + p] * q[kg * 3 * Nfp + fc * Nfp + p];
          yy[fmask<N>(fc, qq)] += sj * mq;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < Np; ++i) out[kg * Np + i] = yy[i];
  }
}

template <typename T, int N>
static int launch_rhs_t(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
                        hipStream_t s) {
  const int K = (int)op.K;
  int grid = (K + 255) / 256;
  const int cap = 8 * device_cu_count();
  grid = grid < cap ? grid : cap;
  hipLaunchKernelGGL((rhs_kernel<T, N>), dim3(grid), dim3(256), 0, s, (const T*)op.tab,
                     (const T*)op.rec, op.nbr, op.perm, (const T*)op.fsj, (const T*)f, (const T*)udir,
                     (const T*)q, (T*)rhs, K);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

int launch_rhs(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
               hipStream_t s) {
  if (op.cv != nullptr) return launch_curved_rhs(op, f, udir, q, rhs, s);
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
