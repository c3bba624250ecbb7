// Block-Jacobi preconditioner: per-element diagonal blocks A_kk of the SIPG
// (or Helmholtz) operator and their inverses (src/ddg/dg2d/sipg.py:671-805).
//
// The reference extracts A_kk with ncolours*Np operator applies on coloured
// probe vectors (sipg.py:671-704).  Here a probe kernel applies the operator
// to the unit vector e_j in every element with all neighbour traces taken as
// zero -- exactly what a colour probe sees, since face neighbours never share
// a colour -- so Np launches give every block's column j.  The blocks are
// inverted in float64 by Gauss-Jordan elimination (SPD blocks, no pivoting),
// one wavefront per element with the augmented block in LDS, and stored
// packed by cyclic diagonals (diag_pk) in the precision of the vectors they
// will multiply.  Blocks of a general (curved-element) operator come from
// an actual colour probe instead (launch_block_probe_general).

#include <vector>

#include "ddgx_internal.h"

namespace ddgx {

// Column j of every diagonal block: y = (A_kk + alpha J_k M) e_j, one thread
// per element (tile position k; the output is indexed by the element id).
template <int N>
__global__ void __launch_bounds__(256)
block_probe_kernel(const double* __restrict__ tab, const double* __restrict__ rec,
                   const int* __restrict__ nbr, const int* __restrict__ perm,
                   double* __restrict__ blocks, double alpha, int K, int j) {
  using S = Shape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= K) return;
  const double* Dr = tab + S::oDr;
  const double* Ds = tab + S::oDs;
  const double* Ar = tab + S::oAr;
  const double* As = tab + S::oAs;
  const double* Mm = tab + S::oM;
  const double* L1 = tab + S::oL1;
  const double* LR = tab + S::oLR;
  const double* LS = tab + S::oLS;
  double rc[DDGX_REC_WORDS];
  for (int w = 0; w < DDGX_REC_WORDS; ++w) rc[w] = rec[(size_t)w * K + k];
  // u = e_j: ur = Dr[:, j], us = Ds[:, j] (tables are stored by columns)
  double wr[Np], ws[Np], y[Np];
  for (int i = 0; i < Np; ++i) {
    const double a = Dr[j * S::RS + i], b = Ds[j * S::RS + i];
    wr[i] = rc[R_VRR] * a + rc[R_VRS] * b;
    ws[i] = rc[R_VRS] * a + rc[R_VSS] * b;
  }
  for (int i = 0; i < Np; ++i) {
    double acc = 0.0;
    for (int c = 0; c < Np; ++c) acc += Ar[c * S::RS + i] * wr[c] + As[c * S::RS + i] * ws[c];
    y[i] = acc + alpha * rc[R_J] * Mm[j * S::RS + i];
  }
  for (int f = 0; f < 3; ++f) {
    const int code = nbr[(size_t)f * K + k];
    if (code == CODE_NONE) continue;  // no coupling
    // neighbours carry no probe (u+ = 0, dn+ = 0); Dirichlet ghost is 0 too
    double P[Nfp], Xa[Nfp], Xb[Nfp];
    for (int p = 0; p < Nfp; ++p) {
      const int n = fmask<N>(f, p);
      const double uM = (n == j) ? 1.0 : 0.0;
      const double dnM = rc[R_A + f] * Dr[j * S::RS + n] + rc[R_B + f] * Ds[j * S::RS + n];
      P[p] = rc[R_PEN + f] * uM - rc[R_SJW + f] * dnM;
      const double X = -rc[R_SJW + f] * uM;
      Xa[p] = rc[R_A + f] * X;
      Xb[p] = rc[R_B + f] * X;
    }
    for (int q = 0; q < Nfp; ++q) {
      double acc = 0.0;
      for (int p = 0; p < Nfp; ++p) acc += L1[(f * Nfp + p) * S::RSF + q] * P[p];
      y[fmask<N>(f, q)] += acc;
    }
    for (int i = 0; i < Np; ++i) {
      double acc = 0.0;
      for (int p = 0; p < Nfp; ++p)
        acc += LR[(f * Nfp + p) * S::RS + i] * Xa[p] + LS[(f * Nfp + p) * S::RS + i] * Xb[p];
      y[i] += acc;
    }
  }
  double* out = blocks + (size_t)perm[k] * Np * Np;
  for (int i = 0; i < Np; ++i) out[i * Np + j] = y[i];
}

template <int N>
static int probe_t(const Op& op, double* blocks, double alpha, hipStream_t s) {
  const int K = (int)op.K;
  for (int j = 0; j < Shape<N>::Np; ++j) {
    hipLaunchKernelGGL((block_probe_kernel<N>), dim3((K + 255) / 256), dim3(256), 0, s,
                       (const double*)op.tab, (const double*)op.rec, op.nbr, op.perm, blocks,
                       alpha, K, j);
    DDGX_HIP(hipGetLastError());
  }
  return DDGX_OK;
}

int launch_block_probe(const Op& op, double* blocks, double alpha, hipStream_t s) {
  if (op.dtype != DDGX_F64) {
    set_error("block probes need a float64 operator");
    return DDGX_ERR_INVALID;
  }
  switch (op.N) {
    case 1: return probe_t<1>(op, blocks, alpha, s);
    case 2: return probe_t<2>(op, blocks, alpha, s);
    case 3: return probe_t<3>(op, blocks, alpha, s);
    case 4: return probe_t<4>(op, blocks, alpha, s);
    case 5: return probe_t<5>(op, blocks, alpha, s);
    case 6: return probe_t<6>(op, blocks, alpha, s);
  }
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

// The general (curved-element) operator has per-node metrics, so its blocks
// come from the reference's colour probe itself (sipg.py:651-704): elements
// coloured so that face neighbours differ, e_j in every element of colour c,
// one apply, and column j of each such element's block read off its own
// output rows (a neighbour of another colour contributes nothing there).
__global__ void __launch_bounds__(256)
probe_vector_kernel(double* __restrict__ u, const int* __restrict__ colour, int c, int j,
                    int64_t K, int Np) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= K * Np) return;
  const int64_t k = i / Np;
  u[i] = (colour[k] == c && i - k * Np == j) ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(256)
probe_gather_kernel(const double* __restrict__ y, const int* __restrict__ colour, int c, int j,
                    double* __restrict__ blocks, int64_t K, int Np) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= K * Np) return;
  const int64_t k = i / Np;
  if (colour[k] == c) blocks[k * Np * Np + (i - k * Np) * Np + j] = y[i];
}

int launch_block_probe_general(const Op& op, double* blocks, double alpha, hipStream_t s) {
  if (op.dtype != DDGX_F64 || op.cv == nullptr) {
    set_error("the colour probe takes a float64 general operator");
    return DDGX_ERR_INVALID;
  }
  const int64_t K = op.K;
  const int Np = op.Np;
  std::vector<int> codes((size_t)K * 3), colour((size_t)K, -1);
  DDGX_HIP(hipMemcpyAsync(codes.data(), op.cv->codes, codes.size() * sizeof(int),
                          hipMemcpyDeviceToHost, s));
  DDGX_HIP(hipStreamSynchronize(s));
  // greedy colouring in element order: the smallest colour no coloured face
  // neighbour has (at most 4 for triangles)
  int ncol = 0;
  for (int64_t k = 0; k < K; ++k) {
    unsigned used = 0;
    for (int f = 0; f < 3; ++f) {
      const int c = codes[(size_t)k * 3 + f];
      if (c >= 0 && colour[nbr_elem(c)] >= 0) used |= 1u << colour[nbr_elem(c)];
    }
    int cc = 0;
    while (used & (1u << cc)) ++cc;
    colour[k] = cc;
    ncol = cc + 1 > ncol ? cc + 1 : ncol;
  }
  double *u = nullptr, *y = nullptr;
  int* dcol = nullptr;
  int rc = DDGX_OK;
  do {
    hipError_t e = hipMalloc((void**)&u, (size_t)K * Np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&y, (size_t)K * Np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&dcol, (size_t)K * sizeof(int));
    if (e == hipSuccess)
      e = hipMemcpyAsync(dcol, colour.data(), (size_t)K * sizeof(int), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { rc = hip_fail(e, "colour probe workspace"); break; }
    const int g = (int)((K * Np + 255) / 256);
    for (int c = 0; c < ncol && rc == DDGX_OK; ++c)
      for (int j = 0; j < Np && rc == DDGX_OK; ++j) {
        hipLaunchKernelGGL(probe_vector_kernel, dim3(g), dim3(256), 0, s, u,
                           (const int*)dcol, c, j, K, Np);
        rc = launch_apply(op, u, y, nullptr, alpha, 0.0, 0, nullptr, nullptr, s);
        if (rc != DDGX_OK) break;
        hipLaunchKernelGGL(probe_gather_kernel, dim3(g), dim3(256), 0, s, (const double*)y,
                           (const int*)dcol, c, j, blocks, K, Np);
      }
    if (rc != DDGX_OK) break;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "colour probe");
  } while (false);
  if (u) (void)hipFree(u);
  if (y) (void)hipFree(y);
  if (dcol) (void)hipFree(dcol);
  return rc;
}

template <typename T, int N, bool FULL>
__global__ void __launch_bounds__(64)
invert_blocks_kernel(const double* __restrict__ blocks, T* __restrict__ binv, int K) {
  constexpr int Np = Shape<N>::Np;
  constexpr int W = 2 * Np;  // augmented width, <= 56
  __shared__ double A[Np][W + 1];
  __shared__ double fac[Np];
  const int t = threadIdx.x;
  for (int k = blockIdx.x; k < K; k += gridDim.x) {
    const double* B = blocks + (size_t)k * Np * Np;
    __syncthreads();
    for (int idx = t; idx < Np * W; idx += 64) {
      const int i = idx / W, j = idx % W;
      A[i][j] = j < Np ? B[i * Np + j] : (j - Np == i ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int c = 0; c < Np; ++c) {
      const double piv = A[c][c];
      __syncthreads();
      if (t < W) A[c][t] /= piv;
      __syncthreads();
      if (t < Np) fac[t] = A[t][c];
      __syncthreads();
      if (t < W) {
        const double acj = A[c][t];
        for (int i = 0; i < Np; ++i)
          if (i != c) A[i][t] -= fac[i] * acj;
      }
      __syncthreads();
    }
    if constexpr (FULL) {
      // the whole inverse, row-major (a non-symmetric block)
      T* out = binv + (size_t)k * Np * Np;
      for (int idx = t; idx < Np * Np; idx += 64) {
        const int i = idx / Np, j = idx % Np;
        out[idx] = (T)A[i][Np + j];
      }
    } else {
      // the symmetric part (the inverse of a symmetric block, symmetric up to
      // rounding), packed by cyclic diagonals (diag_pk): entry (i, j), i <= j,
      // is written once
      T* out = binv + (size_t)k * SymPack<N>::NPK;
      for (int idx = t; idx < Np * Np; idx += 64) {
        const int i = idx / Np, j = idx % Np;
        if (i <= j) out[diag_pk(Np, i, j)] = (T)(0.5 * (A[i][Np + j] + A[j][Np + i]));
      }
    }
  }
}

template <int N, bool FULL>
static int invert_tf(int64_t K, int dtype, const double* blocks, void* binv, hipStream_t s) {
  int grid = 16 * device_cu_count();
  if ((int64_t)grid > K) grid = (int)K;
  if (dtype == DDGX_F64)
    hipLaunchKernelGGL((invert_blocks_kernel<double, N, FULL>), dim3(grid), dim3(64), 0, s,
                       blocks, (double*)binv, (int)K);
  else
    hipLaunchKernelGGL((invert_blocks_kernel<float, N, FULL>), dim3(grid), dim3(64), 0, s,
                       blocks, (float*)binv, (int)K);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

template <int N>
static int invert_t(int64_t K, int dtype, const double* blocks, void* binv, hipStream_t s,
                    bool full) {
  return full ? invert_tf<N, true>(K, dtype, blocks, binv, s)
              : invert_tf<N, false>(K, dtype, blocks, binv, s);
}

size_t packed_block_words(int N) {
  const size_t Np = (size_t)(N + 1) * (N + 2) / 2;
  return Np * (Np + 1) / 2;
}

size_t block_words(int N, bool full) {
  const size_t Np = (size_t)(N + 1) * (N + 2) / 2;
  return full ? Np * Np : packed_block_words(N);
}

int launch_invert_blocks(int N, int64_t K, int dtype, const double* blocks, void* binv,
                         hipStream_t s, bool full) {
  switch (N) {
    case 1: return invert_t<1>(K, dtype, blocks, binv, s, full);
    case 2: return invert_t<2>(K, dtype, blocks, binv, s, full);
    case 3: return invert_t<3>(K, dtype, blocks, binv, s, full);
    case 4: return invert_t<4>(K, dtype, blocks, binv, s, full);
    case 5: return invert_t<5>(K, dtype, blocks, binv, s, full);
    case 6: return invert_t<6>(K, dtype, blocks, binv, s, full);
  }
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

}  // namespace ddgx
