// Block-Jacobi preconditioner: per-element diagonal blocks A_kk of the SIPG
// (or Helmholtz) operator and their inverses (src/ddg/dg2d/sipg.py:671-805).
//
// The reference extracts A_kk with ncolours*Np operator applies on coloured
// probe vectors (sipg.py:671-704).  Here the apply kernel runs in its probe
// mode: every element carries the unit vector e_j and neighbour traces are
// taken as zero -- exactly what a colour probe sees, since face neighbours
// never share a colour -- so Np applies give every block's column j.  The
// blocks are inverted in float64 by Gauss-Jordan elimination (SPD blocks, no
// pivoting), one wavefront per element with the augmented block in LDS, and
// stored transposed in the precision of the vectors they will multiply.

#include "ddgx_internal.h"

namespace ddgx {

template <typename T, int N>
__global__ void __launch_bounds__(64)
invert_blocks_kernel(const double* __restrict__ blocks, T* __restrict__ binvT, int K) {
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
    T* out = binvT + (size_t)k * Np * Np;
    for (int idx = t; idx < Np * Np; idx += 64) {
      const int j = idx / Np, i = idx % Np;  // out[j][i] = inv[i][j]
      out[idx] = (T)A[i][Np + j];
    }
  }
}

template <int N>
static int invert_t(int64_t K, int dtype, const double* blocks, void* binvT, hipStream_t s) {
  int grid = 16 * device_cu_count();
  if ((int64_t)grid > K) grid = (int)K;
  if (dtype == DDGX_F64)
    hipLaunchKernelGGL((invert_blocks_kernel<double, N>), dim3(grid), dim3(64), 0, s, blocks,
                       (double*)binvT, (int)K);
  else
    hipLaunchKernelGGL((invert_blocks_kernel<float, N>), dim3(grid), dim3(64), 0, s, blocks,
                       (float*)binvT, (int)K);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

int launch_invert_blocks(int N, int64_t K, int dtype, const double* blocks, void* binvT,
                         hipStream_t s) {
  switch (N) {
    case 1: return invert_t<1>(K, dtype, blocks, binvT, s);
    case 2: return invert_t<2>(K, dtype, blocks, binvT, s);
    case 3: return invert_t<3>(K, dtype, blocks, binvT, s);
    case 4: return invert_t<4>(K, dtype, blocks, binvT, s);
    case 5: return invert_t<5>(K, dtype, blocks, binvT, s);
    case 6: return invert_t<6>(K, dtype, blocks, binvT, s);
  }
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

}  // namespace ddgx
