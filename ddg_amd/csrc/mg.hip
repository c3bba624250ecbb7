// Multigrid V-cycle preconditioner (src/ddg/dg2d/multigrid.py:53-367).
//
// Levels run fine to coarse.  Level l holds an operator A_l + alpha M_l (an
// Op at its own order and mesh), the inverse diagonal blocks of that
// operator (the block-Jacobi smoother, multigrid.py:87-99), and a transfer
// to level l+1 given as one dense matrix P applied per "coarse unit":
//
//   p-transfer (same mesh, lower order)  P = I, (Np_l x Np_{l+1}), unit = element
//   h-transfer (same order, 4x coarser)  P = [I_0; I_1; I_2; I_3], (4 Np x Np),
//        unit = coarse element k, whose fine children 4k..4k+3 are contiguous
//
//   prolong   hi[k] += P lo[k]        restrict   lo[k] = P^T hi[k]
//
// which is multigrid.py:117-134, 166-184 with the child ordering of the
// refinement forest.  The coarsest level is solved with a dense float64
// inverse assembled on the device (column j = A e_j) and inverted by
// Gauss-Jordan elimination; the reference LU-factors the same matrix
// (multigrid.py:220-228).
//
// The cycle (multigrid.py:232-244), x = 0 on entry:
//   nu times   x += omega Binv (b - A x)       (first sweep: x = omega Binv b)
//   b_c = restrict(b - A x);  x += prolong(vcycle(b_c))
//   nu times   x += omega Binv (b - A x)
// The residuals use the apply's residual mode (one fused kernel); every
// kernel takes the CG stop flag so a chunk of CG iterations can be enqueued
// past convergence without doing work.

#include <cmath>
#include <vector>

#include "ddgx_internal.h"

namespace ddgx {

constexpr int MG_MAXP = 112 * 28;  // largest transfer matrix (h-transfer at N=6)
constexpr int MG_MAXIN = 2560;     // staged input values per workgroup pass
constexpr int MG_MAX_COARSE = 8192;

// x = (accumulate ? x : 0) + omega * Binv r, element by element
template <typename T, int N, bool FULL>
__global__ void __launch_bounds__(256)
bj_smooth_kernel(const T* __restrict__ binv, const T* __restrict__ r, T* __restrict__ x, int K,
                 T omega, int accumulate, const int* __restrict__ skip) {
  using S = Shape<N>;
  constexpr int Np = S::Np, NL = S::NL, EPB = S::EPB;
  if (skip != nullptr && *skip) return;
  __shared__ T s_r[EPB * Np];
  __shared__ T s_b[EPB * BlockWords<N, FULL>::value];
  const int tid = threadIdx.x, e = tid / NL, li = tid % NL;
  const int ntiles = (K + EPB - 1) / EPB;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int k0 = tile * EPB;
    const int ne = min(EPB, K - k0);
    __syncthreads();
    for (int i = tid; i < ne * Np; i += 256) s_r[i] = r[(size_t)k0 * Np + i];
    stage_blocks<T, N, FULL>(binv, s_b, k0, ne, tid);
    __syncthreads();
    if (li < Np && e < ne) {
      const T acc = block_row<T, N, FULL>(s_b, s_r, e, li);
      const size_t g = (size_t)(k0 + e) * Np + li;
      x[g] = accumulate ? x[g] + omega * acc : omega * acc;
    }
  }
}

// lo[k][j] = sum_h P[h][j] hi[k][h]; with x_c != null also the coarse level's
// first smoothing sweep x_c[k] = omega Binv_c[k] lo[k] (the arithmetic of
// bj_smooth_kernel without accumulation), so that level skips its launch
template <typename T>
__global__ void __launch_bounds__(256)
restrict_kernel(const T* __restrict__ P, const T* __restrict__ hi, T* __restrict__ lo, int Ku,
                int nhi, int nlo, const int* __restrict__ skip, const T* __restrict__ binv_c,
                T omega, T* __restrict__ x_c) {
  if (skip != nullptr && *skip) return;
  __shared__ T sP[MG_MAXP];
  __shared__ T sIn[MG_MAXIN];
  __shared__ T sOut[256];
  const int npk = nlo * (nlo + 1) / 2;
  for (int i = threadIdx.x; i < nhi * nlo; i += 256) sP[i] = P[i];
  const int E = 256 / nlo;
  const int e = threadIdx.x / nlo, j = threadIdx.x % nlo;
  for (int k0 = blockIdx.x * E; k0 < Ku; k0 += gridDim.x * E) {
    const int ne = min(E, Ku - k0);
    __syncthreads();
    for (int i = threadIdx.x; i < ne * nhi; i += 256) sIn[i] = hi[(size_t)k0 * nhi + i];
    __syncthreads();
    T acc = 0;
    if (e < ne) {
      for (int h = 0; h < nhi; ++h) acc += sP[h * nlo + j] * sIn[e * nhi + h];
      lo[(size_t)(k0 + e) * nlo + j] = acc;
    }
    if (x_c != nullptr) {
      if (e < E) sOut[e * nlo + j] = acc;
      __syncthreads();
      if (e < ne) {
        const T* bk = binv_c + (size_t)(k0 + e) * npk;
        T a = 0;
        for (int d = 0; d < nlo; ++d) {
          const int i = j + d < nlo ? j + d : j + d - nlo;
          a += bk[diag_pk(nlo, j, i)] * sOut[e * nlo + i];
        }
        x_c[(size_t)(k0 + e) * nlo + j] = omega * a;
      }
    }
  }
}

// hi[k][h] += sum_l P[h][l] lo[k][l]
template <typename T>
__global__ void __launch_bounds__(256)
prolong_add_kernel(const T* __restrict__ P, const T* __restrict__ lo, T* __restrict__ hi, int Ku,
                   int nhi, int nlo, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  __shared__ T sP[MG_MAXP];
  __shared__ T sIn[256];
  for (int i = threadIdx.x; i < nhi * nlo; i += 256) sP[i] = P[i];
  const int E = 256 / nhi;
  const int e = threadIdx.x / nhi, h = threadIdx.x % nhi;
  for (int k0 = blockIdx.x * E; k0 < Ku; k0 += gridDim.x * E) {
    const int ne = min(E, Ku - k0);
    __syncthreads();
    for (int i = threadIdx.x; i < ne * nlo; i += 256) sIn[i] = lo[(size_t)k0 * nlo + i];
    __syncthreads();
    if (e < ne) {
      T acc = 0;
      for (int l = 0; l < nlo; ++l) acc += sP[h * nlo + l] * sIn[e * nlo + l];
      hi[(size_t)(k0 + e) * nhi + h] += acc;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
zero_kernel(T* __restrict__ x, size_t n, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) x[i] = T(0);
}

// Coarse solve x = Ainv b, with Ainv stored so that x_i = sum_j W[j][i] b_j
// (coalesced across i).  Stage 1: partial[jc][i] over a chunk of j;
// stage 2: fixed-order sum over chunks.
// Coarse solve x = W b as column chunks (partial sums, then a fixed-order
// sum): at most GEMV_NCH chunks of at least 16 columns, so the coarse-size
// range (<= 8192 unknowns) gets ~400 workgroups of short dependent chains.
constexpr int GEMV_JC = 128;  // largest chunk (MG_MAX_COARSE / GEMV_NCH)
constexpr int GEMV_NCH = 64;
static_assert(GEMV_JC * GEMV_NCH >= MG_MAX_COARSE, "coarse chunk too narrow");
static int gemv_jc(int n) {
  const int jc = (n + GEMV_NCH - 1) / GEMV_NCH;
  return jc < 16 ? 16 : jc;
}

// Segment blockIdx.z (batched hierarchies: one coarse vector per request,
// the same inverse for all) reads b and writes part at its own offsets.
template <typename T>
__global__ void __launch_bounds__(256)
coarse_gemv_partial_kernel(const double* __restrict__ W, const T* __restrict__ b,
                           double* __restrict__ part, int n, int jc, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  b += (size_t)blockIdx.z * n;
  part += (size_t)blockIdx.z * gridDim.y * n;
  __shared__ double sb[GEMV_JC];
  const int j0 = blockIdx.y * jc;
  const int nj = min(jc, n - j0);
  if (threadIdx.x < nj) sb[threadIdx.x] = (double)b[j0 + threadIdx.x];
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < nj; ++j) acc += W[(size_t)(j0 + j) * n + i] * sb[j];
  part[(size_t)blockIdx.y * n + i] = acc;
}

// x[i] = sum of the chunk partials in chunk order (64-thread workgroups;
// the unrolled loop keeps the independent loads in flight together)
template <typename T>
__global__ void __launch_bounds__(64)
coarse_gemv_sum_kernel(const double* __restrict__ part, T* __restrict__ x, int n, int nchunks,
                       const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  part += (size_t)blockIdx.y * nchunks * n;  // segment blockIdx.y
  x += (size_t)blockIdx.y * n;
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
#pragma unroll 16
  for (int c = 0; c < nchunks; ++c) acc += part[(size_t)c * n + i];
  x[i] = (T)acc;
}

// --- coarse matrix assembly and inversion (set-up only) ---------------------

__global__ void unit_kernel(double* e, int j_on, int j_off) {
  if (threadIdx.x == 0) {
    if (j_off >= 0) e[j_off] = 0.0;
    e[j_on] = 1.0;
  }
}

// d = max_i |A_ii| ; A_ii += eps d   (one workgroup)
__global__ void __launch_bounds__(256)
regularize_kernel(double* A, int n, double eps) {
  __shared__ double s[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(A[(size_t)i * n + i]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  const double add = eps * s[0];
  for (int i = threadIdx.x; i < n; i += 256) A[(size_t)i * n + i] += add;
}

// In-place Gauss-Jordan inversion, pivot k (SPD: no pivoting):
//   row k  <- row k / p, with entry k = 1/p
//   row i  <- row i - A_ik row k (new), with entry k = -A_ik / p
__global__ void __launch_bounds__(256)
gj_pivot_kernel(const double* __restrict__ A, int n, int k, double* __restrict__ prow,
                double* __restrict__ col) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const double p = A[(size_t)k * n + k];
  prow[t] = t == k ? 1.0 / p : A[(size_t)k * n + t] / p;
  col[t] = t == k ? 0.0 : A[(size_t)t * n + k];
}

__global__ void __launch_bounds__(256)
gj_update_kernel(double* __restrict__ A, int n, int k, const double* __restrict__ prow,
                 const double* __restrict__ col) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const size_t g = (size_t)i * n + j;
  if (i == k)
    A[g] = prow[j];
  else
    A[g] = (j == k ? 0.0 : A[g]) - col[i] * prow[j];
}

__global__ void diag_kernel(const double* __restrict__ A, int n, double* __restrict__ d) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = A[(size_t)i * n + i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static size_t dsz(int dtype) { return dtype == DDGX_F64 ? 8 : 4; }

static int vec_grid(size_t n) {
  const size_t g = (n + 255) / 256;
  const size_t cap = 4 * (size_t)device_cu_count();
  return (int)(g < cap ? (g > 0 ? g : 1) : cap);
}

template <typename T, int N>
static int bj_smooth_t(const MGLevel& L, const void* r, void* x, double omega, int acc,
                       const int* skip, hipStream_t s) {
  const int K = (int)L.op->K;
  const int ntiles = (K + Shape<N>::EPB - 1) / Shape<N>::EPB;
  // workgroups per compute unit: float64 blocks of the higher orders gain from
  // one workgroup per tile (N=4 at 2x256x256: 28.5 against 34.7 us with 8 per
  // CU looping over tiles); the lighter sweeps are faster with 8
  constexpr int PER_CU = (sizeof(T) == 8 && N >= 4) ? 32 : 8;
  int grid = PER_CU * device_cu_count();
  grid = ntiles < grid ? ntiles : grid;
  if (L.full)
    hipLaunchKernelGGL((bj_smooth_kernel<T, N, true>), dim3(grid), dim3(256), 0, s,
                       (const T*)L.binv, (const T*)r, (T*)x, K, (T)omega, acc, skip);
  else
    hipLaunchKernelGGL((bj_smooth_kernel<T, N, false>), dim3(grid), dim3(256), 0, s,
                       (const T*)L.binv, (const T*)r, (T*)x, K, (T)omega, acc, skip);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

template <typename T>
static int bj_smooth(const MGLevel& L, const void* r, void* x, double omega, int acc,
                     const int* skip, hipStream_t s) {
  switch (L.op->N) {
    case 1: return bj_smooth_t<T, 1>(L, r, x, omega, acc, skip, s);
    case 2: return bj_smooth_t<T, 2>(L, r, x, omega, acc, skip, s);
    case 3: return bj_smooth_t<T, 3>(L, r, x, omega, acc, skip, s);
    case 4: return bj_smooth_t<T, 4>(L, r, x, omega, acc, skip, s);
    case 5: return bj_smooth_t<T, 5>(L, r, x, omega, acc, skip, s);
    case 6: return bj_smooth_t<T, 6>(L, r, x, omega, acc, skip, s);
  }
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

#define MG_CHECK(x)                    \
  do {                                 \
    int _rc = (x);                     \
    if (_rc != DDGX_OK) return _rc;    \
  } while (0)

template <typename T>
static int coarse_solve(const MG& mg, MGWork& w, const void* b, void* x, const int* skip,
                        hipStream_t s) {
  const int n = mg.nc;
  const int jc = gemv_jc(n);
  const int nch = (n + jc - 1) / jc;
  hipLaunchKernelGGL((coarse_gemv_partial_kernel<T>), dim3((n + 255) / 256, nch, mg.nseg),
                     dim3(256), 0, s, (const double*)mg.Ainv, (const T*)b, w.gemv_part, n, jc,
                     skip);
  DDGX_HIP(hipGetLastError());
  hipLaunchKernelGGL((coarse_gemv_sum_kernel<T>), dim3((n + 63) / 64, mg.nseg), dim3(64), 0, s,
                     (const double*)w.gemv_part, (T*)x, n, nch, skip);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

// prolong: dst += P src; restrict: dst = P^T src, and with x_c != null the
// coarse level's first sweep x_c = omega Binv_c dst
template <typename T>
static int transfer(const MGLevel& L, const MGLevel& C, bool prolong, const void* src, void* dst,
                    const int* skip, hipStream_t s, void* x_c = nullptr, double omega = 0.0) {
  const int Ku = (int)C.op->K;
  const int nlo = C.op->Np;
  const int nhi = (int)((L.op->K * L.op->Np) / C.op->K);
  const int E = prolong ? 256 / nhi : 256 / nlo;
  int grid = (Ku + E - 1) / E;
  const int cap = 4 * device_cu_count();
  grid = grid < cap ? grid : cap;
  if (prolong)
    hipLaunchKernelGGL((prolong_add_kernel<T>), dim3(grid), dim3(256), 0, s, (const T*)L.P,
                       (const T*)src, (T*)dst, Ku, nhi, nlo, skip);
  else
    hipLaunchKernelGGL((restrict_kernel<T>), dim3(grid), dim3(256), 0, s, (const T*)L.P,
                       (const T*)src, (T*)dst, Ku, nhi, nlo, skip, (const T*)C.binv, (T)omega,
                       (T*)x_c);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

// The parent's p-prolongation handed to a coarser level's cycle: its last
// fused sweep also adds P x (P: np x Np of this level) into the parent's
// iterate, and reports that it did.
struct ProlongTo {
  void* x;
  const void* P;
  int np;
  bool done;
};

// V-cycle on level l; first_done: the level's first sweep (omega Binv b) is
// already in its first-sweep buffer (written by the parent's restriction);
// pt: the parent's p-prolongation, done here when the last sweep is fused
template <typename T>
static int vcycle(const MG& mg, MGWork& w, int l, const void* b, void* x, const int* skip,
                  hipStream_t s, bool first_done = false, ProlongTo* pt = nullptr) {
  if (l == mg.nlev - 1) return coarse_solve<T>(mg, w, b, x, skip, s);
  const MGLevel& L = mg.lv[l];
  const MGLevel& C = mg.lv[l + 1];
  const Op& op = *L.op;
  void* Lr = w.r[l];
  void* Cb = w.b[l + 1];
  void* Cx = w.x[l + 1];
  auto resid = [&]() {
    return launch_apply(op, x, Lr, nullptr, mg.alpha, 0.0, M_RESID, nullptr, skip, s, b);
  };
  if (mg.nu > 0 && smooth_fused(op)) {
    // fused sweeps x_out = x_in + omega Binv (b - A x_in), alternating between
    // x and the residual buffer: the first sweep (x = omega Binv b) lands
    // where nu - 1 fused sweeps bring the iterate back to x, the residual of
    // the result goes to Lr for the restriction, and the post-smoother runs
    // its sweeps in pairs x -> Lr -> x (an odd last one as residual + sweep).
    void* bufs[2] = {x, Lr};
    int at = (mg.nu - 1) % 2;  // buffer of the first sweep's output
    if (!first_done) MG_CHECK(bj_smooth<T>(L, b, bufs[at], mg.omega, 0, skip, s));
    for (int i = 1; i < mg.nu; ++i) {
      MG_CHECK(launch_smooth(op, L.binv, mg.omega, mg.alpha, bufs[at], bufs[1 - at], b, skip, s));
      at = 1 - at;
    }
    // a smoothed coarser level gets its first sweep from the restriction
    const bool fuse_first = l + 2 < mg.nlev && smooth_fused(*C.op);
    void* c_first = (mg.nu - 1) % 2 ? w.r[l + 1] : Cx;
    // the residual and its restriction in one launch where that kernel is
    // built (the fine residual is then never stored), else two
    const int rr = launch_resid_restrict(op, L.tkind, C.op->Np, L.P, x, b, mg.alpha, Cb, C.binv,
                                         mg.omega, fuse_first ? c_first : nullptr, skip, s);
    if (rr == DDGX_ERR_UNSUPPORTED) {
      MG_CHECK(resid());
      MG_CHECK(transfer<T>(L, C, false, Lr, Cb, skip, s, fuse_first ? c_first : nullptr,
                           mg.omega));
    } else {
      MG_CHECK(rr);
    }
    // a p-prolongation is added to x by the coarser level's last fused sweep
    ProlongTo pc{x, L.P, op.Np, false};
    MG_CHECK(vcycle<T>(mg, w, l + 1, Cb, Cx, skip, s, fuse_first, L.tkind == 0 ? &pc : nullptr));
    // the h-prolongation into an order-1 level is formed inside the first
    // post-smoothing sweep (which never stores the corrected iterate)
    const bool fuse_prolong = mg.nu >= 2 && L.tkind == 1 && smooth_prolongs(op);
    if (!fuse_prolong && !pc.done) MG_CHECK(transfer<T>(L, C, true, Cx, x, skip, s));
    for (int i = 0; i + 1 < mg.nu; i += 2) {
      const bool pf = fuse_prolong && i == 0;
      MG_CHECK(launch_smooth(op, L.binv, mg.omega, mg.alpha, x, Lr, b, skip, s,
                             pf ? Cx : nullptr, pf ? L.P : nullptr));
      // the last sweep of an even count takes the parent's p-prolongation
      const bool po = pt != nullptr && i + 2 == mg.nu;
      MG_CHECK(launch_smooth(op, L.binv, mg.omega, mg.alpha, Lr, x, b, skip, s, nullptr,
                             nullptr, po ? pt->x : nullptr, po ? pt->P : nullptr,
                             po ? pt->np : 0));
      if (po) pt->done = true;
    }
    if (mg.nu % 2) {
      MG_CHECK(resid());
      MG_CHECK(bj_smooth<T>(L, Lr, x, mg.omega, 1, skip, s));
    }
    return DDGX_OK;
  }
  if (mg.nu == 0) {
    const size_t n = (size_t)op.K * op.Np;
    hipLaunchKernelGGL((zero_kernel<T>), dim3(vec_grid(n)), dim3(256), 0, s, (T*)x, n, skip);
    DDGX_HIP(hipGetLastError());
    MG_CHECK(transfer<T>(L, C, false, b, Cb, skip, s));
  } else {
    MG_CHECK(bj_smooth<T>(L, b, x, mg.omega, 0, skip, s));
    for (int i = 1; i < mg.nu; ++i) {
      MG_CHECK(resid());
      MG_CHECK(bj_smooth<T>(L, Lr, x, mg.omega, 1, skip, s));
    }
    MG_CHECK(resid());
    MG_CHECK(transfer<T>(L, C, false, Lr, Cb, skip, s));
  }
  MG_CHECK(vcycle<T>(mg, w, l + 1, Cb, Cx, skip, s));
  MG_CHECK(transfer<T>(L, C, true, Cx, x, skip, s));
  for (int i = 0; i < mg.nu; ++i) {
    MG_CHECK(resid());
    MG_CHECK(bj_smooth<T>(L, Lr, x, mg.omega, 1, skip, s));
  }
  return DDGX_OK;
}

int mg_vcycle(const MG& mg, MGWork& w, const void* b, void* x, const int* skip, hipStream_t s) {
  if (w.nlev != mg.nlev) {
    set_error("multigrid workspace does not match the hierarchy");
    return DDGX_ERR_INVALID;
  }
  if (mg.dtype == DDGX_F64) return vcycle<double>(mg, w, 0, b, x, skip, s);
  return vcycle<float>(mg, w, 0, b, x, skip, s);
}

void mgwork_free(MGWork* w) {
  if (w == nullptr) return;
  for (int l = 0; l < w->nlev; ++l) {
    if (w->b[l]) (void)hipFree(w->b[l]);
    if (w->x[l]) (void)hipFree(w->x[l]);
    if (w->r[l]) (void)hipFree(w->r[l]);
  }
  if (w->gemv_part) (void)hipFree(w->gemv_part);
  delete w;
}

// Level vectors of one cycle: b and x on levels 1.., r on every level but
// the coarsest, and the coarse-solve partial sums.
int mgwork_create(const MG& mg, MGWork** out) {
  *out = nullptr;
  MGWork* w = new MGWork();
  w->nlev = mg.nlev;
  hipError_t e = hipSuccess;
  for (int l = 0; l < mg.nlev && e == hipSuccess; ++l) {
    const size_t nv = (size_t)mg.lv[l].op->K * mg.lv[l].op->Np * dsz(mg.dtype);
    if (l > 0) {
      e = hipMalloc(&w->b[l], nv);
      if (e == hipSuccess) e = hipMalloc(&w->x[l], nv);
      w->bytes += 2 * nv;
    }
    if (e == hipSuccess && l + 1 < mg.nlev) {
      e = hipMalloc(&w->r[l], nv);
      w->bytes += nv;
    }
  }
  const int jc = gemv_jc(mg.nc);
  const int nch = (mg.nc + jc - 1) / jc;
  const size_t ng = (size_t)nch * mg.nc * mg.nseg * sizeof(double);
  if (e == hipSuccess) e = hipMalloc((void**)&w->gemv_part, ng);
  w->bytes += ng;
  if (e != hipSuccess) {
    mgwork_free(w);
    return hip_fail(e, "hipMalloc(multigrid cycle workspace)");
  }
  *out = w;
  return DDGX_OK;
}

template <typename T>
static int upload_as(const double* src, size_t n, void** dst) {
  std::vector<T> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
  DDGX_HIP(hipMalloc(dst, n * sizeof(T)));
  DDGX_HIP(hipMemcpy(*dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return DDGX_OK;
}

void mg_free(MG* mg) {
  if (mg == nullptr) return;
  for (int l = 0; l < mg->nlev; ++l) {
    MGLevel& L = mg->lv[l];
    if (L.binv) (void)hipFree(L.binv);
    if (L.P) (void)hipFree(L.P);
    if (mg->owns_ops && L.op) {
      op_release(L.op);
      delete L.op;
    }
  }
  delete[] mg->lv;
  if (mg->Ainv) (void)hipFree(mg->Ainv);
  mgwork_free(mg->ws);
  delete mg;
}

// Dense coarse matrix (column j = (A + alpha M) e_j, stored as row j) and its
// in-place inverse.
static int build_coarse(MG& mg, const Op& op64, double alpha, int regularize, double eps,
                        hipStream_t s) {
  const int n = mg.nc;
  double *A = nullptr, *e = nullptr, *prow = nullptr, *col = nullptr;
  DDGX_HIP(hipMalloc((void**)&A, (size_t)n * n * sizeof(double)));
  mg.Ainv = A;
  int rc = DDGX_OK;
  do {
    hipError_t he = hipMalloc((void**)&e, (size_t)n * sizeof(double));
    if (he == hipSuccess) he = hipMalloc((void**)&prow, (size_t)n * sizeof(double));
    if (he == hipSuccess) he = hipMalloc((void**)&col, (size_t)n * sizeof(double));
    if (he == hipSuccess) he = hipMemsetAsync(e, 0, (size_t)n * sizeof(double), s);
    if (he != hipSuccess) { rc = hip_fail(he, "coarse-level workspace"); break; }
    for (int j = 0; j < n && rc == DDGX_OK; ++j) {
      hipLaunchKernelGGL(unit_kernel, dim3(1), dim3(64), 0, s, e, j, j - 1);
      rc = launch_apply(op64, e, A + (size_t)j * n, nullptr, alpha, 0.0, 0, nullptr, nullptr, s);
    }
    if (rc != DDGX_OK) break;
    if (regularize) hipLaunchKernelGGL(regularize_kernel, dim3(1), dim3(256), 0, s, A, n, eps);
    const int g = (n + 255) / 256;
    for (int k = 0; k < n; ++k) {
      hipLaunchKernelGGL(gj_pivot_kernel, dim3(g), dim3(256), 0, s, A, n, k, prow, col);
      hipLaunchKernelGGL(gj_update_kernel, dim3(g, n), dim3(256), 0, s, A, n, k, prow, col);
    }
    hipLaunchKernelGGL(diag_kernel, dim3(g), dim3(256), 0, s, A, n, col);
    he = hipGetLastError();
    if (he != hipSuccess) { rc = hip_fail(he, "coarse-level inversion"); break; }
    std::vector<double> d(n);
    he = hipMemcpyAsync(d.data(), col, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) { rc = hip_fail(he, "coarse-level inversion"); break; }
    for (int i = 0; i < n; ++i) {
      if (!std::isfinite(d[i]) || !(d[i] > 0.0)) {
        set_error("coarsest-level matrix is singular or indefinite (a pure-Neumann operator "
                  "needs regularize=True)");
        rc = DDGX_ERR_INVALID;
        break;
      }
    }
  } while (false);
  if (e) (void)hipFree(e);
  if (prow) (void)hipFree(prow);
  if (col) (void)hipFree(col);
  return rc;
}

int mg_create(int nlev, Op* const* ops, Op* const* ops64, const int* tkind,
              const double* const* P, const ddgx_mg_cfg& cfg, hipStream_t s, MG** out) {
  if (nlev < 1 || nlev > 16) {
    set_error("multigrid needs 1..16 levels");
    return DDGX_ERR_INVALID;
  }
  const int dtype = ops[0]->dtype;
  for (int l = 0; l < nlev; ++l) {
    if (ops[l] == nullptr || ops64[l] == nullptr) {
      set_error("null level operator");
      return DDGX_ERR_INVALID;
    }
    if (ops[l]->dtype != dtype || ops64[l]->dtype != DDGX_F64 || ops64[l]->N != ops[l]->N ||
        ops64[l]->K != ops[l]->K) {
      set_error("level " + std::to_string(l) +
                ": operators must share one dtype, and ops64 must be the float64 twin");
      return DDGX_ERR_INVALID;
    }
    if (l + 1 < nlev) {
      const Op& a = *ops[l];
      const Op& b = *ops[l + 1];
      if (tkind[l] == 0 && !(a.K == b.K && b.N <= a.N)) {
        set_error("level " + std::to_string(l) + ": p-transfer needs the same mesh, lower order");
        return DDGX_ERR_INVALID;
      }
      if (tkind[l] == 1 && !(a.K == 4 * b.K && a.N == b.N)) {
        set_error("level " + std::to_string(l) +
                  ": h-transfer needs 4x the elements of the next level at the same order");
        return DDGX_ERR_INVALID;
      }
      if (tkind[l] != 0 && tkind[l] != 1) {
        set_error("transfer kind must be 0 (p) or 1 (h)");
        return DDGX_ERR_INVALID;
      }
      if (P == nullptr || P[l] == nullptr) {
        set_error("missing transfer matrix");
        return DDGX_ERR_INVALID;
      }
    }
  }
  if (!(cfg.nu_smooth >= 0 && cfg.nu_smooth <= 64) || !std::isfinite(cfg.omega)) {
    set_error("bad smoother settings");
    return DDGX_ERR_INVALID;
  }
  const int64_t nc = ops[nlev - 1]->K * ops[nlev - 1]->Np;
  if (nc > MG_MAX_COARSE) {
    set_error("coarsest level has " + std::to_string(nc) + " unknowns; the dense direct solve "
              "takes at most " + std::to_string(MG_MAX_COARSE) + " (add h-levels)");
    return DDGX_ERR_INVALID;
  }
  DDGX_HIP(hipSetDevice(ops[0]->device));
  MG* mg = new MG();
  mg->dtype = dtype;
  mg->nu = cfg.nu_smooth;
  mg->omega = cfg.omega;
  mg->alpha = cfg.alpha;
  mg->nlev = nlev;
  mg->nc = (int)nc;
  mg->lv = new MGLevel[nlev];
  int rc = DDGX_OK;
  double* blocks = nullptr;
  do {
    const size_t nb0 = (size_t)ops64[0]->K * ops64[0]->Np * ops64[0]->Np;
    if (nlev > 1) {
      hipError_t he = hipMalloc((void**)&blocks, nb0 * sizeof(double));
      if (he != hipSuccess) { rc = hip_fail(he, "hipMalloc(blocks)"); break; }
    }
    for (int l = 0; l < nlev && rc == DDGX_OK; ++l) {
      MGLevel& L = mg->lv[l];
      L.op = ops[l];
      L.alpha = cfg.alpha;
      if (l + 1 == nlev) break;
      L.full = ops[l]->cv != nullptr;  // a curved level's blocks are not symmetric
      const size_t nb = (size_t)ops[l]->K * block_words(ops[l]->N, L.full);
      hipError_t he = hipMalloc(&L.binv, nb * dsz(dtype));
      if (he != hipSuccess) { rc = hip_fail(he, "hipMalloc(level)"); break; }
      rc = ops64[l]->cv != nullptr ? launch_block_probe_general(*ops64[l], blocks, cfg.alpha, s)
                                   : launch_block_probe(*ops64[l], blocks, cfg.alpha, s);
      if (rc == DDGX_OK)
        rc = launch_invert_blocks(ops[l]->N, ops[l]->K, dtype, blocks, L.binv, s, L.full);
      if (rc != DDGX_OK) break;
      he = hipStreamSynchronize(s);  // blocks is reused by the next level
      if (he != hipSuccess) { rc = hip_fail(he, "block-Jacobi build"); break; }
      L.tkind = tkind[l];
      const Op& C = *ops[l + 1];
      const int nlo = C.Np;
      const int nhi = (int)((ops[l]->K * ops[l]->Np) / C.K);
      L.np_lo = nlo;
      if (nhi * nlo > MG_MAXP || (256 / nlo) * nhi > MG_MAXIN || nhi > 256) {
        set_error("transfer matrix too large");
        rc = DDGX_ERR_UNSUPPORTED;
        break;
      }
      rc = dtype == DDGX_F64 ? upload_as<double>(P[l], (size_t)nhi * nlo, &L.P)
                             : upload_as<float>(P[l], (size_t)nhi * nlo, &L.P);
    }
    if (rc != DDGX_OK) break;
    if (blocks) {
      (void)hipFree(blocks);
      blocks = nullptr;
    }
    rc = mgwork_create(*mg, &mg->ws);
    if (rc != DDGX_OK) break;
    rc = build_coarse(*mg, *ops64[nlev - 1], cfg.alpha, cfg.regularize, cfg.eps, s);
  } while (false);
  if (blocks) (void)hipFree(blocks);
  if (rc != DDGX_OK) {
    (void)hipStreamSynchronize(s);
    mg_free(mg);
    return rc;
  }
  *out = mg;
  return DDGX_OK;
}

}  // namespace ddgx
