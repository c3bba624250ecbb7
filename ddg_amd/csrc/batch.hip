// Batched solves: nseg requests of one problem class -- the same mesh, order
// and preconditioner, each with its own right-hand side and Helmholtz
// coefficient -- advanced together, one launch per kernel for all of them.
// The reference solves one problem per call (src/ddg/dg2d/sipg.py:536-612);
// batching is how a serving loop keeps small problems from paying one GPU
// launch latency per kernel each.
//
// The batch operator is nseg disjoint copies of the class operator: its
// element records, face codes, halo lists and Neumann data are replicated on
// the device with element indices offset by s*K for copy s, so every kernel
// of the single-problem path runs unchanged on the batch, and copy s sees
// exactly the data (and tile structure: K is a multiple of the 64-element
// tile) of the class operator.  Per-request Helmholtz coefficients go into
// the Jacobian record word of each copy (alpha_s * J, applied with alpha = 1:
// the same product the single apply forms).  The multigrid hierarchy is
// replicated the same way, level by level, with one coarse inverse applied
// to every segment.

#include <vector>

#include "ddgx_internal.h"

namespace ddgx {

static size_t dsz_b(int dtype) { return dtype == DDGX_F64 ? 8 : 4; }

// dst[s*n + i] = src[i] + (src[i] >= 0 ? s * add : 0)   (s = blockIdx.y)
__global__ void __launch_bounds__(256)
replicate_codes_kernel(const int* __restrict__ src, int* __restrict__ dst, int64_t n, int add) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int v = src[i];
  dst[(int64_t)blockIdx.y * n + i] = v >= 0 ? v + (int)blockIdx.y * add : v;
}

// rec plane R_J of the batch operator: segment s = alpha[s] * j1
template <typename T>
__global__ void __launch_bounds__(256)
scale_jacobian_kernel(T* __restrict__ plane, const T* __restrict__ j1,
                      const double* __restrict__ alpha, int64_t k1) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= k1) return;
  plane[(int64_t)blockIdx.y * k1 + i] = (T)(alpha[blockIdx.y] * (double)j1[i]);
}

static int grid_of(int64_t n) { return (int)((n + 255) / 256); }

// nseg copies of `bytes` at src into dst (contiguous)
static int copy_rep(void* dst, const void* src, size_t bytes, int nseg, hipStream_t s) {
  for (int c = 0; c < nseg; ++c)
    DDGX_HIP(hipMemcpyAsync((char*)dst + (size_t)c * bytes, src, bytes,
                            hipMemcpyDeviceToDevice, s));
  return DDGX_OK;
}

// planes x [nseg copies of one plane of n items]
static int copy_planes(void* dst, const void* src, int planes, size_t plane_bytes, int nseg,
                       hipStream_t s) {
  for (int w = 0; w < planes; ++w)
    for (int c = 0; c < nseg; ++c)
      DDGX_HIP(hipMemcpyAsync((char*)dst + ((size_t)w * nseg + c) * plane_bytes,
                              (const char*)src + (size_t)w * plane_bytes, plane_bytes,
                              hipMemcpyDeviceToDevice, s));
  return DDGX_OK;
}

// planes x [nseg copies of one plane of n codes, element indices offset]
static int codes_planes(int* dst, const int* src, int planes, int64_t n, int add, int nseg,
                        hipStream_t s) {
  for (int w = 0; w < planes; ++w)
    hipLaunchKernelGGL(replicate_codes_kernel, dim3(grid_of(n), nseg), dim3(256), 0, s,
                       src + (size_t)w * n, dst + (size_t)w * n * nseg, n, add);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

int op_replicate(const Op& src, int nseg, hipStream_t s, Op** out) {
  *out = nullptr;
  if (src.cv != nullptr) {
    set_error("batched solves take affine operators only");
    return DDGX_ERR_UNSUPPORTED;
  }
  if (src.K % TILE != 0) {
    set_error("batched solves need a multiple of 64 elements per problem (whole tiles)");
    return DDGX_ERR_INVALID;
  }
  if (nseg < 1 || src.K * nseg >= (int64_t(1) << 28) ||
      src.K * nseg * src.Np * 8 > MAX_VECTOR_BYTES) {
    set_error("batch size out of range (K * Np * 8 bytes of a batched vector below 4 GiB)");
    return DDGX_ERR_INVALID;
  }
  const int64_t K1 = src.K, KB = K1 * nseg, T1 = K1 / TILE;
  const size_t ds = dsz_b(src.dtype);
  Op* o = new Op();
  o->N = src.N;
  o->dtype = src.dtype;
  o->device = src.device;
  o->K = KB;
  o->Np = src.Np;
  o->Nfp = src.Nfp;
  o->tab = src.tab;
  o->frag = src.frag;
  o->owns_tab = false;
  o->perm_id = src.perm_id;
  o->max_halo = src.max_halo;
  o->halo_cap = src.halo_cap;
  o->nseg = nseg;
  int rc = DDGX_OK;
  do {
    hipError_t e = hipMalloc(&o->rec, (size_t)DDGX_REC_WORDS * KB * ds);
    if (e == hipSuccess) e = hipMalloc(&o->fsj, (size_t)4 * KB * ds);
    if (e == hipSuccess) e = hipMalloc(&o->hab, (size_t)T1 * nseg * 6 * src.halo_cap * ds);
    if (e == hipSuccess) e = hipMalloc((void**)&o->nbr, (size_t)3 * KB * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&o->nbrp, (size_t)3 * KB * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&o->fcode, (size_t)3 * KB * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&o->perm, (size_t)KB * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&o->hpos, (size_t)T1 * nseg * src.halo_cap * sizeof(int));
    if (e != hipSuccess) { rc = hip_fail(e, "batch operator allocation"); break; }
    if ((rc = copy_planes(o->rec, src.rec, DDGX_REC_WORDS, (size_t)K1 * ds, nseg, s))) break;
    if ((rc = copy_rep(o->fsj, src.fsj, (size_t)4 * K1 * ds, nseg, s))) break;
    if ((rc = copy_rep(o->hab, src.hab, (size_t)T1 * 6 * src.halo_cap * ds, nseg, s))) break;
    // the apply's packed codes hold tile-local slots only: copied as they are
    if ((rc = copy_planes(o->fcode, src.fcode, 3, (size_t)K1 * sizeof(int), nseg, s))) break;
    // neighbour codes carry the element (or tile position) in bits 3..
    if ((rc = codes_planes(o->nbr, src.nbr, 3, K1, (int)(K1 << 3), nseg, s))) break;
    if ((rc = codes_planes(o->nbrp, src.nbrp, 3, K1, (int)(K1 << 3), nseg, s))) break;
    if ((rc = codes_planes(o->perm, src.perm, 1, K1, (int)K1, nseg, s))) break;
    if ((rc = codes_planes(o->hpos, src.hpos, 1, T1 * src.halo_cap, (int)K1, nseg, s))) break;
    o->grid_apply = (int)(T1 * nseg);
    const int64_t gv = (KB * src.Np + 255) / 256;
    const int64_t cap = 4 * (int64_t)device_cu_count();
    o->grid_vec = (int)(gv < cap ? gv : cap);
  } while (false);
  if (rc != DDGX_OK) {
    (void)hipStreamSynchronize(s);
    op_release(o);
    delete o;
    return rc;
  }
  *out = o;
  return DDGX_OK;
}

int mg_replicate(const MG& src, int nseg, hipStream_t s, MG** out) {
  *out = nullptr;
  MG* mg = new MG();
  mg->dtype = src.dtype;
  mg->nu = src.nu;
  mg->omega = src.omega;
  mg->alpha = src.alpha;
  mg->nlev = src.nlev;
  mg->nc = src.nc;
  mg->nseg = nseg;
  mg->owns_ops = true;
  mg->lv = new MGLevel[src.nlev];
  const size_t ds = dsz_b(src.dtype);
  int rc = DDGX_OK;
  do {
    for (int l = 0; l < src.nlev && rc == DDGX_OK; ++l) {
      const MGLevel& S = src.lv[l];
      MGLevel& L = mg->lv[l];
      rc = op_replicate(*S.op, nseg, s, &L.op);
      if (rc != DDGX_OK) break;
      L.alpha = S.alpha;
      L.tkind = S.tkind;
      L.np_lo = S.np_lo;
      if (S.binv != nullptr) {
        L.full = S.full;
        const size_t nb = (size_t)S.op->K * block_words(S.op->N, S.full) * ds;
        hipError_t e = hipMalloc(&L.binv, nb * nseg);
        if (e != hipSuccess) { rc = hip_fail(e, "batch multigrid blocks"); break; }
        rc = copy_rep(L.binv, S.binv, nb, nseg, s);
        if (rc != DDGX_OK) break;
      }
      if (S.P != nullptr && l + 1 < src.nlev) {
        const Op& C = *src.lv[l + 1].op;
        const size_t nhi = (size_t)(S.op->K * S.op->Np) / C.K;
        const size_t words = nhi * S.np_lo;
        hipError_t e = hipMalloc(&L.P, words * ds);
        if (e == hipSuccess)
          e = hipMemcpyAsync(L.P, S.P, words * ds, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) { rc = hip_fail(e, "batch transfer matrices"); break; }
      }
    }
    if (rc != DDGX_OK) break;
    const size_t na = (size_t)src.nc * src.nc * sizeof(double);
    hipError_t e = hipMalloc((void**)&mg->Ainv, na);
    if (e == hipSuccess) e = hipMemcpyAsync(mg->Ainv, src.Ainv, na, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) { rc = hip_fail(e, "batch coarse inverse"); break; }
    rc = mgwork_create(*mg, &mg->ws);
  } while (false);
  if (rc != DDGX_OK) {
    (void)hipStreamSynchronize(s);
    mg_free(mg);
    return rc;
  }
  *out = mg;
  return DDGX_OK;
}

void batch_free(Batch* b) {
  if (b == nullptr) return;
  if (b->chunk) (void)hipGraphExecDestroy(b->chunk);
  void* ptrs[] = {b->j1, b->alpha, b->r, b->p, b->z, b->Ap, b->pc_r, b->pc_z, b->partials,
                  b->apart, b->scal};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (b->alpha_host) (void)hipHostFree(b->alpha_host);
  if (b->scal_host) (void)hipHostFree(b->scal_host);
  if (b->pc) {
    mg_free(b->pc->mg);  // frees its replicated level operators (and b->mgw = its ws)
    delete b->pc;
  }
  if (b->op) {
    op_release(b->op);
    delete b->op;
  }
  delete b;
}

int batch_create(const Op& op, const PC* pc, int nseg, hipStream_t s, Batch** out) {
  *out = nullptr;
  if (op.dtype != DDGX_F64) {
    set_error("batched solves take a float64 operator");
    return DDGX_ERR_INVALID;
  }
  if (pc != nullptr && pc->type != 1) {
    set_error("batched solves take a multigrid preconditioner or none");
    return DDGX_ERR_UNSUPPORTED;
  }
  if (pc != nullptr && (pc->op->K != op.K || pc->op->N != op.N)) {
    set_error("preconditioner was built for a different mesh or order");
    return DDGX_ERR_INVALID;
  }
  if (nseg < 1 || nseg > 64) {
    set_error("batch size must be 1..64");
    return DDGX_ERR_INVALID;
  }
  DDGX_HIP(hipSetDevice(op.device));
  Batch* b = new Batch();
  b->nseg = nseg;
  b->n1 = op.K * op.Np;
  b->g1 = op.grid_vec;
  b->t1 = op.grid_apply;
  int rc = DDGX_OK;
  do {
    if ((rc = op_replicate(op, nseg, s, &b->op))) break;
    if (pc != nullptr) {
      b->pc = new PC();
      b->pc->type = 1;
      b->pc->dtype = pc->dtype;
      b->pc->K = pc->K * nseg;
      b->pc->N = pc->N;
      b->pc->op = b->op;
      if ((rc = mg_replicate(*pc->mg, nseg, s, &b->pc->mg))) break;
      b->mgw = b->pc->mg->ws;
    }
    const size_t nv = (size_t)b->n1 * nseg * sizeof(double);
    const size_t k1b = (size_t)op.K * sizeof(double);
    hipError_t e = hipMalloc(&b->j1, k1b);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->j1, (const char*)op.rec + (size_t)R_J * k1b, k1b,
                         hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMalloc((void**)&b->alpha, nseg * sizeof(double));
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&b->alpha_host, nseg * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&b->r, nv);
    if (e == hipSuccess) e = hipMalloc(&b->p, nv);
    if (e == hipSuccess) e = hipMalloc(&b->z, nv);
    if (e == hipSuccess) e = hipMalloc(&b->Ap, nv);
    if (e == hipSuccess && pc != nullptr && pc->dtype == DDGX_F32) {
      e = hipMalloc(&b->pc_r, nv / 2);
      if (e == hipSuccess) e = hipMalloc(&b->pc_z, nv / 2);
    }
    if (e == hipSuccess)
      e = hipMalloc((void**)&b->partials, (size_t)nseg * 2 * b->g1 * sizeof(double));
    if (e == hipSuccess)
      e = hipMalloc((void**)&b->apart, (size_t)nseg * b->t1 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->scal, nseg * sizeof(CGScalars));
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&b->scal_host, nseg * sizeof(CGScalars), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { rc = hip_fail(e, "batch workspace"); break; }
    b->bytes = 4 * nv + (b->pc_r ? nv : 0) + (size_t)nseg * (2 * b->g1 + b->t1) * 8 +
               nseg * sizeof(CGScalars) + k1b +
               (size_t)DDGX_REC_WORDS * op.K * nseg * 8 + (size_t)op.K * nseg * 4 * 8;
    if (b->pc) {
      const MG& m = *b->pc->mg;
      for (int l = 0; l < m.nlev; ++l)
        b->bytes += (size_t)m.lv[l].op->K * (block_words(m.lv[l].op->N, m.lv[l].full) +
                                             DDGX_REC_WORDS + 4) * dsz_b(m.dtype);
      b->bytes += m.ws->bytes + (size_t)m.nc * m.nc * 8;
    }
  } while (false);
  if (rc != DDGX_OK) {
    (void)hipStreamSynchronize(s);
    batch_free(b);
    return rc;
  }
  *out = b;
  return DDGX_OK;
}

int launch_scale_jacobian(Batch& bt, hipStream_t s) {
  const int64_t k1 = bt.op->K / bt.nseg;
  DDGX_HIP(hipMemcpyAsync(bt.alpha, bt.alpha_host, bt.nseg * sizeof(double),
                          hipMemcpyHostToDevice, s));
  double* plane = (double*)bt.op->rec + (size_t)R_J * bt.op->K;
  hipLaunchKernelGGL((scale_jacobian_kernel<double>), dim3(grid_of(k1), bt.nseg), dim3(256), 0,
                     s, plane, (const double*)bt.j1, (const double*)bt.alpha, k1);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

}  // namespace ddgx
