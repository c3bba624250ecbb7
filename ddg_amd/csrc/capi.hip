// C ABI of libddgx (declarations and reference correspondences in
// include/ddgx.h).  Handles own their device memory; every compute entry is
// stream-ordered on the caller's stream.

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ddgx_internal.h"
#include "sipg_apply_mfma.h"

struct ddgx_op_s {
  ddgx::Op op;
};
struct ddgx_pc_s {
  ddgx::PC pc;
};
struct ddgx_batch_s {
  ddgx::Batch* b = nullptr;
  int busy = 0;  // a solve is running on this context (Claim)
};
struct ddgx_solver_s {
  ddgx::Op* op = nullptr;     // operator of the solve (float64 for refinement)
  ddgx::Op* op32 = nullptr;   // float32 twin (refinement solver) or null
  ddgx::PC* pc = nullptr;     // preconditioner or null
  ddgx::Work* w = nullptr;    // Krylov state of op
  ddgx::Work* w32 = nullptr;  // Krylov state of op32
  ddgx::MGWork* mgw = nullptr;  // V-cycle buffers when pc is a multigrid
  int busy = 0;               // a solve is running on this context (Claim)
};

namespace ddgx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? DDGX_ERR_NOMEM : DDGX_ERR_HIP;
}

// Compute units of the current device, queried once per device (launch code
// runs inside stream capture, where device queries are best avoided).
int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    return 256;
  if (dev >= 0 && dev < 64) cached[dev] = cus;
  return cus;
}

static int valu_table_size(int N) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  const int RS = (Np + 7) / 8 * 8, RSF = (Nfp + 7) / 8 * 8;
  return 5 * Np * RS + 3 * Nfp * RSF + 6 * Nfp * RS;
}

static int frag_size(int N) {
  const int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1;
  // face k-steps: 12 when the face rows are formed in registers (one
  // 16-row block), else the 9 Nfp rows four per k-step
  const int NB = (Np + 15) / 16, KB = 4 * NB, FK = NB == 1 ? 12 : (9 * Nfp + 3) / 4;
  return 64 * (2 * NB * KB + NB * 2 * KB + NB * FK + NB * KB);
}

// host table: column tables followed by one copy of the 64-lane MFMA fragments
static int table_size(int N) { return valu_table_size(N) + frag_size(N); }

template <typename T>
static int upload(const double* src, size_t n, void** dst) {
  std::vector<T> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
  DDGX_HIP(hipMalloc(dst, n * sizeof(T)));
  DDGX_HIP(hipMemcpy(*dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return DDGX_OK;
}

static int upload_as(int dtype, const double* src, size_t n, void** dst) {
  return dtype == DDGX_F64 ? upload<double>(src, n, dst) : upload<float>(src, n, dst);
}

void op_release(Op* op) {
  Op& o = *op;
  void* ptrs[] = {o.owns_tab ? o.tab : nullptr, o.rec, o.nbr, o.nbrp, o.fcode, o.perm, o.hpos,
                  o.hab, o.fsj};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  work_free(o.ws);
  curved_free(o.cv);
  o = Op{};
}

static void free_op(Op& o) { op_release(&o); }

}  // namespace ddgx

using namespace ddgx;

extern "C" {

const char* ddgx_last_error(void) { return g_err.c_str(); }

int ddgx_table_size(int N) { return table_size(N); }

int ddgx_op_create(int N, int64_t K, int dtype, const double* tables, const double* rec,
                   const int32_t* nbr, const double* fsj, const int32_t* order, ddgx_op* out) {
  if (out == nullptr || tables == nullptr || rec == nullptr || nbr == nullptr || fsj == nullptr) {
    set_error("ddgx_op_create: null argument");
    return DDGX_ERR_INVALID;
  }
  *out = nullptr;
  if (N < 1 || N > 6) {
    set_error("ddgx_op_create: polynomial order " + std::to_string(N) + " not built (1..6)");
    return DDGX_ERR_UNSUPPORTED;
  }
  if (dtype != DDGX_F64 && dtype != DDGX_F32) {
    set_error("ddgx_op_create: dtype must be DDGX_F64 or DDGX_F32");
    return DDGX_ERR_INVALID;
  }
  if (K < 1 || K >= (int64_t(1) << 28)) {
    set_error("ddgx_op_create: element count out of range [1, 2^28)");
    return DDGX_ERR_INVALID;
  }
  if (K * ((N + 1) * (N + 2) / 2) * 8 > MAX_VECTOR_BYTES) {
    set_error("ddgx_op_create: a nodal vector of K * Np float64 values must stay below 4 GiB");
    return DDGX_ERR_INVALID;
  }
  // validate neighbour codes on the host: they index device memory
  for (int64_t k = 0; k < K; ++k)
    for (int f = 0; f < 3; ++f) {
      const int c = nbr[k * 4 + f];
      if (c >= 0 && (nbr_elem(c) >= K || nbr_face(c) > 2)) {
        set_error("ddgx_op_create: face code out of range at element " + std::to_string(k));
        return DDGX_ERR_INVALID;
      }
      if (c < 0 && c != CODE_DIRICHLET && c != CODE_NONE) {
        set_error("ddgx_op_create: invalid negative face code");
        return DDGX_ERR_INVALID;
      }
    }
  // tile order: position i holds element perm[i]
  std::vector<int> perm(K), inv(K, -1);
  for (int64_t i = 0; i < K; ++i) {
    const int64_t k = order ? order[i] : i;
    if (k < 0 || k >= K || inv[k] >= 0) {
      set_error("ddgx_op_create: order is not a permutation of 0..K-1");
      return DDGX_ERR_INVALID;
    }
    perm[i] = (int)k;
    inv[k] = (int)i;
  }
  // per-tile halo lists (tile positions of out-of-tile neighbour elements)
  // and the packed apply code of every face (neighbour trace slot: 0..63 in
  // the tile, 64+h in the halo)
  const int64_t ntiles = (K + TILE - 1) / TILE;
  std::vector<int> hall(ntiles * HALO_MAX, -1), fcode((size_t)K * 3);
  int max_halo = 0;
  bool perm_id = true;
  for (int64_t i = 0; i < K; ++i) perm_id = perm_id && perm[i] == i;
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t i0 = t * TILE, i1 = std::min<int64_t>(K, i0 + TILE);
    int nh = 0;
    for (int64_t i = i0; i < i1; ++i) {
      for (int f = 0; f < 3; ++f) {
        const int c = nbr[(int64_t)perm[i] * 4 + f];
        int code = fc_pack(0, 0, 0, c == CODE_DIRICHLET ? FT_DIRICHLET : FT_NONE);
        if (c >= 0) {
          const int in = inv[nbr_elem(c)];
          int sl;
          if (in >= i0 && in < i1) {
            sl = (int)(in - i0);
          } else {
            int h = 0;
            while (h < nh && hall[t * HALO_MAX + h] != in) ++h;
            if (h == nh) {
              if (nh == HALO_MAX) {
                set_error("ddgx_op_create: a 64-element tile has more than 64 out-of-tile "
                          "neighbours; pass an element order that keeps tiles compact");
                return DDGX_ERR_INVALID;
              }
              hall[t * HALO_MAX + nh] = in;
              ++nh;
            }
            sl = TILE + h;
          }
          code = fc_pack(sl, nbr_face(c), nbr_rev(c), FT_INTERIOR);
        }
        fcode[(size_t)f * K + i] = code;
      }
    }
    max_halo = std::max(max_halo, nh);
  }
  // the smaller-LDS kernel variant when every halo fits its capacity
  const int hcap = max_halo <= HALO_SMALL ? HALO_SMALL : HALO_MAX;
  // the neighbour read of every face as an LDS address (fc_nbase, fc_nstep):
  // interior faces the neighbour's trace, Neumann faces their own, Dirichlet
  // faces the zero slot after the halo (slot TILE + hcap); a slot holds u and
  // dn at the 3 Nfp face nodes
  {
    const int Nfp = N + 1, SW = trace_slot_words(N);
    for (int64_t i = 0; i < K; ++i) {
      const int own = (int)(i % TILE);
      for (int f = 0; f < 3; ++f) {
        int& c = fcode[(size_t)f * K + i];
        const int ty = fc_type(c);
        if (ty == FT_INTERIOR)
          c = fc_addr_pack(c, fc_slot(c), fc_face(c), fc_rev(c), SW, Nfp);
        else if (ty == FT_DIRICHLET)
          c = fc_addr_pack(c, TILE + hcap, 0, 0, SW, Nfp);
        else
          c = fc_addr_pack(c, own, f, 0, SW, Nfp);
      }
    }
  }
  // the halo lists hold element ids (the tile positions mapped through
  // perm), so the apply loads a halo element's values without a dependent
  // load of perm
  std::vector<int> hpos(ntiles * hcap);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int h = 0; h < hcap; ++h) {
      const int pos = hall[t * HALO_MAX + h];
      hpos[t * hcap + h] = pos >= 0 ? perm[pos] : -1;
    }
  // halo face coefficients a_f, b_f gathered per tile, hab[t][c][h], so the
  // apply reads them with contiguous loads
  std::vector<double> hab((size_t)ntiles * 6 * hcap, 0.0);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int h = 0; h < hcap; ++h) {
      const int64_t kid = hpos[t * hcap + h];
      if (kid < 0) continue;
      for (int c = 0; c < 6; ++c)
        hab[((size_t)t * 6 + c) * hcap + h] = rec[kid * DDGX_REC_WORDS + R_A + c];
    }
  auto* h = new ddgx_op_s();
  Op& o = h->op;
  o.N = N;
  o.K = K;
  o.dtype = dtype;
  o.Np = (N + 1) * (N + 2) / 2;
  o.Nfp = N + 1;
  o.max_halo = max_halo;
  o.halo_cap = hcap;
  o.perm_id = perm_id;
  int rc = DDGX_OK;
  auto put = [&](int** dst, const std::vector<int>& v, const char* what) -> bool {
    hipError_t e = hipMalloc((void**)dst, v.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      rc = hip_fail(e, what);
      return false;
    }
    return true;
  };
  do {
    if (hipGetDevice(&o.device) != hipSuccess) { rc = hip_fail(hipGetLastError(), "hipGetDevice"); break; }
    {
      // device table: column tables, then FRAG_COPIES replicas of the fragments
      const int nv = valu_table_size(N), nf = frag_size(N);
      std::vector<double> dev((size_t)nv + (size_t)nf * FRAG_COPIES);
      std::copy(tables, tables + nv, dev.begin());
      for (int c = 0; c < FRAG_COPIES; ++c)
        std::copy(tables + nv, tables + nv + nf, dev.begin() + nv + (size_t)c * nf);
      if ((rc = upload_as(dtype, dev.data(), dev.size(), &o.tab)) != DDGX_OK) break;
      o.frag = (char*)o.tab + (size_t)nv * (dtype == DDGX_F64 ? 8 : 4);
    }
    // records, face codes and Neumann Jacobians in tile order (SoA records)
    std::vector<double> recT((size_t)K * DDGX_REC_WORDS), fsjT((size_t)K * 4);
    std::vector<int> nbrT((size_t)K * 3), nbrP((size_t)K * 3);
    for (int64_t i = 0; i < K; ++i) {
      const int64_t k = perm[i];
      for (int w = 0; w < DDGX_REC_WORDS; ++w) recT[(size_t)w * K + i] = rec[k * DDGX_REC_WORDS + w];
      for (int f = 0; f < 3; ++f) {
        const int c = nbr[k * 4 + f];
        nbrT[(size_t)f * K + i] = c;
        nbrP[(size_t)f * K + i] = c >= 0 ? (inv[nbr_elem(c)] << 3) | (c & 7) : c;
      }
      for (int f = 0; f < 4; ++f) fsjT[(size_t)i * 4 + f] = fsj[k * 4 + f];
    }
    if ((rc = upload_as(dtype, recT.data(), recT.size(), &o.rec)) != DDGX_OK) break;
    if ((rc = upload_as(dtype, fsjT.data(), fsjT.size(), &o.fsj)) != DDGX_OK) break;
    if ((rc = upload_as(dtype, hab.data(), hab.size(), &o.hab)) != DDGX_OK) break;
    if (!put(&o.nbr, nbrT, "nbr") || !put(&o.nbrp, nbrP, "nbrp") || !put(&o.fcode, fcode, "fcode") ||
        !put(&o.perm, perm, "perm") ||
        !put(&o.hpos, hpos, "hpos"))
      break;
    o.grid_apply = (int)ntiles;
    const int64_t nvec = K * o.Np;
    int64_t gv = (nvec + 255) / 256;
    const int64_t cap = 4 * (int64_t)device_cu_count();
    o.grid_vec = (int)(gv < cap ? gv : cap);
  } while (false);
  if (rc != DDGX_OK) {
    free_op(o);
    delete h;
    return rc;
  }
  *out = h;
  return DDGX_OK;
}

int ddgx_op_create_curved(int N, int64_t K, int dtype, int use_cub, const double* reftab,
                          const double* geo, const double* fnorm, const double* fpar,
                          const int32_t* codes, const double* fmass, const double* estiff,
                          const double* emass, ddgx_op* out) {
  if (out == nullptr || reftab == nullptr || geo == nullptr || fnorm == nullptr ||
      fpar == nullptr || codes == nullptr) {
    set_error("ddgx_op_create_curved: null argument");
    return DDGX_ERR_INVALID;
  }
  *out = nullptr;
  if (N < 1 || N > 6) {
    set_error("ddgx_op_create_curved: polynomial order " + std::to_string(N) +
              " not built (1..6)");
    return DDGX_ERR_UNSUPPORTED;
  }
  if (dtype != DDGX_F64 && dtype != DDGX_F32) {
    set_error("ddgx_op_create_curved: dtype must be DDGX_F64 or DDGX_F32");
    return DDGX_ERR_INVALID;
  }
  if (K < 1 || K >= (int64_t(1) << 28)) {
    set_error("ddgx_op_create_curved: element count out of range [1, 2^28)");
    return DDGX_ERR_INVALID;
  }
  auto* h = new ddgx_op_s();
  Op& o = h->op;
  o.N = N;
  o.K = K;
  o.dtype = dtype;
  o.Np = (N + 1) * (N + 2) / 2;
  o.Nfp = N + 1;
  int rc = DDGX_OK;
  if (hipGetDevice(&o.device) != hipSuccess) rc = hip_fail(hipGetLastError(), "hipGetDevice");
  if (rc == DDGX_OK)
    rc = curved_create(o, use_cub, reftab, geo, fnorm, fpar, codes, fmass, estiff, emass);
  if (rc != DDGX_OK) {
    free_op(o);
    delete h;
    return rc;
  }
  const int64_t gv = (K * o.Np + 255) / 256;
  const int64_t cap = 4 * (int64_t)device_cu_count();
  o.grid_vec = (int)(gv < cap ? gv : cap);
  *out = h;
  return DDGX_OK;
}

int ddgx_op_destroy(ddgx_op op) {
  if (op == nullptr) return DDGX_OK;
  free_op(op->op);
  delete op;
  return DDGX_OK;
}

int ddgx_op_info(ddgx_op op, int* N, int64_t* K, int* dtype) {
  if (op == nullptr) {
    set_error("ddgx_op_info: null handle");
    return DDGX_ERR_INVALID;
  }
  if (N) *N = op->op.N;
  if (K) *K = op->op.K;
  if (dtype) *dtype = op->op.dtype;
  return DDGX_OK;
}

int ddgx_sipg_apply(ddgx_op op, const void* u, void* y, const void* u_dir, double alpha,
                    double shift, void* stream) {
  if (op == nullptr || u == nullptr || y == nullptr) {
    set_error("ddgx_sipg_apply: null argument");
    return DDGX_ERR_INVALID;
  }
  if (u == y) {
    set_error("ddgx_sipg_apply: u and y must not alias");
    return DDGX_ERR_INVALID;
  }
  return launch_apply(op->op, u, y, u_dir, alpha, shift, 0, nullptr, nullptr,
                      (hipStream_t)stream);
}

int ddgx_sipg_rhs(ddgx_op op, const void* f, const void* u_dir, const void* q_neu, void* rhs,
                  void* stream) {
  if (op == nullptr || rhs == nullptr) {
    set_error("ddgx_sipg_rhs: null argument");
    return DDGX_ERR_INVALID;
  }
  return launch_rhs(op->op, f, u_dir, q_neu, rhs, (hipStream_t)stream);
}

int ddgx_block_jacobi_create(ddgx_op op64, double alpha, int dtype, void* stream, ddgx_pc* out) {
  if (op64 == nullptr || out == nullptr) {
    set_error("ddgx_block_jacobi_create: null argument");
    return DDGX_ERR_INVALID;
  }
  if (op64->op.dtype != DDGX_F64) {
    set_error("ddgx_block_jacobi_create: blocks are built from a float64 operator");
    return DDGX_ERR_INVALID;
  }
  if (dtype != DDGX_F64 && dtype != DDGX_F32) {
    set_error("ddgx_block_jacobi_create: bad dtype");
    return DDGX_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  Op& o = op64->op;
  auto* h = new ddgx_pc_s();
  PC& pc = h->pc;
  pc.op = &o;
  pc.dtype = dtype;
  const size_t nb = (size_t)o.K * o.Np * o.Np;
  int rc = DDGX_OK;
  do {
    hipError_t e = hipMalloc((void**)&pc.blocks, nb * sizeof(double));
    if (e != hipSuccess) { rc = hip_fail(e, "hipMalloc(blocks)"); break; }
    pc.full = o.cv != nullptr;  // a curved operator's blocks are not symmetric
    e = hipMalloc(&pc.binv, (size_t)o.K * block_words(o.N, pc.full) * (dtype == DDGX_F64 ? 8 : 4));
    if (e != hipSuccess) { rc = hip_fail(e, "hipMalloc(binv)"); break; }
    rc = o.cv != nullptr ? launch_block_probe_general(o, pc.blocks, alpha, s)
                         : launch_block_probe(o, pc.blocks, alpha, s);
    if (rc != DDGX_OK) break;
    rc = launch_invert_blocks(o.N, o.K, dtype, pc.blocks, pc.binv, s, pc.full);
    if (rc != DDGX_OK) break;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) { rc = hip_fail(e, "block-Jacobi build"); break; }
  } while (false);
  if (rc != DDGX_OK) {
    if (pc.blocks) (void)hipFree(pc.blocks);
    if (pc.binv) (void)hipFree(pc.binv);
    delete h;
    return rc;
  }
  *out = h;
  return DDGX_OK;
}

int ddgx_block_jacobi_blocks(ddgx_pc pc, void* blocks, void* stream) {
  if (pc == nullptr || blocks == nullptr) {
    set_error("ddgx_block_jacobi_blocks: null argument");
    return DDGX_ERR_INVALID;
  }
  if (pc->pc.type != 0) {
    set_error("ddgx_block_jacobi_blocks: not a block-Jacobi preconditioner");
    return DDGX_ERR_INVALID;
  }
  const Op& o = *pc->pc.op;
  DDGX_HIP(hipMemcpyAsync(blocks, pc->pc.blocks, (size_t)o.K * o.Np * o.Np * sizeof(double),
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return DDGX_OK;
}

// Refusal of a call whose shared buffers another call holds (Claim).
static int busy(const char* who, const char* what) {
  set_error(std::string(who) + ": " + what +
            " is in use by a concurrent call (a solver context, ddgx_solver_create, gives "
            "each concurrent solve its own)");
  return DDGX_ERR_BUSY;
}

static const char* kOpState = "the operator's default solve state";
static const char* kMgState = "the multigrid preconditioner's default cycle buffers";

int ddgx_pc_apply(ddgx_pc pc, const void* r, void* z, void* stream) {
  if (pc == nullptr || r == nullptr || z == nullptr) {
    set_error("ddgx_pc_apply: null argument");
    return DDGX_ERR_INVALID;
  }
  if (r == z) {
    set_error("ddgx_pc_apply: r and z must not alias");
    return DDGX_ERR_INVALID;
  }
  const Op& o = *pc->pc.op;
  if (pc->pc.type == 1) {
    // held while the cycle is enqueued: refused during a solve that uses the
    // same buffers (the cycle itself runs after this call returns)
    Claim claim;
    if (!claim.take(&pc->pc.mg->ws_busy)) return busy("ddgx_pc_apply", kMgState);
    return mg_vcycle(*pc->pc.mg, *pc->pc.mg->ws, r, z, nullptr, (hipStream_t)stream);
  }
  return launch_bj_apply(o.N, o.K, pc->pc.dtype, pc->pc.binv, r, z, (hipStream_t)stream,
                         pc->pc.full);
}

int ddgx_pc_destroy(ddgx_pc pc) {
  if (pc == nullptr) return DDGX_OK;
  if (pc->pc.blocks) (void)hipFree(pc->pc.blocks);
  if (pc->pc.binv) (void)hipFree(pc->pc.binv);
  if (pc->pc.mg) mg_free(pc->pc.mg);
  delete pc;
  return DDGX_OK;
}

int ddgx_mg_create(int n_levels, const ddgx_op* ops, const ddgx_op* ops64,
                   const int32_t* transfer_kind, const double* const* transfer,
                   const ddgx_mg_cfg* cfg, void* stream, ddgx_pc* out) {
  if (ops == nullptr || ops64 == nullptr || cfg == nullptr || out == nullptr ||
      (n_levels > 1 && (transfer_kind == nullptr || transfer == nullptr))) {
    set_error("ddgx_mg_create: null argument");
    return DDGX_ERR_INVALID;
  }
  if (n_levels < 1 || n_levels > 16) {
    set_error("ddgx_mg_create: n_levels must be 1..16");
    return DDGX_ERR_INVALID;
  }
  std::vector<Op*> lv(n_levels), lv64(n_levels);
  for (int l = 0; l < n_levels; ++l) {
    if (ops[l] == nullptr || ops64[l] == nullptr) {
      set_error("ddgx_mg_create: null level operator");
      return DDGX_ERR_INVALID;
    }
    lv[l] = &ops[l]->op;
    lv64[l] = &ops64[l]->op;
  }
  MG* mg = nullptr;
  const int rc = mg_create(n_levels, lv.data(), lv64.data(), transfer_kind, transfer, *cfg,
                           (hipStream_t)stream, &mg);
  if (rc != DDGX_OK) return rc;
  auto* h = new ddgx_pc_s();
  h->pc.type = 1;
  h->pc.op = lv[0];
  h->pc.dtype = lv[0]->dtype;
  h->pc.mg = mg;
  h->pc.K = lv[0]->K;
  h->pc.N = lv[0]->N;
  *out = h;
  return DDGX_OK;
}

int ddgx_pcg_solve(ddgx_op op, double alpha, double shift, ddgx_pc pc, const void* b, void* x,
                   const ddgx_cg_cfg* cfg, ddgx_cg_info* info, void* stream) {
  if (op == nullptr || b == nullptr || x == nullptr || cfg == nullptr) {
    set_error("ddgx_pcg_solve: null argument");
    return DDGX_ERR_INVALID;
  }
  if (cfg->maxiter < 0) {
    set_error("ddgx_pcg_solve: maxiter must be >= 0");
    return DDGX_ERR_INVALID;
  }
  if (pc != nullptr && (pc->pc.op->K != op->op.K || pc->pc.op->N != op->op.N)) {
    set_error("ddgx_pcg_solve: preconditioner built for a different mesh/order");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (!claim.take(&op->op.ws_busy)) return busy("ddgx_pcg_solve", kOpState);
  if (pc != nullptr && pc->pc.type == 1 && !claim.take(&pc->pc.mg->ws_busy))
    return busy("ddgx_pcg_solve", kMgState);
  Work* w = nullptr;
  const int rc = default_work(op->op, false, &w);
  if (rc != DDGX_OK) return rc;
  return pcg_solve(op->op, *w, alpha, shift, pc ? &pc->pc : nullptr, nullptr, b, x, *cfg, info,
                   (hipStream_t)stream);
}

int ddgx_pcg_solve_noalloc(ddgx_op op, double alpha, double shift, ddgx_pc pc, const void* b,
                           void* x, const ddgx_cg_cfg* cfg, ddgx_cg_info* info, void* stream) {
  if (op == nullptr || b == nullptr || x == nullptr || cfg == nullptr) {
    set_error("ddgx_pcg_solve_noalloc: null argument");
    return DDGX_ERR_INVALID;
  }
  if (cfg->maxiter < 0) {
    set_error("ddgx_pcg_solve_noalloc: maxiter must be >= 0");
    return DDGX_ERR_INVALID;
  }
  if (pc != nullptr && (pc->pc.op->K != op->op.K || pc->pc.op->N != op->op.N)) {
    set_error("ddgx_pcg_solve_noalloc: preconditioner built for a different mesh/order");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (pc != nullptr && pc->pc.type == 1 && !claim.take(&pc->pc.mg->ws_busy))
    return busy("ddgx_pcg_solve_noalloc", kMgState);
  return pcg_solve_noalloc(op->op, alpha, shift, pc ? &pc->pc : nullptr, b, x, *cfg, info,
                           (hipStream_t)stream);
}

int ddgx_ir_solve(ddgx_op op64, ddgx_op op32, double alpha, double shift, ddgx_pc pc32,
                  const void* b, void* x, const ddgx_ir_cfg* cfg, ddgx_ir_info* info,
                  void* stream) {
  if (op64 == nullptr || op32 == nullptr || b == nullptr || x == nullptr || cfg == nullptr) {
    set_error("ddgx_ir_solve: null argument");
    return DDGX_ERR_INVALID;
  }
  const Op &a = op64->op, &c = op32->op;
  if (a.dtype != DDGX_F64 || c.dtype != DDGX_F32 || a.K != c.K || a.N != c.N) {
    set_error("ddgx_ir_solve: needs a float64 and a float32 operator of one mesh and order");
    return DDGX_ERR_INVALID;
  }
  if (pc32 != nullptr && (pc32->pc.dtype != DDGX_F32 || pc32->pc.op->K != c.K ||
                          pc32->pc.op->N != c.N)) {
    set_error("ddgx_ir_solve: the preconditioner must be float32 and of the same mesh/order");
    return DDGX_ERR_INVALID;
  }
  if (cfg->max_outer < 0 || cfg->inner_maxiter < 0 || !(cfg->inner_tol > 0.0)) {
    set_error("ddgx_ir_solve: bad refinement settings");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (!claim.take(&op64->op.ws_busy) || !claim.take(&op32->op.ws_busy))
    return busy("ddgx_ir_solve", kOpState);
  if (pc32 != nullptr && pc32->pc.type == 1 && !claim.take(&pc32->pc.mg->ws_busy))
    return busy("ddgx_ir_solve", kMgState);
  Work *w64 = nullptr, *w32 = nullptr;
  int rc = default_work(op64->op, false, &w64);
  if (rc == DDGX_OK) rc = default_work(op32->op, true, &w32);
  if (rc != DDGX_OK) return rc;
  return ir_solve(op64->op, *w64, op32->op, *w32, alpha, shift, pc32 ? &pc32->pc : nullptr,
                  nullptr, b, x, *cfg, info, (hipStream_t)stream);
}

int ddgx_sipg_param_grad(ddgx_op op, const double* u, const double* lam, const double* u_dir,
                         const double* kappa, const double* tau_scale, double* g_kappa,
                         double* g_tau, void* stream) {
  if (op == nullptr || u == nullptr || lam == nullptr || tau_scale == nullptr ||
      g_kappa == nullptr || g_tau == nullptr) {
    set_error("ddgx_sipg_param_grad: null argument");
    return DDGX_ERR_INVALID;
  }
  if (op->op.cv != nullptr) {
    set_error("ddgx_sipg_param_grad: needs an affine operator");
    return DDGX_ERR_UNSUPPORTED;
  }
  return launch_param_grad(op->op, u, lam, u_dir, kappa, tau_scale, g_kappa, g_tau,
                           (hipStream_t)stream);
}

int ddgx_dot(ddgx_op op, const void* a, const void* b, double* out, void* stream) {
  if (op == nullptr || a == nullptr || b == nullptr || out == nullptr) {
    set_error("ddgx_dot: null argument");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (!claim.take(&op->op.ws_busy)) return busy("ddgx_dot", kOpState);
  Work* w = nullptr;
  const int rc = default_work(op->op, false, &w);
  if (rc != DDGX_OK) return rc;
  return dot(op->op, *w, a, b, out, (hipStream_t)stream);
}

static void free_solver(ddgx_solver_s* h) {
  work_free(h->w);
  work_free(h->w32);
  mgwork_free(h->mgw);
  delete h;
}

int ddgx_solver_create(ddgx_op op, ddgx_op op32, ddgx_pc pc, ddgx_solver* out) {
  if (op == nullptr || out == nullptr) {
    set_error("ddgx_solver_create: null argument");
    return DDGX_ERR_INVALID;
  }
  *out = nullptr;
  const Op& a = op->op;
  if (op32 != nullptr) {
    const Op& c = op32->op;
    if (a.dtype != DDGX_F64 || c.dtype != DDGX_F32 || a.K != c.K || a.N != c.N) {
      set_error("ddgx_solver_create: refinement needs a float64 and a float32 operator of one "
                "mesh and order");
      return DDGX_ERR_INVALID;
    }
  }
  if (pc != nullptr) {
    const int want = op32 != nullptr ? DDGX_F32 : a.dtype;
    // a float32 multigrid may also precondition a float64 solve directly
    const bool mixed_mg = op32 == nullptr && a.dtype == DDGX_F64 && pc->pc.dtype == DDGX_F32 &&
                          pc->pc.type == 1;
    if ((pc->pc.dtype != want && !mixed_mg) || pc->pc.op->K != a.K || pc->pc.op->N != a.N) {
      set_error("ddgx_solver_create: preconditioner of another mesh, order or precision");
      return DDGX_ERR_INVALID;
    }
  }
  auto* h = new ddgx_solver_s();
  h->op = &op->op;
  h->op32 = op32 ? &op32->op : nullptr;
  h->pc = pc ? &pc->pc : nullptr;
  int rc = work_create(a, false, &h->w);
  if (rc == DDGX_OK && op32 != nullptr) rc = work_create(op32->op, true, &h->w32);
  // a float64 solve with a float32 multigrid: its conversion buffers now, so
  // the reported footprint is complete and no solve allocates
  if (rc == DDGX_OK && op32 == nullptr && pc != nullptr && pc->pc.dtype != a.dtype)
    rc = work_alloc_pc32(*h->w);
  // a solver context replays captured chunks of CG iterations
  if (rc == DDGX_OK) h->w->use_graph = true;
  if (rc == DDGX_OK && h->w32) h->w32->use_graph = true;
  if (rc == DDGX_OK && pc != nullptr && pc->pc.type == 1) rc = mgwork_create(*pc->pc.mg, &h->mgw);
  if (rc != DDGX_OK) {
    free_solver(h);
    return rc;
  }
  *out = h;
  return DDGX_OK;
}

int ddgx_solver_bytes(ddgx_solver sv, int64_t* bytes) {
  if (sv == nullptr || bytes == nullptr) {
    set_error("ddgx_solver_bytes: null argument");
    return DDGX_ERR_INVALID;
  }
  *bytes = (int64_t)(sv->w->bytes + (sv->w32 ? sv->w32->bytes : 0) +
                     (sv->mgw ? sv->mgw->bytes : 0));
  return DDGX_OK;
}

int ddgx_solver_pcg(ddgx_solver sv, double alpha, double shift, const void* b, void* x,
                    const ddgx_cg_cfg* cfg, ddgx_cg_info* info, void* stream) {
  if (sv == nullptr || b == nullptr || x == nullptr || cfg == nullptr) {
    set_error("ddgx_solver_pcg: null argument");
    return DDGX_ERR_INVALID;
  }
  if (sv->op32 != nullptr) {
    set_error("ddgx_solver_pcg: this solver was created for refinement (ddgx_solver_ir)");
    return DDGX_ERR_INVALID;
  }
  if (cfg->maxiter < 0) {
    set_error("ddgx_solver_pcg: maxiter must be >= 0");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (!claim.take(&sv->busy)) return busy("ddgx_solver_pcg", "this solver context");
  return pcg_solve(*sv->op, *sv->w, alpha, shift, sv->pc, sv->mgw, b, x, *cfg, info,
                   (hipStream_t)stream);
}

int ddgx_solver_ir(ddgx_solver sv, double alpha, double shift, const void* b, void* x,
                   const ddgx_ir_cfg* cfg, ddgx_ir_info* info, void* stream) {
  if (sv == nullptr || b == nullptr || x == nullptr || cfg == nullptr) {
    set_error("ddgx_solver_ir: null argument");
    return DDGX_ERR_INVALID;
  }
  if (sv->op32 == nullptr) {
    set_error("ddgx_solver_ir: this solver has no float32 operator (use ddgx_solver_pcg)");
    return DDGX_ERR_INVALID;
  }
  if (cfg->max_outer < 0 || cfg->inner_maxiter < 0 || !(cfg->inner_tol > 0.0)) {
    set_error("ddgx_solver_ir: bad refinement settings");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (!claim.take(&sv->busy)) return busy("ddgx_solver_ir", "this solver context");
  return ir_solve(*sv->op, *sv->w, *sv->op32, *sv->w32, alpha, shift, sv->pc, sv->mgw, b, x,
                  *cfg, info, (hipStream_t)stream);
}

int ddgx_solver_destroy(ddgx_solver sv) {
  if (sv == nullptr) return DDGX_OK;
  free_solver(sv);
  return DDGX_OK;
}

int ddgx_batch_create(ddgx_op op, ddgx_pc pc, int nseg, void* stream, ddgx_batch* out) {
  if (op == nullptr || out == nullptr) {
    set_error("ddgx_batch_create: null argument");
    return DDGX_ERR_INVALID;
  }
  *out = nullptr;
  Batch* b = nullptr;
  const int rc = batch_create(op->op, pc ? &pc->pc : nullptr, nseg, (hipStream_t)stream, &b);
  if (rc != DDGX_OK) return rc;
  auto* h = new ddgx_batch_s();
  h->b = b;
  *out = h;
  return DDGX_OK;
}

int ddgx_batch_bytes(ddgx_batch bt, int64_t* bytes) {
  if (bt == nullptr || bytes == nullptr) {
    set_error("ddgx_batch_bytes: null argument");
    return DDGX_ERR_INVALID;
  }
  *bytes = (int64_t)bt->b->bytes;
  return DDGX_OK;
}

int ddgx_batch_pcg(ddgx_batch bt, const double* alpha, const void* b, void* x,
                   const ddgx_cg_cfg* cfg, ddgx_cg_info* info, void* stream) {
  if (bt == nullptr || alpha == nullptr || b == nullptr || x == nullptr || cfg == nullptr) {
    set_error("ddgx_batch_pcg: null argument");
    return DDGX_ERR_INVALID;
  }
  if (cfg->maxiter < 0) {
    set_error("ddgx_batch_pcg: maxiter must be >= 0");
    return DDGX_ERR_INVALID;
  }
  Claim claim;
  if (!claim.take(&bt->busy)) return busy("ddgx_batch_pcg", "this batch context");
  DDGX_HIP(hipSetDevice(bt->b->op->device));
  return pcg_batch(*bt->b, alpha, b, x, *cfg, info, (hipStream_t)stream);
}

int ddgx_batch_destroy(ddgx_batch bt) {
  if (bt == nullptr) return DDGX_OK;
  batch_free(bt->b);
  delete bt;
  return DDGX_OK;
}

}  // extern "C"
