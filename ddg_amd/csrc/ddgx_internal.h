// Internal declarations shared by the ddgx HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/ddgx.h"

namespace ddgx {

// ---------------------------------------------------------------------------
// Reference-element shapes and table layout (order N is a template constant)
// ---------------------------------------------------------------------------
template <int N>
struct Shape {
  static constexpr int Np = (N + 1) * (N + 2) / 2;
  static constexpr int Nfp = N + 1;
  // lanes per element in the thread-per-node kernels (power of two >= Np)
  static constexpr int NL = Np <= 4 ? 4 : Np <= 8 ? 8 : Np <= 16 ? 16 : 32;
  static constexpr int EPB = 256 / NL;  // elements per 256-thread tile
  // table offsets (see RefElement.kernel_tables in ddg_amd/reference.py)
  static constexpr int oDr = 0;
  static constexpr int oDs = Np * Np;
  static constexpr int oAr = 2 * Np * Np;
  static constexpr int oAs = 3 * Np * Np;
  static constexpr int oM = 4 * Np * Np;
  static constexpr int oL1 = 5 * Np * Np;
  static constexpr int oLR = oL1 + 3 * Nfp * Nfp;
  static constexpr int oLS = oLR + 3 * Np * Nfp;
  static constexpr int size = oLS + 3 * Np * Nfp;
};

// Node index of face node p on face f.  Nodes are ordered in rows of
// constant s (row a has N+1-a nodes, starting at a(N+1) - a(a-1)/2):
// face 0 = row 0, face 1 = last node of each row, face 2 = first node of
// each row -- the ascending-index Fmask of the reference element.
template <int N>
__host__ __device__ constexpr int fmask(int f, int p) {
  return f == 0 ? p : (p * (N + 1) - p * (p - 1) / 2 + (f == 1 ? (N - p) : 0));
}

// Position of node i on face f, or -1.
template <int N>
__host__ __device__ constexpr int facepos(int f, int i) {
  for (int p = 0; p <= N; ++p)
    if (fmask<N>(f, p) == i) return p;
  return -1;
}

// Per-element record (DDGX_REC_WORDS words)
enum : int {
  R_VRR = 0, R_VRS = 1, R_VSS = 2, R_J = 3,
  R_A = 4,     // a_f = kappa (nx rx + ny ry), f = 0..2
  R_B = 7,     // b_f = kappa (nx sx + ny sy)
  R_SJW = 10,  // adjoint/consistency face weight
  R_PEN = 13,  // penalty face weight
};

// Face code helpers
__host__ __device__ inline int nbr_elem(int c) { return c >> 3; }
__host__ __device__ inline int nbr_face(int c) { return (c >> 1) & 3; }
__host__ __device__ inline int nbr_rev(int c) { return c & 1; }
constexpr int CODE_DIRICHLET = -1;
constexpr int CODE_NONE = -2;

// Kernel mode bits for the apply
enum : int { M_HELM = 1, M_UDIR = 2, M_PROBE = 4, M_DOT = 8, M_RESID = 16 };

// Device scalars of one CG solve
struct CGScalars {
  double gamma, pAp, rr, bb, atol2, alpha, beta, rs;
  int k, done, maxiter, use_pc;
  double tol, atol;
};

struct Op {
  int N = 0, dtype = 0, device = 0;
  int64_t K = 0;
  int Np = 0, Nfp = 0;
  void* tab = nullptr;   // table (dtype)
  void* rec = nullptr;   // K x 16 (dtype)
  int* nbr = nullptr;    // K x 4
  void* fsj = nullptr;   // K x 4 (dtype)
  int grid_apply = 0;    // persistent grid for the apply
  int grid_vec = 0;      // grid for vector kernels / reductions
  // CG workspace
  void* r = nullptr;
  void* p = nullptr;
  void* z = nullptr;
  void* Ap = nullptr;
  double* partials = nullptr;  // 2 * max(grid) doubles
  CGScalars* scal = nullptr;
  CGScalars* scal_host = nullptr;  // pinned
};

struct MG;

struct PC {
  int type = 0;              // 0 block-Jacobi, 1 multigrid V-cycle
  Op* op = nullptr;          // BJ: the float64 operator the blocks came from
  int dtype = 0;             // precision of binvT and of the vectors
  void* binvT = nullptr;     // K x Np x Np, transposed inverse blocks (dtype)
  double* blocks = nullptr;  // K x Np x Np float64 blocks (kept for inspection)
  MG* mg = nullptr;          // type 1
  int64_t K = 0;             // vector shape the PC applies to
  int N = 0;
};

// One level of a multigrid hierarchy (fine -> coarse); the last level is
// solved directly with a dense float64 inverse.
struct MGLevel {
  Op* op = nullptr;        // operator at this level (MG dtype)
  void* binvT = nullptr;   // block-Jacobi inverse blocks of this level (MG dtype)
  double alpha = 0.0;      // Helmholtz coefficient of this level's operator
  int tkind = 0;           // transfer to the next-coarser level: 0 = p, 1 = h
  int np_lo = 0;           // Np of the next-coarser level (p transfer)
  void* P = nullptr;       // p: Np x np_lo matrix; h: 4 x Np x Np child matrices (MG dtype)
  void* b = nullptr;       // level right-hand side / solution / residual (MG dtype)
  void* x = nullptr;
  void* r = nullptr;
};

struct MG {
  int dtype = 0, nu = 2;
  double omega = 0.7;
  MGLevel* lv = nullptr;
  int nlev = 0;
  double* Ainv = nullptr;  // coarsest level dense inverse (float64, symmetric)
  int nc = 0;
};

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define DDGX_HIP(call)                                   \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return ::ddgx::hip_fail(_e, #call); \
  } while (0)

// launchers (defined per translation unit)
int launch_apply(const Op& op, const void* u, void* y, const void* udir, double alpha,
                 double shift, int mode, int probe_j, double* partial, const int* skip,
                 hipStream_t s, const void* bvec = nullptr);
int pc_apply(PC& pc, const void* r, void* z, hipStream_t s);
int mg_vcycle(MG& mg, const void* b, void* x, hipStream_t s);
int launch_rhs(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
               hipStream_t s);
int launch_invert_blocks(int N, int64_t K, int dtype, const double* blocks, void* binvT,
                         hipStream_t s);
int launch_bj_apply(int N, int64_t K, int dtype, const void* binvT, const void* r, void* z,
                    hipStream_t s);

int pcg_solve(Op& op, double alpha, double shift, PC* pc, const void* b, void* x,
              const ddgx_cg_cfg& cfg, ddgx_cg_info* info, hipStream_t s);
int dot(Op& op, const void* a, const void* b, double* out, hipStream_t s);

int device_cu_count();

}  // namespace ddgx
