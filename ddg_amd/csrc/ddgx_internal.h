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
  // table layout (see RefElement.kernel_tables in ddg_amd/reference.py):
  // every matrix is stored by columns, each column zero-padded to RS (RSF
  // for face-length columns) entries, a multiple of 8, so columns start on
  // 64-byte boundaries for the scalar loads.
  //   Dr, Ds, ArT, AsT, M : Np columns of length Np
  //   L1[f]               : Nfp columns (index p) of length Nfp (index q)
  //   LR[f], LS[f]        : Nfp columns (index p) of length Np (index i)
  static constexpr int RS = (Np + 7) / 8 * 8;
  static constexpr int RSF = (Nfp + 7) / 8 * 8;
  static constexpr int oDr = 0;
  static constexpr int oDs = Np * RS;
  static constexpr int oAr = 2 * Np * RS;
  static constexpr int oAs = 3 * Np * RS;
  static constexpr int oM = 4 * Np * RS;
  static constexpr int oL1 = 5 * Np * RS;
  static constexpr int oLR = oL1 + 3 * Nfp * RSF;
  static constexpr int oLS = oLR + 3 * Nfp * RS;
  static constexpr int size = oLS + 3 * Nfp * RS;
};

// A symmetric Np x Np matrix packed as its upper triangle: entry (i, j),
// i <= j, at pk(i, j), rows one after another (the curved operator's volume
// stiffness and mass).  NPK words per matrix, also the size of a
// block-Jacobi inverse block, which is packed by diagonals (diag_pk).
template <int N>
struct SymPack {
  static constexpr int Np = Shape<N>::Np;
  static constexpr int NPK = Np * (Np + 1) / 2;
  __host__ __device__ static constexpr int pk(int i, int j) {
    return i <= j ? i * Np - i * (i - 1) / 2 + (j - i) : j * Np - j * (j - 1) / 2 + (i - j);
  }
};

// Block-Jacobi inverse blocks (symmetric) are stored by cyclic diagonals:
// diagonal q = (j - i) mod n holds the entries (i, i + q mod n), i = 0..n-1,
// at q * n + i, for q = 0 .. n/2; the diagonal n - q is the same entries read
// from the other end (entry (i, j) with (j - i) mod n > n/2 is entry (j, i)),
// and for even n the middle diagonal q = n/2 holds its n/2 distinct entries.
// n (n + 1) / 2 words per block, as for the packed upper triangle.  A row of
// the block product read diagonal by diagonal (row i takes diagonal d at
// column i + d mod n, block_row_diag) has the n lanes of an element read n
// consecutive words per step: one or two cache lines, where reading a
// packed upper triangle column by column spreads them over the whole block.
__host__ __device__ constexpr int diag_pk(int n, int i, int j) {
  int d = j - i;
  if (d < 0) d += n;
  int q = d, row = i;
  if (2 * d > n) {
    q = n - d;
    row = j;
  }
  if (2 * q == n && row >= n / 2) row -= n / 2;
  return q * n + row;
}

// Row li of B x for one element's block b (diag_pk layout) and its vector v,
// summed over the cyclic diagonals d = 0..n-1 (column li + d mod n)
template <typename T, int n, typename V>
__device__ __forceinline__ T block_row_diag(const T* b, const V& v, int li) {
  T acc = 0;
#pragma unroll
  for (int d = 0; d < n; ++d) {
    const int j = li + d < n ? li + d : li + d - n;
    acc += b[diag_pk(n, li, j)] * v(j);
  }
  return acc;
}

// Inverse blocks of a general (curved-element) operator are not symmetric
// (its nodal volume form weights M_ref by the per-node Jacobian on one side,
// sipg.py:188-189): they are stored in full, row-major, Np*Np words per
// element (FULL = true) instead of packed.
template <int N, bool FULL>
struct BlockWords {
  static constexpr int value = FULL ? Shape<N>::Np * Shape<N>::Np : SymPack<N>::NPK;
};

// Inverse blocks of the EPB elements of a tile into LDS (contiguous loads).
// Every load is issued (from a clamped index) before the first LDS store,
// so a tile pays one memory round trip for its blocks, not one per pass.
template <typename T, int N, bool FULL = false>
__device__ __forceinline__ void stage_blocks(const T* __restrict__ binv, T* s_b, int k0, int ne,
                                             int tid) {
  constexpr int NW = BlockWords<N, FULL>::value;
  constexpr int IT = (Shape<N>::EPB * NW + 255) / 256;  // loads per thread
  const T* src = binv + (size_t)k0 * NW;
  const int n = ne * NW;
  T v[IT];
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = tid + 256 * j;
    v[j] = src[i < n ? i : n - 1];
  }
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = tid + 256 * j;
    if (i < n) s_b[i] = v[j];
  }
}

// Row li of B_e r_e from the staged block and the staged vector.
template <typename T, int N, bool FULL = false>
__device__ __forceinline__ T block_row(const T* s_b, const T* s_r, int e, int li) {
  constexpr int Np = Shape<N>::Np, NW = BlockWords<N, FULL>::value;
  const T* b = s_b + e * NW;
  const T* v = s_r + e * Np;
  if constexpr (!FULL) return block_row_diag<T, Np>(b, [&](int j) { return v[j]; }, li);
  T acc = 0;
#pragma unroll
  for (int j = 0; j < Np; ++j) acc += b[li * Np + j] * v[j];
  return acc;
}

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

// Kernel mode bits for the apply (M_SMOOTH: one block-Jacobi sweep
// x_out = x + omega Binv (b - A x), launched through launch_smooth;
// M_PROLONG: with x = u + the h-prolongation of a coarser vector, formed on
// the fly, order 1 only; M_RESTRICT, with M_RESID: the residual b - A x is
// not stored but restricted to the next-coarser multigrid level, with that
// level's first smoothing sweep, launched through launch_resid_restrict;
// M_PROLONG_OUT, with M_SMOOTH: the sweep's result is also prolongated into
// the next-finer level's iterate, p-transfers, launched through
// launch_smooth's prolong_to; M_PUPDATE, with M_DOT: CG's direction update
// p = z + beta p_old runs inside the apply, launched through
// launch_apply_pupdate)
enum : int {
  M_HELM = 1, M_UDIR = 2, M_DOT = 8, M_RESID = 16, M_SMOOTH = 32, M_PROLONG = 64,
  M_RESTRICT = 128, M_PROLONG_OUT = 256, M_PUPDATE = 512
};

// Device scalars of one CG solve
struct CGScalars {
  double gamma, pAp, rr, bb, atol2, alpha, beta, rs;
  int k, done, maxiter, use_pc;
  double tol, atol;
  // gamma of iteration i at gq[i & 1]: the kernel that forms the next gamma
  // writes the other slot while its blocks still read this one
  double gq[2];
  // iterations begun, by the parity of the iteration about to begin (the
  // apply that forms the direction reads kq[i & 1] and its block 0 writes
  // kq[(i + 1) & 1] and k)
  int kq[2];
};

// Sum over a 256-thread block (every thread gets the sum; s_red: 4 doubles
// of LDS); the fixed order makes the bits independent of timing.
__device__ inline double block_sum_256(double v, double* s_red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
  __syncthreads();
  return t;
}

// Both sums of partial[0..nparts) and partial[nparts..2 nparts) by a
// 256-thread block, in one fixed order: each thread adds entries i, i + 256,
// ... in that order (eight loaded before they are added, so the loads
// overlap), then block_sum_256.
__device__ inline void block_sum2_partials(const double* __restrict__ partial, int nparts,
                                           double* s_red, double& r0, double& r1) {
  double s0 = 0.0, s1 = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * 256 < nparts; i += 8 * 256) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = partial[i + u * 256];
      b[u] = partial[nparts + i + u * 256];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      s0 += a[u];
      s1 += b[u];
    }
  }
  for (; i < nparts; i += 256) {
    s0 += partial[i];
    s1 += partial[nparts + i];
  }
  r0 = block_sum_256(s0, s_red);
  r1 = block_sum_256(s1, s_red);
}

// Per-element data of an operator is stored in tile order: tile position
// k (0..K-1, 64 per tile) holds element perm[k].  Vectors (u, y, ...) keep
// the caller's element order and are addressed through perm.
constexpr int TILE = 64;      // elements per matrix-core workgroup
// Largest nodal vector (in float64 bytes) an operator takes: the apply
// addresses vectors with 32-bit byte offsets from a uniform base
// (ld_off / st_off), so K * Np * 8 stays below 2^32.
constexpr int64_t MAX_VECTOR_BYTES = (int64_t(1) << 32) - 1;

// Element of a device array at a 32-bit byte offset from a base held in
// scalar registers: one global load or store with a scalar base and a
// 32-bit vector offset, instead of a 64-bit address per lane.
template <typename T>
__device__ __forceinline__ T ld_off(const T* base, unsigned boff) {
  return *(const T*)((const char*)base + boff);
}
template <typename T>
__device__ __forceinline__ void st_off(T* base, unsigned boff, T v) {
  *(T*)((char*)base + boff) = v;
}
constexpr int LDS_BYTES_PER_CU = 160 * 1024;  // gfx950
constexpr int HALO_MAX = 64;  // out-of-tile neighbour elements per tile (largest capacity)
constexpr int HALO_SMALL = 32;  // capacity of the smaller-LDS variant of the apply
// Packed per-face code of the apply: bits 0-7 LDS slot of the neighbour
// trace (0..63 in the tile, 64+h halo), bits 8-9 neighbour face, bit 10
// reversed node order, bits 11-12 face type.
//
// Bits 13-25 and 26-28 give the same neighbour read as an address: the LDS
// word of the trace the face reads as its neighbour's at face node p is
// fc_nbase + p * fc_nstep (2 words per face node: u, dn).  An interior face
// reads the neighbour's trace; a Neumann face reads its own, so the jump
// u^- - u^+ and the flux sum dn^- - dn_nbr vanish exactly; a Dirichlet face
// reads the zero slot after the halo (u^+ = 0, no neighbour flux).  The
// apply's face rows then need no test of the face type.
enum : int { FT_INTERIOR = 0, FT_DIRICHLET = 1, FT_NONE = 2 };
__host__ __device__ inline int fc_slot(int c) { return c & 0xff; }
__host__ __device__ inline int fc_face(int c) { return (c >> 8) & 3; }
__host__ __device__ inline int fc_rev(int c) { return (c >> 10) & 1; }
__host__ __device__ inline int fc_type(int c) { return (c >> 11) & 3; }
__host__ __device__ inline int fc_pack(int slot, int face, int rev, int type) {
  return slot | (face << 8) | (rev << 10) | (type << 11);
}
// LDS words of an element's trace slot in the apply: u and dn at the 3 (N+1)
// face nodes, plus one word of padding that makes the stride odd
__host__ __device__ constexpr int trace_slot_words(int N) { return 6 * (N + 1) + 1; }
__host__ __device__ inline int fc_nbase(int c) { return (c >> 13) & 0x1fff; }
__host__ __device__ inline int fc_nstep(int c) { return (int)((unsigned)c << 3) >> 29; }
// c with the address fields of a read of slot `slot`, face `face` (node
// order reversed when rev), for SW LDS words per element slot and Nfp nodes
// per face
__host__ __device__ inline int fc_addr_pack(int c, int slot, int face, int rev, int SW, int Nfp) {
  const int base = slot * SW + 2 * face * Nfp + (rev ? 2 * (Nfp - 1) : 0);
  const int step = rev ? -2 : 2;
  return (c & 0x1fff) | (base << 13) | ((step & 7) << 26);
}

// General operator for curved (non-affine) elements: per-node metrics,
// per-side face data, optional volume and face cubature (csrc/curved.hip).
// Its apply is three launches -- face traces, face sides, element gather --
// through the trace and cotangent buffers below, so one curved operator
// runs one apply at a time.
struct Curved {
  int use_cub = 0;         // volume and face integrals by cubature (sipg.py:190-201, 281-299)
  void* geo = nullptr;     // K x 5 x Np: rx, sx, ry, sy, J at the nodes (dtype)
  void* fnorm = nullptr;   // K x 3 x Nfp x 2: nx, ny at the face nodes
  void* fpar = nullptr;    // K x 3 x 3: w, w tau, sJ at face node 0 of each element side
  void* fmass = nullptr;   // K x 3 x Nfp x Nfp face cubature mass of each side (cubature)
  void* estiff = nullptr;  // K x NPK volume cubature stiffness of each element, packed upper
  void* emass = nullptr;   // K x NPK cubature mass of each element, packed upper (cubature)
  void* reftab = nullptr;  // Dr, Ds, M (Np x Np, row-major), then m1d (3 x Nfp x Nfp)
  int* codes = nullptr;    // K x 3 face codes (nbr_elem / nbr_face / nbr_rev, CODE_*)
  void* trace = nullptr;   // 3 planes of K*3*Nfp: u, du/dx, du/dy at the face nodes
  void* cown = nullptr;    // 3 planes: face cotangents of the own element (v, dv/dx, dv/dy)
  void* cnbr = nullptr;    // 3 planes: face cotangents for the neighbour across each side
  int epb = 1;             // elements per workgroup of the element kernel
  int grid_face = 0, grid_side = 0, grid_elem = 0, grid_node = 0;
};

struct Op {
  Curved* cv = nullptr;  // set: the general (curved-element) operator; the fields
                         // of the affine apply below are then unused
  int N = 0, dtype = 0, device = 0;
  int64_t K = 0;
  int Np = 0, Nfp = 0;
  void* tab = nullptr;   // table (dtype): column tables, then MFMA fragments
  void* frag = nullptr;  // = tab + Shape<N>::size (MFMA A fragments)
  void* rec = nullptr;   // 16 x K records, tile order (dtype)
  int* nbr = nullptr;    // 3 x K face codes (see above), tile order
  int* nbrp = nullptr;   // the same codes with the neighbour's tile position in place of its id
  int* fcode = nullptr;  // 3 x K packed apply codes (fc_* below), tile order
  int* perm = nullptr;   // K: element id at each tile position
  bool perm_id = true;   // perm is the identity (the apply then skips it)
  int* hpos = nullptr;   // ntiles x halo_cap tile positions of halo elements (-1: none)
  void* hab = nullptr;   // ntiles x 6 x halo_cap halo face coefficients a_f, b_f (dtype)
  void* fsj = nullptr;   // K x 4 Neumann face Jacobians, tile order (dtype)
  int grid_apply = 0;    // workgroups of one apply: one per 64-element tile (= dot partials)
  int max_halo = 0;      // largest halo of a tile
  int halo_cap = 0;      // HALO_SMALL or HALO_MAX: halo entries stored per tile
  int grid_vec = 0;      // grid for vector kernels / reductions
  struct Work* ws = nullptr;  // default Krylov state (ddgx_pcg_solve / ddgx_ir_solve / ddgx_dot)
  int ws_busy = 0;            // ws claimed by a running call (Claim below)
  // A batch replica (op_replicate): nseg disjoint copies of a source operator
  // of K / nseg elements each, sharing the source's table (not owned).
  int nseg = 1;
  bool owns_tab = true;
};

// Krylov state of one solve on an operator's vector shape and dtype.  The
// operator itself is immutable after creation; everything a solve writes
// lives here, so solves on one operator with different Work run at once.
struct Work {
  int64_t K = 0;
  int Np = 0, dtype = 0, device = 0;
  void* r = nullptr;
  void* p = nullptr;
  void* z = nullptr;
  void* Ap = nullptr;
  double* partials = nullptr;  // 3 * npart doubles: vector reductions, then <p, Ap>
  int npart = 0;
  void* ir_b = nullptr;        // refinement (float32 work): scaled residual / correction
  void* ir_x = nullptr;
  void* pc_r = nullptr;        // float64 solve with a float32 multigrid: residual and
  void* pc_z = nullptr;        // preconditioned residual in float32 (allocated on first use)
  // second direction vector: with the direction update inside the apply,
  // iteration i reads p_old from one and writes p to the other (by parity);
  // allocated on first use unless the state lives in fixed memory
  void* p2 = nullptr;
  bool fixed = false;
  CGScalars* scal = nullptr;
  CGScalars* scal_host = nullptr;  // pinned
  size_t bytes = 0;            // device bytes held
  // Captured chunk of CG iterations (HIP graph) and what it was captured for
  bool use_graph = false;
  hipGraphExec_t chunk = nullptr;
  const void* g_x = nullptr;
  const void* g_pc = nullptr;
  const void* g_mgw = nullptr;
  double g_alpha = 0.0, g_shift = 0.0;
  int g_every = 0;
  hipStream_t g_stream = nullptr;
};
int work_create(const Op& op, bool refine, Work** out);
// float32 buffers of a float64 solve preconditioned by a float32 multigrid
int work_alloc_pc32(Work& w);
void work_free(Work* w);

struct MG;

struct PC {
  int type = 0;              // 0 block-Jacobi, 1 multigrid V-cycle
  Op* op = nullptr;          // BJ: the float64 operator the blocks came from
  int dtype = 0;             // precision of binv and of the vectors
  void* binv = nullptr;      // K x Np(Np+1)/2 packed symmetric inverse blocks (dtype),
                             // or K x Np x Np full ones (full: a curved operator's)
  bool full = false;
  double* blocks = nullptr;  // K x Np x Np float64 blocks (kept for inspection)
  MG* mg = nullptr;          // type 1
  int64_t K = 0;             // vector shape the PC applies to
  int N = 0;
};

// One level of a multigrid hierarchy (fine -> coarse); the last level is
// solved directly with a dense float64 inverse.
struct MGLevel {
  Op* op = nullptr;        // operator at this level (MG dtype)
  void* binv = nullptr;    // block-Jacobi inverse blocks of this level (MG dtype): packed,
                           // or full for a curved level (full = true)
  bool full = false;
  double alpha = 0.0;      // Helmholtz coefficient of this level's operator
  int tkind = 0;           // transfer to the next-coarser level: 0 = p, 1 = h
  int np_lo = 0;           // Np of the next-coarser level (p transfer)
  void* P = nullptr;       // p: Np x np_lo matrix; h: 4 x Np x Np child matrices (MG dtype)
};

// Buffers one V-cycle writes (per-level right-hand side, iterate, residual
// and the coarse-solve partial sums): one MGWork per concurrent solve.
struct MGWork {
  int nlev = 0;
  void* b[16] = {};
  void* x[16] = {};
  void* r[16] = {};
  double* gemv_part = nullptr;
  size_t bytes = 0;
};

struct MG {
  int dtype = 0, nu = 2;
  double omega = 0.7;
  double alpha = 0.0;
  MGLevel* lv = nullptr;
  int nlev = 0;
  double* Ainv = nullptr;       // coarsest level dense inverse (float64), n x n
  int nc = 0;                   // unknowns of the coarsest level (of one segment)
  MGWork* ws = nullptr;         // default cycle buffers (ddgx_pc_apply, ddgx_pcg_solve)
  int ws_busy = 0;              // ws claimed by a running call (Claim below)
  // A batch replica (mg_replicate): every level is nseg disjoint copies and
  // the coarse solve applies Ainv to each segment; the level operators are
  // then owned by the hierarchy.
  int nseg = 1;
  bool owns_ops = false;
};
int mgwork_create(const MG& mg, MGWork** out);
void mgwork_free(MGWork* w);

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Claims shared buffers (a handle's default Krylov state or cycle buffers, a
// GPU's fixed solve workspace) for the duration of one call: take() sets the
// flag unless another call holds it, and the destructor releases every flag
// taken.  Calls that return only after their device work ends (the solves,
// ddgx_dot) hold the claim over all of that work, so two such calls on one
// set of buffers cannot overlap: the second is refused instead of racing.
class Claim {
 public:
  Claim() = default;
  Claim(const Claim&) = delete;
  Claim& operator=(const Claim&) = delete;
  ~Claim() {
    for (int i = 0; i < n_; ++i) __atomic_store_n(held_[i], 0, __ATOMIC_RELEASE);
  }
  bool take(int* flag) {
    if (n_ == kMax || __atomic_exchange_n(flag, 1, __ATOMIC_ACQUIRE) != 0) return false;
    held_[n_++] = flag;
    return true;
  }

 private:
  static constexpr int kMax = 4;
  int* held_[kMax] = {};
  int n_ = 0;
};

#define DDGX_HIP(call)                                   \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return ::ddgx::hip_fail(_e, #call); \
  } while (0)

// launchers (defined per translation unit)
int launch_apply(const Op& op, const void* u, void* y, const void* udir, double alpha,
                 double shift, int mode, double* partial, const int* skip, hipStream_t s,
                 const void* bvec = nullptr);
// A p with CG's direction update in the same launch (affine float64
// operators of order 2..4, apply_pupdate_ok): every workgroup forms gamma = <r, z> (and <r, r>) from the update's
// partials rz (2 x nrz), beta = gamma / gq[q ^ 1] and the stop test, then
// applies p = z + beta p_old (tile and halo), writes p for its tile and the
// partial <p, A p>; block 0 records gamma in gq[q], the residual and the
// iteration count (kq), or the stop flag.
bool apply_pupdate_ok(const Op& op);
int launch_apply_pupdate(const Op& op, const void* z, const void* p_old, void* p, void* Ap,
                         double alpha, double shift, double* partial, CGScalars* sc,
                         const double* rz, int nrz, int q, hipStream_t s);
// One fused block-Jacobi sweep x_out = x_in + omega Binv (b - (A + alpha M) x_in)
// (x_out must not alias x_in; Binv packed, element order, the operator's dtype).
bool smooth_fused(const Op& op);
// With prolong_to != null (p-transfers, affine levels): the sweep's result is
// also added, through prolong_to_mat (np_to x Np, row-major), into the
// next-finer level's iterate prolong_to -- the prolongation x_f += P x of the
// V-cycle, element by element, in the coarser level's last sweep.
int launch_smooth(const Op& op, const void* binv, double omega, double alpha, const void* x_in,
                  void* x_out, const void* b, const int* skip, hipStream_t s,
                  const void* prolong_from = nullptr, const void* prolong_mat = nullptr,
                  void* prolong_to = nullptr, const void* prolong_to_mat = nullptr,
                  int np_to = 0);
// Whether launch_smooth can take the h-prolongation of a coarser level as part
// of its input (prolong_from / prolong_mat) on this operator
bool smooth_prolongs(const Op& op);
// The multigrid residual and its restriction in one launch:
//   b_c = P^T (b - (A + alpha M) x)   per coarse unit (tkind 0: the element,
//         P is Np x np_lo; tkind 1: the four children 4k..4k+3 of coarse
//         element k, P is 4 Np x Np), and with x_c != null also the coarser
//         level's first sweep x_c = omega Binv_c b_c (Binv_c packed, np_lo).
// The fine residual is not stored.  Returns DDGX_ERR_UNSUPPORTED where the
// fused kernel is not built (the caller then runs the residual apply and the
// restriction kernel): an h-transfer needs order 1 in the identity tile order.
int launch_resid_restrict(const Op& op, int tkind, int np_lo, const void* P, const void* x,
                          const void* b, double alpha, void* b_c, const void* binv_c,
                          double omega, void* x_c, const int* skip, hipStream_t s);
int mg_vcycle(const MG& mg, MGWork& w, const void* b, void* x, const int* skip, hipStream_t s);
int mg_create(int nlev, Op* const* ops, Op* const* ops64, const int* tkind,
              const double* const* P, const ddgx_mg_cfg& cfg, hipStream_t s, MG** out);
void mg_free(MG* mg);
int launch_rhs(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
               hipStream_t s);
// curved operator (csrc/curved.hip)
int curved_create(Op& op, int use_cub, const double* reftab, const double* geo,
                  const double* fnorm, const double* fpar, const int32_t* codes,
                  const double* fmass, const double* estiff, const double* emass);
void curved_free(Curved* cv);
int launch_curved(const Op& op, const void* u, void* y, const void* udir, double alpha,
                  double shift, int mode, double* partial, const int* skip, hipStream_t s,
                  const void* bvec, bool face_nodal);
int launch_curved_rhs(const Op& op, const void* f, const void* udir, const void* q, void* rhs,
                      hipStream_t s);
int launch_block_probe(const Op& op, double* blocks, double alpha, hipStream_t s);
int launch_block_probe_general(const Op& op, double* blocks, double alpha, hipStream_t s);
int launch_invert_blocks(int N, int64_t K, int dtype, const double* blocks, void* binv,
                         hipStream_t s, bool full = false);
int launch_bj_apply(int N, int64_t K, int dtype, const void* binv, const void* r, void* z,
                    hipStream_t s, bool full = false);
size_t packed_block_words(int N);
size_t block_words(int N, bool full);  // per element: packed or full

// Batched solves: nseg requests of one problem class in every launch.
int op_replicate(const Op& src, int nseg, hipStream_t s, Op** out);
void op_release(Op* op);  // frees what the operator owns (any Op made here)
int mg_replicate(const MG& src, int nseg, hipStream_t s, MG** out);
struct Batch {
  int nseg = 0;
  Op* op = nullptr;       // float64 operator on nseg copies (owned)
  PC* pc = nullptr;       // multigrid on nseg copies (owned), or null
  MGWork* mgw = nullptr;  // its cycle buffers
  int64_t n1 = 0;         // unknowns of one segment
  int g1 = 0, t1 = 0;     // vector grid and tiles of one segment (the source operator's)
  void* j1 = nullptr;     // the source operator's Jacobian record plane (tile order)
  double* alpha = nullptr;  // nseg Helmholtz coefficients on the device
  double* alpha_host = nullptr;
  void *r = nullptr, *p = nullptr, *z = nullptr, *Ap = nullptr;
  void *pc_r = nullptr, *pc_z = nullptr;  // float32 copies around a float32 V-cycle
  double* partials = nullptr;  // nseg x 2 g1 vector reductions
  double* apart = nullptr;     // nseg x t1 apply <p, Ap> partials (one per tile)
  CGScalars* scal = nullptr;
  CGScalars* scal_host = nullptr;  // pinned
  size_t bytes = 0;
  // captured chunk of batched iterations and what it was captured for
  hipGraphExec_t chunk = nullptr;
  const void* g_x = nullptr;
  double g_a = 0.0;
  int g_every = 0;
  hipStream_t g_stream = nullptr;
};
int batch_create(const Op& op, const PC* pc, int nseg, hipStream_t s, Batch** out);
void batch_free(Batch* b);
int pcg_batch(Batch& bt, const double* alpha, const void* b, void* x, const ddgx_cg_cfg& cfg,
              ddgx_cg_info* info, hipStream_t s);
// rec plane R_J of segment s (tile order) = alpha[s] * j1, s < nseg
int launch_scale_jacobian(Batch& bt, hipStream_t s);

int pcg_solve_noalloc(const Op& op, double alpha, double shift, const PC* pc, const void* b,
                      void* x, const ddgx_cg_cfg& cfg, ddgx_cg_info* info, hipStream_t s);
int pcg_solve(const Op& op, Work& w, double alpha, double shift, const PC* pc, MGWork* mgw,
              const void* b, void* x, const ddgx_cg_cfg& cfg, ddgx_cg_info* info,
              hipStream_t s);
int dot(const Op& op, Work& w, const void* a, const void* b, double* out, hipStream_t s);
int launch_param_grad(const Op& op, const double* u, const double* lam, const double* udir,
                      const double* kappa, const double* tau, double* gk, double* gt,
                      hipStream_t s);
int ir_solve(const Op& op64, Work& w64, const Op& op32, Work& w32, double alpha, double shift,
             const PC* pc32, MGWork* mgw, const void* b, void* x, const ddgx_ir_cfg& cfg,
             ddgx_ir_info* info, hipStream_t s);
// default Work of an operator, created on first use
int default_work(Op& op, bool refine, Work** out);

int device_cu_count();

}  // namespace ddgx
