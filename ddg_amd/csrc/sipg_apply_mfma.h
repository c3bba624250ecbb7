// SIPG apply on the matrix cores (v_mfma_*_16x16x4) for gfx950:
//   y = (A_sipg + alpha*J*M + shift*I) u          (or y = b - that, residual mode)
// with A_sipg the mass-weighted SIPG Laplacian of src/ddg/dg2d/sipg.py:142-337
// in the closed form of its reverse-mode gradient:
//
//   volume : y_k = ArT (vrr ur + vrs us) + AsT (vrs ur + vss us),
//            ur = Dr u_k, us = Ds u_k, ArT = Dr^T M^T, AsT = Ds^T M^T
//   face f : P = pen_f [u] - sJw_f (dn^- + dn^+)    -> y[Fmask_f] += m1d_f^T P
//            X = -sJw_f [u]                          -> y += (a_f LR_f + b_f LS_f) X
//
// [u] = u^- - u^+, dn^- = a_f ur + b_f us at the own face nodes, dn^+ the
// neighbour's normal derivative taken with the own normal.
//
// The element-local contractions are 16x16x4 MFMA products with 16 elements
// of a wave as the N dimension:
//
//   [ur; us] = [Dr; Ds] . U                      (GEMM 1)
//   y = ArT . Wr + AsT . Ws + G . F  (+ M . aJu)  (GEMM 2)
//
// U, Wr, Ws and y are held in the accumulator layout of the MFMA: lane
// (g = lane/16, e = lane%16), register r of block b holds output row
// 16b + crow(g, r) of element e, with crow(g, r) = g + 4r for float64 and
// 4g + r for float32 (the output layouts of the two instructions).  Which
// node a row stands for is free (RowMap: face-major for N <= 4), and so is
// the k order of a product, so the constant A fragments are laid out on the
// host (reference.mfma_fragments) with the matching row and column
// permutations, and an output feeds the next product as its B operand
// without any lane shuffle.  F holds
// the per-face-node values P, a*X, b*X (9*Nfp rows); G = [E_f m1d_f^T | LR_f |
// LS_f] lifts them.
//
// A workgroup is 4 waves and handles one tile of 64 elements; the grid has
// one workgroup per tile, mapped XCD-contiguously (xcd_tile).  Face traces
// (u, dn) of the tile's elements and of its halo -- the out-of-tile
// neighbours, listed per tile on the host -- are computed with the same
// GEMM 1 and exchanged through LDS, so every face reads its neighbour trace
// from LDS through a host-computed slot index.  All global loads of a tile
// are issued up front; every element owns its output row (no atomics).

#pragma once

#include "ddgx_internal.h"

namespace ddgx {

template <typename T>
struct MfmaT;
template <>
struct MfmaT<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __host__ __device__ static constexpr int crow(int g, int r) { return g + 4 * r; }
};
template <>
struct MfmaT<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __host__ __device__ static constexpr int crow(int g, int r) { return 4 * g + r; }
};

// Replicas of the fragment table in device memory (workgroup b reads replica
// b % FRAG_COPIES), so the burst of identical loads at kernel start is spread
// over many L2 channels instead of hammering the few that hold one copy.
constexpr int FRAG_COPIES = 32;

// fragment counts / offsets (in 64-lane fragments) of the MFMA operand table
template <int N>
struct MS {
  static constexpr int Np = Shape<N>::Np, Nfp = Shape<N>::Nfp;
  static constexpr int NB = (Np + 15) / 16;     // 16-row blocks of a nodal vector
  static constexpr int KB = 4 * NB;             // k-steps over one nodal vector
  static constexpr int NF = 9 * Nfp;            // face rows: P, aX, bX
  // one 16-row block: the face rows are formed in registers, four face
  // nodes per lane (face_rows_inreg), 3 quantities each -- 12 k-steps;
  // two blocks: the rows are staged in LDS, 4 per k-step
  static constexpr bool INREG = NB == 1;
  static constexpr int FK = INREG ? 12 : (NF + 3) / 4;  // k-steps over the face rows
  static constexpr int FS = 4 * FK + 1;         // LDS row stride of the face rows (odd)
  static constexpr int TRW = 3 * Nfp * 2;       // trace words per element: u, dn per face node
  // LDS stride of an element's trace slot: odd, so the 16 elements of a
  // wave reading the same trace word fall on different banks
  static constexpr int SW = trace_slot_words(N);
  static constexpr int NFP3 = 3 * Nfp;          // face nodes of an element
  static constexpr int FPW = NFP3 <= 8 ? 8 : NFP3 <= 16 ? 16 : 32;  // face-phase lanes per element
  static constexpr int f1 = 0;                  // GEMM1: 2*NB out blocks x KB
  static constexpr int f2 = f1 + 2 * NB * KB;   // GEMM2 volume: NB x 2*KB
  static constexpr int fF = f2 + NB * 2 * KB;   // GEMM2 face: NB x FK
  static constexpr int fM = fF + NB * FK;       // mass: NB x KB
  static constexpr int nfrag = fM + NB * KB;
};

// Which node each accumulator position holds: lane group g, register r of
// 16-row block b.  Orders with one block (N <= 4) use a face-major layout:
// register r = f (0..2) holds N nodes of face f -- lane group g the face
// node p = g on faces 0 and 1, p = g + 1 on face 2 -- and register 3 the
// interior nodes; every other position is padding (index Np).  Each vertex
// is held once: node 0 by (0, 0), node N by (0, 1), node Np-1 by (N-1, 2).
// So every face node sits in a register of its own face and the face traces
// are written by all lanes together, plus one write per vertex for its
// second face.  Orders with two blocks keep the instructions' row order
// (node 16b + crow(g, r)).  reference.mfma_node_map is the host copy that
// lays out the A fragments to match.
template <typename T, int N>
struct RowMap {
  static constexpr int Np = Shape<N>::Np;
  static constexpr bool FACE = MS<N>::NB == 1;
  static constexpr int NINT = (N - 1) * (N - 2) / 2;  // interior nodes
  static constexpr int interior(int i) {
    for (int n = 0; n < Np; ++n) {
      bool on = false;
      for (int f = 0; f < 3; ++f) on = on || facepos<N>(f, n) >= 0;
      if (!on && i-- == 0) return n;
    }
    return Np;
  }
  static constexpr int face_node(int g, int r) {
    if (r < 3) return g < N ? fmask<N>(r, r < 2 ? g : g + 1) : Np;
    return g < NINT ? interior(g) : Np;
  }
  // register r's nodes for lane groups 0..3, one byte each
  static constexpr unsigned pack(int r) {
    return (unsigned)face_node(0, r) | ((unsigned)face_node(1, r) << 8) |
           ((unsigned)face_node(2, r) << 16) | ((unsigned)face_node(3, r) << 24);
  }
  __device__ static int node(int b, int g, int r) {
    if constexpr (FACE)
      return (int)((pack(r) >> (8 * g)) & 0xffu);
    else
      return 16 * b + MfmaT<T>::crow(g, r);
  }
  // the same, recomputed at each use: g passes through an empty asm so the
  // compiler neither hoists nor keeps the values live between uses (they
  // cost two instructions; holding them costs registers the kernels lack)
  __device__ static int node_here(int b, int g, int r) {
    asm volatile("" : "+v"(g));
    return node(b, g, r);
  }
  // compile-time check: each node is held once, each face node in a
  // register of its face (vertices by the first face listed above)
  static constexpr bool valid() {
    if (!FACE) return true;
    for (int n = 0; n < Np; ++n) {
      int seen = 0;
      for (int g = 0; g < 4; ++g)
        for (int r = 0; r < 4; ++r) seen += face_node(g, r) == n;
      if (seen != 1) return false;
    }
    for (int r = 0; r < 3; ++r)
      for (int g = 0; g < N; ++g)
        if (facepos<N>(r, face_node(g, r)) != (r < 2 ? g : g + 1)) return false;
    return face_node(0, 0) == fmask<N>(2, 0) && face_node(0, 1) == fmask<N>(0, N) &&
           face_node(N - 1, 2) == fmask<N>(1, N);
  }
};

// normal-derivative coefficients a_f, b_f of one element's three faces
template <typename T>
struct FaceAB {
  T a0, a1, a2, b0, b1, b2;
  __device__ T a(int f) const { return f == 0 ? a0 : (f == 1 ? a1 : a2); }
  __device__ T b(int f) const { return f == 0 ? b0 : (f == 1 ? b1 : b2); }
};

// GEMM 1 for 16 elements of a wave and their face traces into LDS slots.
template <typename T, int N>
__device__ __forceinline__ void grad_and_traces(const T (&A1)[2 * MS<N>::NB][MS<N>::KB],
                                                const T (&U)[MS<N>::NB][4],
                                                T (&UR)[MS<N>::NB][4], T (&US)[MS<N>::NB][4],
                                                const FaceAB<T>& ab, T* tr, int g) {
  using Q = MS<N>;
  using MF = MfmaT<T>;
  using v4 = typename MF::v4;
  constexpr int NB = Q::NB, KB = Q::KB, Nfp = Q::Nfp, Np = Q::Np;
  using RM = RowMap<T, N>;
  static_assert(RM::valid(), "face-major node layout");
#pragma unroll
  for (int b = 0; b < 2 * NB; ++b) {
    v4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < KB; ++t) acc = MF::mma(A1[b][t], U[t / 4][t % 4], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (b < NB)
        UR[b][r] = acc[r];
      else
        US[b - NB][r] = acc[r];
    }
  }
  if constexpr (RM::FACE) {
    // register f holds face f's nodes p = g (p = g + 1 on face 2)
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const int j = f * Nfp + (f < 2 ? g : g + 1);
      if (g < N) {
        tr[j * 2] = U[0][f];
        tr[j * 2 + 1] = ab.a(f) * UR[0][f] + ab.b(f) * US[0][f];
      }
    }
    // the vertices' second faces: node 0 is face 2's p = 0, node N face 0's
    // p = N (both lane group 0), node Np-1 face 1's p = N (lane group N-1)
    if (g == 0) {
      tr[(2 * Nfp) * 2] = U[0][0];
      tr[(2 * Nfp) * 2 + 1] = ab.a(2) * UR[0][0] + ab.b(2) * US[0][0];
      tr[N * 2] = U[0][1];
      tr[N * 2 + 1] = ab.a(0) * UR[0][1] + ab.b(0) * US[0][1];
    }
    if (g == N - 1) {
      tr[(Nfp + N) * 2] = U[0][2];
      tr[(Nfp + N) * 2 + 1] = ab.a(1) * UR[0][2] + ab.b(1) * US[0][2];
    }
    return;
  }
  // two blocks: face node (f, p) is node n = fmask(f, p), held by lane group
  // gn in register (b, r): all three are constants after unrolling, so
  // every face node is one write predicated on the lane group
#pragma unroll
  for (int f = 0; f < 3; ++f)
#pragma unroll
    for (int p = 0; p < Nfp; ++p) {
      const int n = fmask<N>(f, p), b = n / 16, m = n % 16;
      const int gn = MF::crow(1, 0) == 1 ? m % 4 : m / 4;  // inverse of crow
      const int r = MF::crow(1, 0) == 1 ? m / 4 : m % 4;
      if (g == gn) {
        tr[(f * Nfp + p) * 2] = U[b][r];
        tr[(f * Nfp + p) * 2 + 1] = ab.a(f) * UR[b][r] + ab.b(f) * US[b][r];
      }
    }
}

// Face rows of GEMM 2 in the accumulator layout (one 16-row block): lane
// (g, e) forms P, aX, bX of its element at the face nodes its registers
// hold -- register f's node p = g (p = g + 1 on face 2) for g < N -- and,
// for g < 3, at one vertex's second face (g = 0: face 2, p = 0; g = 1:
// face 0, p = N; g = 2: face 1, p = N), so the 3 Nfp face nodes are covered
// once.  Own and neighbour traces both come from LDS.  Slot s (s = f for the
// register faces, 3 for the vertex) gives the B operands of k-steps 3 s + q,
// q = 0, 1, 2 for P, aX, bX (reference.mfma_face_row is the host copy that
// lays out the lift's columns to match; a lane with no node in a slot still
// forms the finite values of face node (0, 0), which meet zero columns).
template <typename T, int N, bool UDIR>
__device__ __forceinline__ void face_rows_inreg(T (&Fv)[12], const T* s_x, const T* re,
                                                const int* s_fc, int el, int g,
                                                const T* __restrict__ udir, int kg) {
  constexpr int Nfp = N + 1, Np = Shape<N>::Np, SW = MS<N>::SW;
  // the vertex slot's face and node for lane groups 0..3, 4 bits each
  constexpr unsigned VF = 0x2102u, VP = (unsigned)(N << 4) | (unsigned)(N << 8);
  auto slot = [&](int s, int& f, int& p) {
    if (s < 3) {
      f = s;
      p = g < N ? (s < 2 ? g : g + 1) : 0;
    } else {
      f = (int)((VF >> (4 * g)) & 15u);
      p = (int)((VP >> (4 * g)) & 15u);
    }
  };
  // Dirichlet data at the four face nodes, loaded up front (no load in a
  // branch; only Dirichlet faces use them)
  T uDv[4] = {};
  if constexpr (UDIR) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      int f, p;
      slot(s, f, p);
      uDv[s] = udir[(size_t)kg * Np + fmask<N>(f, p)];
    }
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    int f, p;
    slot(s, f, p);
    const int j = f * Nfp + p;
    const T uM = s_x[el * SW + 2 * j], dnM = s_x[el * SW + 2 * j + 1];
    const int c = s_fc[f * TILE + el];
    if constexpr (!UDIR) {
      // the neighbour's trace at the face code's address: an interior
      // face's neighbour (its dn taken with its own outward normal), a
      // Neumann face's own trace (jump and flux sum exactly 0), a
      // Dirichlet face's zero slot (u^+ = 0, no neighbour flux)
      const T* xp = s_x + (fc_nbase(c) + p * fc_nstep(c));
      const T uN = xp[0], dN = xp[1];
      const T ju = uM - uN;
      const T sj = re[R_SJW + f];
      const T X = -sj * ju;
      Fv[3 * s] = re[R_PEN + f] * ju - sj * (dnM - dN);
      Fv[3 * s + 1] = re[R_A + f] * X;
      Fv[3 * s + 2] = re[R_B + f] * X;
      continue;
    }
    const int ty = fc_type(c);
    const int pn = fc_rev(c) ? Nfp - 1 - p : p;
    // the neighbour's trace, read for every face (slot 0 for a boundary
    // face) and weighted by the face type, so no read sits in a branch; its
    // dn was taken with its own outward normal
    const T* xp = s_x + fc_slot(c) * SW + 2 * (fc_face(c) * Nfp + pn);
    const T uN = xp[0], dN = xp[1];
    const T cI = ty == FT_INTERIOR ? T(1) : T(0);
    const T cD = ty == FT_DIRICHLET ? T(1) : T(0);
    T ju = cI * (uM - uN);  // [u]: u^- - u^+ (Neumann: 0)
    if constexpr (UDIR) {
      ju += cD * (uM - uDv[s]);
    } else {
      ju += cD * uM;
    }
    const T dnP = -cI * dN;
    const T sj = re[R_SJW + f];
    const T X = -sj * ju;
    Fv[3 * s] = re[R_PEN + f] * ju - sj * (dnM + dnP);
    Fv[3 * s + 1] = re[R_A + f] * X;
    Fv[3 * s + 2] = re[R_B + f] * X;
  }
}

// Tile handled by workgroup b of a G-workgroup launch.  Workgroups are
// dispatched to the 8 XCDs round-robin (b % 8); each XCD gets one contiguous
// run of tiles, so a tile's halo neighbours -- mostly the adjacent tiles of a
// compact element order -- are fetched into the same XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int G) {
  constexpr int NX = 8;
  const int q = G / NX, r = G % NX, x = b % NX;
  return x * q + (x < r ? x : r) + b / NX;
}

// Minimum waves per SIMD the register allocator is held to: N <= 4 fits 128
// registers in float64 (4 waves) and 64 in float32 (8 waves, which the
// float32 LDS footprint also allows; 7 at N = 3, which spills at 64).
// Orders with two 16-row blocks (N >= 5) hold 2 x 48 A fragments and fit
// 256 registers in float64 (2 waves: twice the workgroups of one wave at
// 512, N = 6 91 -> 67 us at 2x256x256; 3 waves spill) and 128 in float32.
template <typename T, int N>
struct ApplyWaves {
  static constexpr int value =
      N <= 4 ? (sizeof(T) == 8 ? 4 : (N == 3 ? 7 : 8)) : (sizeof(T) == 8 ? 2 : 4);
};

// One 64-element tile per workgroup (grid = number of tiles).  HC is the
// halo capacity of a tile (32 or 64 out-of-tile neighbour elements).
template <typename T, int N, int MODE, int HC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ApplyWaves<T, N>::value)))
sipg_apply_mfma_kernel(const T* __restrict__ frag, const T* __restrict__ rec,
                       const int* __restrict__ fcode, const int* __restrict__ perm,
                       const int* __restrict__ hpos, const T* __restrict__ hab,
                       const T* __restrict__ u, T* __restrict__ y, const T* __restrict__ udir,
                       T alpha, T shift, int K, double* __restrict__ partial,
                       const int* __restrict__ skip, const T* __restrict__ bvec,
                       const T* __restrict__ binv, T omega, const T* __restrict__ pmat,
                       int npt, T* __restrict__ xco, CGScalars* __restrict__ sc,
                       const double* __restrict__ rzp, int nrz, int q) {
  using Q = MS<N>;
  using MF = MfmaT<T>;
  using v4 = typename MF::v4;
  constexpr int Np = Q::Np, Nfp = Q::Nfp, NB = Q::NB, KB = Q::KB, FK = Q::FK;
  constexpr int NFP3 = Q::NFP3, FPW = Q::FPW;
  // LDS layout of GEMM 1's output: u and dn at the face nodes only (the
  // traces a neighbour reads), so four float64 / eight float32 workgroups
  // fit in a compute unit's LDS -- eight float32 tiles per CU cover the 2048
  // tiles of a 2x256x256 mesh in one round.
  constexpr int SW = Q::SW;                 // LDS words per element slot
  constexpr bool HELM = MODE & M_HELM, UDIR = MODE & M_UDIR, DOT = MODE & M_DOT;
  constexpr bool RESID = MODE & M_RESID;
  // smoothing sweep: y = u + omega Binv (bvec - A u) (the epilogue below)
  constexpr bool SMOOTH = MODE & M_SMOOTH;
  // with RESID: the residual restricted to the coarser p-level instead of
  // stored (the epilogue below): y receives P^T (bvec - A u) per element
  // (pmat: Np x np_lo), xco, when given, the coarse level's first sweep
  // omega Binv_c (that) with binv the coarse level's packed blocks
  constexpr bool RESTRICT = MODE & M_RESTRICT;
  static_assert(!RESTRICT || (RESID && !SMOOTH && !DOT), "restriction of a residual");
  // with SMOOTH: the sweep's result is also prolongated into the next-finer
  // p-level's iterate, xco[k] += pmat x_new[k] (pmat: npt x Np)
  constexpr bool PROLONG_OUT = MODE & M_PROLONG_OUT;
  static_assert(!PROLONG_OUT || SMOOTH, "prolongation of a sweep's result");
  // with DOT: CG's direction update first, p = u + beta bvec (u = z, bvec =
  // p_old), p written to xco for the tile (the prologue below)
  constexpr bool PUPD = MODE & M_PUPDATE;
  static_assert(!PUPD || (DOT && !UDIR && !RESID && !SMOOTH), "direction update of a CG apply");
  // the transfer's other level: np_lo = npt coarse nodes (RESTRICT)
  const int np_lo = npt;
  constexpr bool ROWS = SMOOTH || RESTRICT;  // epilogue on element rows in LDS
  constexpr int RECS = DDGX_REC_WORDS + 1;  // LDS stride of a record (odd: no bank conflicts)
  // element slots: tile, halo, then the zero slot Dirichlet faces read
  constexpr int XT = (TILE + HC + 1) * SW;
  // face rows staged in LDS (two blocks), or the smoothing epilogue's rows
  constexpr int XF = Q::INREG ? (ROWS ? 2 * TILE * Np : 0) : TILE * Q::FS;
  constexpr int XW = XT > XF ? XT : XF;
  static_assert(!ROWS || 2 * TILE * Np <= XW, "epilogue rows fit the face-row buffer");
  // s_x holds GEMM 1's output for the tile and its halo during the face
  // phase and, after a barrier, the face rows F that feed GEMM 2 (and, in
  // the smoothing epilogue, the residual and iterate rows).
  __shared__ T s_x[XW];
  __shared__ T s_rec[TILE * RECS];
  __shared__ T s_hab[6 * HC];
  __shared__ int s_fc[3 * TILE];
  __shared__ double s_red[4];
  // Fragments through LDS (one 16-row block): the workgroup loads each
  // fragment once -- 256 threads, a quarter of the words each -- instead of
  // every wave loading all of them, a quarter of the L2-to-CU traffic.
  // GEMM 1's fragments have a region of their own.  GEMM 2's (and the mass
  // fragments of the Helmholtz modes, which follow them in the table) have
  // one too where the LDS still holds ApplyWaves workgroups per compute unit
  // (A2OWN: float64 N = 2, float32 N = 2, 3); otherwise they are staged into
  // s_x once the face rows have been formed and the traces there are no
  // longer read, at the cost of two more barriers.
  constexpr bool FLDS = Q::INREG;
  constexpr int NA1 = 2 * NB * KB;                              // GEMM 1
  constexpr int NA2 = NB * (2 * KB + FK) + (HELM ? NB * KB : 0);  // GEMM 2 (+ mass)
  // (LDS is allocated in 512-byte granules)
  constexpr int LDS_OWN = ((int)sizeof(T) * (XW + TILE * RECS + 6 * HC + (NA1 + NA2) * 64) +
                           (int)sizeof(int) * 3 * TILE + (int)sizeof(double) * 4 + 511) / 512 * 512;
  constexpr bool A2OWN = FLDS && (LDS_BYTES_PER_CU / LDS_OWN) >= ApplyWaves<T, N>::value;
  constexpr int A1PT = FLDS ? NA1 / 4 : 1, A2PT = FLDS ? NA2 / 4 : 1;
  static_assert(!FLDS || (NA1 % 4 == 0 && NA2 % 4 == 0 && NA2 * 64 <= XW),
                "fragment staging split and fit");
  static_assert(Q::fM == Q::f2 + NB * (2 * KB + FK), "mass fragments follow GEMM 2's");
  __shared__ T s_a1[FLDS ? (A2OWN ? NA1 + NA2 : NA1) * 64 : 1];
  T* s_a2 = A2OWN ? s_a1 + NA1 * 64 : s_x;

  if constexpr (!PUPD) {
    if (skip != nullptr && *skip) return;
  }

  const int tid = threadIdx.x;
  // the wave index is uniform: held in a scalar register, so the staging
  // addresses below are scalar bases plus the lane
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int g = lane >> 4, e = lane & 15;
  const int el = 16 * wv + e;  // this lane's element (and halo entry) within the tile
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int k0 = tile * TILE;
  const int ne = min(TILE, K - k0);
  const int k = k0 + el;  // this lane's tile position
  const bool valid = el < ne;

  // ---- global loads of the tile ------------------------------------------
  // Every load of the tile is issued before the first wait on one, with
  // clamped addresses and the out-of-range values selected away afterwards
  // (no load sits in a lane-divergent branch): a workgroup pays one memory
  // round trip for its own data and one more for the halo, whose element
  // list is the only index a nodal load depends on.  The node index of each
  // accumulator position is computed once here, so the loads can be
  // scheduled back to back.
  int kg = valid ? k : 0;  // this lane's element id
  if (perm != nullptr) {   // uniform branch
    const int pv = perm[valid ? k : k0];
    kg = valid ? pv : 0;
  }
  int nn[NB][4];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) nn[b][r] = RowMap<T, N>::node(b, g, r);
  // Nodal values are loaded from clamped addresses and used unmasked: a
  // padding position (n = Np) holds a finite copy of node Np - 1 and meets
  // zero columns in every fragment table, and a lane past the tile's end
  // computes a column of its own whose results are never stored.
  T U[NB][4];
  // the right-hand side of the float64 residual and smoothing modes, loaded
  // with u so the epilogue does not wait on a load issued after GEMM 2
  // (float32 reads it in the epilogue: it has no registers to spare)
  constexpr bool NEEDB = (RESID || SMOOTH) && sizeof(T) == 8;
  T BV[NB][4];  // unused (and dropped by the compiler) unless NEEDB
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = nn[b][r];
      const unsigned go = (unsigned)(kg * Np + (n < Np ? n : Np - 1)) * (unsigned)sizeof(T);
      U[b][r] = ld_off(u, go);
      if constexpr (NEEDB || PUPD) BV[b][r] = ld_off(bvec, go);
    }
  // halo: the element ids of the tile's out-of-tile neighbours, 16 per
  // wave, listed from entry 0 (-1 past the end); a wave past the list skips
  // the halo work
  const int hkv = hpos[(size_t)tile * HC + (el < HC ? el : HC - 1)];
  // records, face codes and halo coefficients into registers, then LDS
  constexpr int RPT = DDGX_REC_WORDS * TILE / 256;  // record words per thread
  constexpr int HPT = (6 * HC + 255) / 256;         // halo coefficients per thread
  static_assert(DDGX_REC_WORDS * TILE % 256 == 0 && 3 * TILE <= 256, "staging split");
  T rc[RPT], hb[HPT];
  // thread (wave wv, lane) stages record word wv + 4 j and face code wv
  // (waves 0-2) of tile element lane, and halo coefficient 64 (wv + 4 j) + lane
  static_assert(TILE == 64, "staging by wave and lane");
  const int le = lane < ne ? lane : ne - 1;
#pragma unroll
  for (int j = 0; j < RPT; ++j)
    rc[j] = ld_off(rec + (size_t)(wv + 4 * j) * K, (unsigned)(k0 + le) * (unsigned)sizeof(T));
  const int fcv = ld_off(fcode + (size_t)(wv < 3 ? wv : 2) * K, (unsigned)(k0 + le) * 4u);
#pragma unroll
  for (int j = 0; j < HPT; ++j) {
    const int i = 64 * (wv + 4 * j) + lane;
    hb[j] = hab[(size_t)tile * 6 * HC + (i < 6 * HC ? i : 0)];
  }
  // fragments, from one of the table's replicas (the many workgroups that
  // start together spread these loads over L2 channels): with FLDS a quarter
  // of GEMM 1's and GEMM 2's words per thread, else GEMM 1's per lane
  const T* fr = frag + (size_t)(blockIdx.x % FRAG_COPIES) * Q::nfrag * 64;
  T A1[2 * NB][KB];
  T a1v[A1PT], a2v[A2PT];
  if constexpr (FLDS) {
#pragma unroll
    for (int j = 0; j < A1PT; ++j) a1v[j] = fr[(size_t)Q::f1 * 64 + tid + 256 * j];
#pragma unroll
    for (int j = 0; j < A2PT; ++j) a2v[j] = fr[(size_t)Q::f2 * 64 + tid + 256 * j];
  } else {
#pragma unroll
    for (int b = 0; b < 2 * NB; ++b)
#pragma unroll
      for (int t = 0; t < KB; ++t) A1[b][t] = fr[(size_t)(Q::f1 + b * KB + t) * 64 + lane];
  }
  // the halo's nodal values, once its element list has arrived (a lane past
  // the list loads element 0's); no wait on these loads comes before the
  // staging writes below
  const int hk = el < HC ? hkv : -1;
  const bool hwave = __ballot(hk >= 0) != 0;  // uniform per wave
  T UH[NB][4] = {};
  T PH[NB][4] = {};  // the halo's p_old (direction update only)
  if (hwave) {
    const int hg = hk >= 0 ? hk : 0;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = nn[b][r];
        const unsigned ho = (unsigned)(hg * Np + (n < Np ? n : Np - 1)) * (unsigned)sizeof(T);
        UH[b][r] = ld_off(u, ho);
        if constexpr (PUPD) PH[b][r] = ld_off(bvec, ho);
      }
  }
  if constexpr (PUPD) {
    // CG's direction update, with the tile's and the halo's loads in flight:
    // gamma = <r, z> and <r, r> from the update's partials, in the order every
    // other reduction of the solve uses.  Every workgroup takes the same stop
    // decision from them, so none depends on what block 0 writes here.
    __shared__ double s_pr[4];
    double s0, s1;
    block_sum2_partials(rzp, nrz, s_pr, s0, s1);
    const double rs = sc->use_pc ? s1 : s0, g_old = sc->gq[q ^ 1];
    const int kb = sc->kq[q];
    const bool stop = sc->done || !(rs > sc->atol2 && kb < sc->maxiter && s0 != 0.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->rr = rs;
      sc->rs = rs;
      if (stop) {
        sc->done = 1;
      } else {
        sc->gamma = s0;
        sc->gq[q] = s0;
        sc->beta = s0 / g_old;
        sc->kq[q ^ 1] = kb + 1;
        sc->k = kb + 1;
      }
    }
    if (stop) return;
    const T beta = (T)(s0 / g_old);
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        U[b][r] += beta * BV[b][r];
        UH[b][r] += beta * PH[b][r];
      }
    // the tile's new direction (padding positions repeat node Np - 1)
    if (valid) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (nn[b][r] < Np)
            st_off(xco, (unsigned)(kg * Np + nn[b][r]) * (unsigned)sizeof(T), U[b][r]);
    }
  }
  // records and codes past the tile's end are the clamped copies of its
  // last element's, read only by the lanes past the end
#pragma unroll
  for (int j = 0; j < RPT; ++j) s_rec[lane * RECS + wv + 4 * j] = rc[j];
  if (wv < 3) s_fc[tid] = fcv;
#pragma unroll
  for (int j = 0; j < HPT; ++j) {
    const int i = tid + 256 * j;
    if (i < 6 * HC) s_hab[i] = hb[j];
  }
  if (tid < SW) s_x[(TILE + HC) * SW + tid] = T(0);  // the zero slot
  if constexpr (FLDS) {
#pragma unroll
    for (int j = 0; j < A1PT; ++j) s_a1[tid + 256 * j] = a1v[j];
    if constexpr (A2OWN) {
#pragma unroll
      for (int j = 0; j < A2PT; ++j) s_a2[tid + 256 * j] = a2v[j];
    }
  }
  __syncthreads();
  if constexpr (FLDS) {
#pragma unroll
    for (int b = 0; b < 2 * NB; ++b)
#pragma unroll
      for (int t = 0; t < KB; ++t) A1[b][t] = s_a1[(b * KB + t) * 64 + lane];
  }

  // ---- GEMM 1: own elements and halo into LDS ----------------------------
  const T* re = s_rec + el * RECS;
  T UR[NB][4], US[NB][4];
  {
    const FaceAB<T> ab{re[R_A], re[R_A + 1], re[R_A + 2], re[R_B], re[R_B + 1], re[R_B + 2]};
    grad_and_traces<T, N>(A1, U, UR, US, ab, s_x + el * SW, g);
    if (hwave) {
      // a lane past the halo list computes the gradient of element 0 (its
      // clamped load) into a slot no face code names
      const FaceAB<T> abh{s_hab[el], s_hab[HC + el], s_hab[2 * HC + el],
                          s_hab[3 * HC + el], s_hab[4 * HC + el], s_hab[5 * HC + el]};
      T HR[NB][4], HS[NB][4];
      grad_and_traces<T, N>(A1, UH, HR, HS, abh, s_x + (TILE + el) * SW, g);
    }
  }
  // metric products (in place): Wr, Ws
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const T a = UR[b][r], c = US[b][r];
      UR[b][r] = re[R_VRR] * a + re[R_VRS] * c;
      US[b][r] = re[R_VRS] * a + re[R_VSS] * c;
    }
  // GEMM 2 fragments (two blocks), issued here so their latency overlaps
  // the face phase
  T A2[NB][2 * KB + FK];
  if constexpr (!FLDS) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
      for (int t = 0; t < 2 * KB; ++t) A2[b][t] = fr[(size_t)(Q::f2 + b * 2 * KB + t) * 64 + lane];
#pragma unroll
      for (int t = 0; t < FK; ++t) A2[b][2 * KB + t] = fr[(size_t)(Q::fF + b * FK + t) * 64 + lane];
    }
  }
  __syncthreads();  // the tile's and the halo's traces are in LDS

  // ---- face phase: P, aX, bX per (element, face node) ---------------------
  T Fv[Q::INREG ? 12 : 1];
  if constexpr (Q::INREG) {
    face_rows_inreg<T, N, UDIR>(Fv, s_x, re, s_fc, el, g, udir, kg);
    // GEMM 2's fragments into s_x once every wave has read its traces
    if constexpr (FLDS && !A2OWN) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < A2PT; ++j) s_x[tid + 256 * j] = a2v[j];
      __syncthreads();
    }
    if constexpr (FLDS) {
#pragma unroll
      for (int t = 0; t < 2 * KB + FK; ++t) A2[0][t] = s_a2[t * 64 + lane];
    }
  } else {
    // thread -> one face node (f, p) of elements eg, eg + EG, ...
    constexpr int EG = 256 / FPW, ITE = TILE / EG;
    const int fp = tid % FPW, eg = tid / FPW;
    const bool fpa = fp < NFP3;
    const int ff = fpa ? fp / Nfp : 0, pp = fpa ? fp % Nfp : 0;
    const int nM = fmask<N>(ff, pp);
    T fv[ITE][3];
#pragma unroll
    for (int it = 0; it < ITE; ++it) {
      const int e2 = eg + EG * it;
      fv[it][0] = fv[it][1] = fv[it][2] = T(0);
      if (fpa && e2 < ne) {
        const T* rr = s_rec + e2 * RECS;
        const T* xm = s_x + e2 * SW;
        const T aM = rr[R_A + ff], bM = rr[R_B + ff];
        const T uM = xm[fp * 2], dnM = xm[fp * 2 + 1];
        const int c = s_fc[ff * TILE + e2];
        const int ty = fc_type(c);
        T uP, dnP;
        if (ty == FT_INTERIOR) {
          const int sl = fc_slot(c), fn = fc_face(c);
          const int pn = fc_rev(c) ? (Nfp - 1 - pp) : pp;
          // the neighbour's trace; its dn was taken with its own outward normal
          const T* xp = s_x + sl * SW;
          uP = xp[(fn * Nfp + pn) * 2];
          dnP = -xp[(fn * Nfp + pn) * 2 + 1];
        } else if (ty == FT_DIRICHLET) {
          if constexpr (UDIR) {
            const int kk = k0 + e2;
            uP = udir[(size_t)(perm ? perm[kk] : kk) * Np + nM];
          } else {
            uP = T(0);
          }
          dnP = T(0);
        } else {
          uP = uM;
          dnP = T(0);
        }
        const T ju = uM - uP;
        const T X = -rr[R_SJW + ff] * ju;
        fv[it][0] = rr[R_PEN + ff] * ju - rr[R_SJW + ff] * (dnM + dnP);
        fv[it][1] = aM * X;
        fv[it][2] = bM * X;
      }
    }
    __syncthreads();  // nodal values consumed: s_x becomes the face rows
#pragma unroll
    for (int it = 0; it < ITE; ++it) {
      const int e2 = eg + EG * it;
      if (fpa) {
        s_x[e2 * Q::FS + fp] = fv[it][0];
        s_x[e2 * Q::FS + NFP3 + fp] = fv[it][1];
        s_x[e2 * Q::FS + 2 * NFP3 + fp] = fv[it][2];
      }
    }
    constexpr int PADC = 4 * FK - Q::NF;
    if constexpr (PADC > 0) {
      for (int i = tid; i < TILE * PADC; i += 256) s_x[(i / PADC) * Q::FS + Q::NF + i % PADC] = T(0);
    }
    __syncthreads();
  }

  // ---- GEMM 2: volume + face lift (+ mass) -------------------------------
  T Jm = T(0);
  if constexpr (HELM) Jm = alpha * re[R_J];
  double dacc = 0.0;
  T vv[NB][4];  // A u, kept for the smoothing epilogue
#pragma unroll
  for (int bo = 0; bo < NB; ++bo) {
    // three independent accumulator chains (volume r, volume s, faces) so
    // consecutive matrix-core products do not wait on each other
    v4 acc = {0, 0, 0, 0}, acs = {0, 0, 0, 0}, acf = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < KB; ++t) {
      acc = MF::mma(A2[bo][t], UR[t / 4][t % 4], acc);
      acs = MF::mma(A2[bo][KB + t], US[t / 4][t % 4], acs);
    }
#pragma unroll
    for (int t = 0; t < FK; ++t) {
      T fb;
      if constexpr (Q::INREG)
        fb = Fv[t];
      else
        fb = s_x[el * Q::FS + 4 * t + g];
      acf = MF::mma(A2[bo][2 * KB + t], fb, acf);
    }
    if constexpr (HELM) {
#pragma unroll
      for (int t = 0; t < KB; ++t)
        acc = MF::mma(FLDS ? s_a2[(NB * (2 * KB + FK) + bo * KB + t) * 64 + lane]
                           : fr[(size_t)(Q::fM + bo * KB + t) * 64 + lane],
                      Jm * U[t / 4][t % 4], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = RowMap<T, N>::node_here(bo, g, r);
      T v = acc[r] + acs[r] + acf[r];
      if constexpr (HELM) v += shift * U[bo][r];
      if constexpr (ROWS) {
        vv[bo][r] = v;
      } else if (valid && n < Np) {
        const unsigned go = (unsigned)(kg * Np + n) * (unsigned)sizeof(T);
        if constexpr (RESID)
          st_off(y, go, (NEEDB ? BV[bo][r] : ld_off(bvec, go)) - v);
        else
          st_off(y, go, v);
        if constexpr (DOT) dacc += (double)U[bo][r] * (double)v;
      }
    }
  }

  // the float32 right-hand side of the row epilogues, every value loaded
  // (clamped) before the first is used: one memory round trip, not one per
  // register
  T BR[NB][4];
  if constexpr (ROWS && !NEEDB) {
#pragma unroll
    for (int bo = 0; bo < NB; ++bo)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = RowMap<T, N>::node_here(bo, g, r);
        BR[bo][r] = ld_off(bvec, (unsigned)(kg * Np + (n < Np ? n : Np - 1)) * (unsigned)sizeof(T));
      }
  }

  if constexpr (SMOOTH) {
    // the residual r = b - A u and u, by element rows, into LDS; then one
    // block-Jacobi row per lane (Np of NL lanes per element): the same
    // arithmetic as a residual apply followed by bj_smooth_kernel
    constexpr int NPK = SymPack<N>::NPK, NL = Shape<N>::NL, EPB = 256 / NL;
    T* s_r = s_x;
    T* s_u = s_x + TILE * Np;
    __syncthreads();  // every wave is done reading the face rows
#pragma unroll
    for (int bo = 0; bo < NB; ++bo)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = RowMap<T, N>::node_here(bo, g, r);
        if (valid && n < Np) {
          s_r[el * Np + n] = (NEEDB ? BV[bo][r] : BR[bo][r]) - vv[bo][r];
          s_u[el * Np + n] = U[bo][r];
        }
      }
    __syncthreads();
    const int li = tid % NL, ep = tid / NL;
    for (int e0 = 0; e0 < TILE; e0 += EPB) {
      const int e2 = e0 + ep;
      if (li < Np && e2 < ne) {
        const int kk = k0 + e2;
        const size_t k2 = (size_t)(perm ? perm[kk] : kk);
        const T* bk = binv + k2 * NPK;
        const T a = block_row_diag<T, Np>(bk, [&](int j) { return s_r[e2 * Np + j]; }, li);
        const T xn = s_u[e2 * Np + li] + omega * a;
        y[k2 * Np + li] = xn;
        // only this lane reads s_u[e2 * Np + li]: the new value takes its place
        if constexpr (PROLONG_OUT) s_u[e2 * Np + li] = xn;
      }
    }
    if constexpr (PROLONG_OUT) {
      // one thread per (element, fine node): x_f += P x_new, element-local
      __syncthreads();
      for (int it = tid; it < ne * npt; it += 256) {
        const int e2 = it / npt, i = it % npt;
        const int kk = k0 + e2;
        const size_t k2 = (size_t)(perm ? perm[kk] : kk);
        T acc = 0;
#pragma unroll
        for (int l = 0; l < Np; ++l) acc += pmat[i * Np + l] * s_u[e2 * Np + l];
        xco[k2 * npt + i] += acc;
      }
    }
  }

  if constexpr (RESTRICT) {
    // the residual rows into LDS, then one thread per (element, coarse node)
    T* s_r = s_x;
    T* s_lo = s_x + TILE * Np;
    __syncthreads();  // every wave is done reading the traces
#pragma unroll
    for (int bo = 0; bo < NB; ++bo)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = RowMap<T, N>::node_here(bo, g, r);
        if (valid && n < Np) s_r[el * Np + n] = (NEEDB ? BV[bo][r] : BR[bo][r]) - vv[bo][r];
      }
    __syncthreads();
    for (int it = tid; it < ne * np_lo; it += 256) {
      const int e2 = it / np_lo, j = it % np_lo;
      T lo = 0;
#pragma unroll
      for (int h = 0; h < Np; ++h) lo += pmat[h * np_lo + j] * s_r[e2 * Np + h];
      const int kk = k0 + e2;
      y[(size_t)(perm ? perm[kk] : kk) * np_lo + j] = lo;
      s_lo[it] = lo;
    }
    if (xco != nullptr) {
      __syncthreads();
      const int npk = np_lo * (np_lo + 1) / 2;
      for (int it = tid; it < ne * np_lo; it += 256) {
        const int e2 = it / np_lo, j = it % np_lo;
        const int kk = k0 + e2;
        const size_t k2 = (size_t)(perm ? perm[kk] : kk);
        const T* bk = binv + k2 * npk;
        T a = 0;
        for (int d = 0; d < np_lo; ++d) {
          const int i = j + d < np_lo ? j + d : j + d - np_lo;
          a += bk[diag_pk(np_lo, j, i)] * s_lo[e2 * np_lo + i];
        }
        xco[k2 * np_lo + j] = omega * a;
      }
    }
  }

  if constexpr (DOT) {
    for (int off = 32; off > 0; off >>= 1) dacc += __shfl_down(dacc, off, 64);
    if (lane == 0) s_red[wv] = dacc;
    __syncthreads();
    // one partial per tile, at the tile's index (the tiles of a batched
    // solve's request are then one contiguous range of partials)
    if (tid == 0) partial[tile] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
  }
}

}  // namespace ddgx
