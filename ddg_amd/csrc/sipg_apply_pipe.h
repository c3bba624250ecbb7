// Software-pipelined variant of the matrix-core SIPG apply (same arithmetic
// as sipg_apply_mfma_kernel, src/ddg/dg2d/sipg.py:142-337 in closed form):
// persistent workgroups that overlap the global loads of their next tile
// with the matrix-core and face work of the current one.
//
// Why a second kernel: with one tile per workgroup all workgroups of a
// round load, compute and store in lockstep, so the kernel takes about the
// sum of its memory and compute phases.  Here a workgroup walks a
// contiguous run of tiles (inside its XCD's run) and, at the top of each
// tile, issues the next tile's loads into registers.  Two details make the
// overlap real:
//   * the A fragments sit in LDS for the workgroup's lifetime, so the body
//     reads them with LDS loads (lgkmcnt) -- a global fragment load would
//     make the body wait on the in-flight prefetch, whose vector-memory
//     counter completes in order;
//   * the only dependent index, the tile's halo list, is loaded with the
//     previous tile's data, so no prefetch waits on an index, and no value
//     of an in-flight load is consumed before the next tile's wait;
//   * the body's barriers order LDS only (lds_barrier).
// Float64 orders 2-4 (one 16-row block) in the identity tile order (forest
// meshes), for the plain and Helmholtz applies; everything else uses the
// one-tile-per-workgroup kernel.  At 2x256x256 this takes the float64 apply
// from 26.4 to 25.6 us (N=4), 23.6 to 22.6 (N=3), 20.9 to 20.5 (N=2).

#pragma once

#include "sipg_apply_mfma.h"

namespace ddgx {

// Contiguous tile run [t0, t1) of workgroup b of G: each XCD (b % 8) owns one
// contiguous run of tiles, split evenly over that XCD's workgroups (with
// fewer than 8 workgroups, G groups of one workgroup each cover all tiles).
__device__ __forceinline__ void pipe_tiles(int b, int G, int ntiles, int& t0, int& t1) {
  const int NX = G < 8 ? G : 8;
  const int x = b % NX, j = b / NX;
  const int gx = G / NX + (x < G % NX ? 1 : 0);
  const int q = ntiles / NX, r = ntiles % NX;
  const int c0 = x * q + (x < r ? x : r), cn = q + (x < r ? 1 : 0);
  const int qq = cn / gx, rr = cn % gx;
  t0 = c0 + j * qq + (j < rr ? j : rr);
  t1 = t0 + qq + (j < rr ? 1 : 0);
}

// Workgroup barrier that orders LDS only.  __syncthreads() also orders
// global memory, which makes every barrier wait for the next tile's
// prefetch still in flight; the body's barriers only guard LDS.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T, int N, int HC>
struct PipeIn {
  static constexpr int RPT = DDGX_REC_WORDS * TILE / 256;  // record words per thread
  static constexpr int HPT = (6 * HC + 255) / 256;           // halo coefficients per thread
  T U[4], UH[4], B[4];
  T rc[RPT];
  T hb[HPT];
  int fc;
  int hkn;  // this lane's halo entry of the following tile (clamped load)
};

template <typename T, int N, int MODE, int HC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
sipg_apply_pipe_kernel(const T* __restrict__ frag, const T* __restrict__ rec,
                       const int* __restrict__ fcode, const int* __restrict__ hpos,
                       const T* __restrict__ hab, const T* __restrict__ u, T* __restrict__ y,
                       const T* __restrict__ udir, T alpha, T shift, int K,
                       double* __restrict__ partial, const int* __restrict__ skip,
                       const T* __restrict__ bvec) {
  using Q = MS<N>;
  using MF = MfmaT<T>;
  using v4 = typename MF::v4;
  using PI = PipeIn<T, N, HC>;
  constexpr int Np = Q::Np, Nfp = Q::Nfp, NB = Q::NB, KB = Q::KB, FK = Q::FK;
  constexpr int SW = Q::TRW;
  static_assert(NB == 1, "one 16-row block per nodal vector");
  static_assert(DDGX_REC_WORDS * TILE % 256 == 0, "record words split evenly");
  constexpr bool HELM = MODE & M_HELM, UDIR = MODE & M_UDIR, DOT = MODE & M_DOT;
  constexpr bool RESID = MODE & M_RESID;
  constexpr int RECS = DDGX_REC_WORDS + 1;
  static_assert(Q::INREG, "face rows formed in registers");
  constexpr int XW = (TILE + HC) * SW;  // element slots: tile then halo
  __shared__ T s_fr[Q::nfrag * 64];
  __shared__ T s_x[XW];
  __shared__ T s_rec[TILE * RECS];
  __shared__ T s_hab[6 * HC];
  __shared__ int s_fc[3 * TILE];
  __shared__ double s_red[4];

  if (skip != nullptr && *skip) return;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int g = lane >> 4, e = lane & 15;
  const int el = 16 * wv + e;
  const int ntiles = (K + TILE - 1) / TILE;
  int t0, t1;
  pipe_tiles(blockIdx.x, gridDim.x, ntiles, t0, t1);

  // fragments into LDS for the workgroup's lifetime
  {
    const T* fr = frag + (size_t)(blockIdx.x % FRAG_COPIES) * Q::nfrag * 64;
    for (int i = tid; i < Q::nfrag * 64; i += 256) s_fr[i] = fr[i];
  }

  // node held by accumulator register r of this lane, recomputed where it
  // is used rather than held across the loop
  auto node_of = [&](int r) { return RowMap<T, N>::node_here(0, g, r); };
  PI P;
  auto prefetch = [&](int tile, int hk) {
    const int k0 = tile * TILE;
    const int ne = min(TILE, K - k0);
    const bool valid = el < ne;
    // the tile's nodal rows from a uniform base with 32-bit lane offsets
    const T* ut = u + (size_t)k0 * Np;
    const T* uh = u + (size_t)(hk >= 0 ? hk : 0) * Np;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = node_of(r);
      const unsigned o = (unsigned)(el * Np + n);
      P.U[r] = (valid && n < Np) ? ut[o] : T(0);
      P.UH[r] = (hk >= 0 && n < Np) ? uh[(unsigned)n] : T(0);
      if constexpr (RESID) P.B[r] = (valid && n < Np) ? bvec[(size_t)k0 * Np + o] : T(0);
    }
#pragma unroll
    for (int j = 0; j < PI::RPT; ++j) {
      const int i = tid + 256 * j, w = i / TILE, ee = i % TILE;
      P.rc[j] = ee < ne ? rec[(size_t)w * K + k0 + ee] : T(0);
    }
#pragma unroll
    for (int j = 0; j < PI::HPT; ++j) {
      const int i = tid + 256 * j;
      P.hb[j] = i < 6 * HC ? hab[(size_t)tile * 6 * HC + i] : T(0);
    }
    if (tid < 3 * TILE) {
      const int f = tid / TILE, ee = tid % TILE;
      P.fc = ee < ne ? fcode[(size_t)f * K + k0 + ee] : (FT_NONE << 11);
    }
    // the halo list of the following tile, loaded unconditionally from a
    // valid (clamped) address and masked where it is used: consuming it
    // here would make this prefetch wait on itself
    const int tn = tile + 1 < ntiles ? tile + 1 : tile;
    P.hkn = hpos[(size_t)tn * HC + (el < HC ? el : HC - 1)];
  };

  int hk_cur = (t0 < t1 && el < HC) ? hpos[(size_t)t0 * HC + el] : -1;
  if (t0 < t1) prefetch(t0, hk_cur);
  double dacc = 0.0;

  for (int tile = t0; tile < t1; ++tile) {
    const int k0 = tile * TILE;
    const int ne = min(TILE, K - k0);
    const bool valid = el < ne;
    lds_barrier();  // the previous tile's LDS reads are done (and fragments visible)
    // commit the prefetched tile
    T U[NB][4], UH[NB][4], Bv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      U[0][r] = P.U[r];
      UH[0][r] = P.UH[r];
      if constexpr (RESID)
        Bv[r] = P.B[r];
      else
        Bv[r] = T(0);
    }
#pragma unroll
    for (int j = 0; j < PI::RPT; ++j) {
      const int i = tid + 256 * j, w = i / TILE, ee = i % TILE;
      s_rec[ee * RECS + w] = P.rc[j];
    }
#pragma unroll
    for (int j = 0; j < PI::HPT; ++j) {
      const int i = tid + 256 * j;
      if (i < 6 * HC) s_hab[i] = P.hb[j];
    }
    if (tid < 3 * TILE) s_fc[tid] = P.fc;
    const bool hwave = __ballot(hk_cur >= 0) != 0;  // uniform per wave
    // the next tile's loads go out now; its halo list came with this tile's
    if (tile + 1 < t1) {
      const int hkn = el < HC ? P.hkn : -1;
      prefetch(tile + 1, hkn);
      hk_cur = hkn;
    }
    lds_barrier();

    // ---- GEMM 1: own elements and halo into LDS --------------------------
    const T* re = s_rec + el * RECS;
    T A1[2 * NB][KB];
#pragma unroll
    for (int b = 0; b < 2 * NB; ++b)
#pragma unroll
      for (int t = 0; t < KB; ++t) A1[b][t] = s_fr[(Q::f1 + b * KB + t) * 64 + lane];
    T UR[NB][4], US[NB][4];
    {
      const FaceAB<T> ab{re[R_A], re[R_A + 1], re[R_A + 2], re[R_B], re[R_B + 1], re[R_B + 2]};
      grad_and_traces<T, N>(A1, U, UR, US, ab, s_x + el * SW, g);
      if (hwave) {
        const FaceAB<T> abh{s_hab[el], s_hab[HC + el], s_hab[2 * HC + el],
                            s_hab[3 * HC + el], s_hab[4 * HC + el], s_hab[5 * HC + el]};
        T HR[NB][4], HS[NB][4];
        grad_and_traces<T, N>(A1, UH, HR, HS, abh, s_x + (TILE + el) * SW, g);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const T a = UR[0][r], c = US[0][r];
      UR[0][r] = re[R_VRR] * a + re[R_VRS] * c;
      US[0][r] = re[R_VRS] * a + re[R_VSS] * c;
    }
    lds_barrier();

    // ---- face phase: P, aX, bX in registers -------------------------------
    T Fv[12];
    face_rows_inreg<T, N, UDIR>(Fv, s_x, re, s_fc, el, g, udir, k0 + el);

    // ---- GEMM 2: volume + face lift (+ mass) -----------------------------
    T Jm = T(0);
    if constexpr (HELM) Jm = alpha * re[R_J];
    // four independent accumulator chains (volume r, volume s, the face
    // rows split in two) so consecutive matrix-core products rarely wait
    v4 acc = {0, 0, 0, 0}, acs = {0, 0, 0, 0}, acf = {0, 0, 0, 0}, acg = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < KB; ++t) {
      acc = MF::mma(s_fr[(Q::f2 + t) * 64 + lane], UR[0][t % 4], acc);
      acs = MF::mma(s_fr[(Q::f2 + KB + t) * 64 + lane], US[0][t % 4], acs);
    }
#pragma unroll
    for (int t = 0; t < FK; ++t) {
      if (t % 2 == 0)
        acf = MF::mma(s_fr[(Q::fF + t) * 64 + lane], Fv[t], acf);
      else
        acg = MF::mma(s_fr[(Q::fF + t) * 64 + lane], Fv[t], acg);
    }
    if constexpr (HELM) {
#pragma unroll
      for (int t = 0; t < KB; ++t)
        acc = MF::mma(s_fr[(Q::fM + t) * 64 + lane], Jm * U[0][t % 4], acc);
    }
    T* yt = y + (size_t)k0 * Np;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = node_of(r);
      T v = acc[r] + acs[r] + (acf[r] + acg[r]);
      if constexpr (HELM) v += shift * U[0][r];
      if (valid && n < Np) {
        const unsigned gi = (unsigned)(el * Np + n);
        if constexpr (RESID)
          yt[gi] = Bv[r] - v;
        else
          yt[gi] = v;
        if constexpr (DOT) dacc += (double)U[0][r] * (double)v;
      }
    }
  }

  if constexpr (DOT) {
    for (int off = 32; off > 0; off >>= 1) dacc += __shfl_down(dacc, off, 64);
    __syncthreads();
    if (lane == 0) s_red[wv] = dacc;
    __syncthreads();
    if (tid == 0) partial[blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
  }
}

}  // namespace ddgx
