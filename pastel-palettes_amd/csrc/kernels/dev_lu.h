// Workgroup-level LU, triangular solves and low-rank accumulation.
//
// Semantics follow the reference device routines (src/kernel.cu):
//   getrf   in-place LU without pivoting, unit-lower L, upper U
//           (DenseGetrf / blockDenseGetrf, src/kernel.cu:104-147, 820-882)
//   trsml   B <- L^{-1} B, L unit lower (src/kernel.cu:150-182, 667-715);
//           b_T: B stored transposed (the reference leaves this branch empty,
//           src/kernel.cu:1423-1424)
//   trsmr   B <- B U^{-1}, U upper non-unit (src/kernel.cu:185-226, 718-766);
//           b_T: B stored transposed (src/kernel.cu:229-285, 769-817)
//   trsmu   B <- U^{-1} B (left, upper non-unit; backward substitution)
//   accum   U1 V1^T <- Q Q^T (U1 V1^T + U2 V2^T), Q an orthonormal basis of
//           (U1 V1^T + U2 V2^T) Omega, Omega = the nx x k1 sketch
//           (blockLowRankAccum, src/kernel.cu:1255-1280).  The reference
//           orthonormalises with Givens rotations; Householder is used here.
//           Q Q^T depends only on range(Y), so the product U1 V1^T is the same.
//
// Organisation (MI355X): everything is blocked by NB = 32.
//   * getrf: right-looking.  In the 32-column panel one wave factors the
//     diagonal block (pivot rows broadcast from registers), while a second
//     wave forms L11^{-1} and two more eliminate the rows below, following
//     the first through LDS counters (no workgroup barrier per column).
//     U12 = L11^{-1} A12 and the trailing update A22 -= L21 U12 run on MFMA
//     through gemm_block.
//   * trsm*: the inverses of the 32 x 32 diagonal blocks of the triangular
//     factor are formed first (one wave per block, column-wise substitution),
//     then the solve is a sequence of block products X_i = inv_i B_i and
//     block updates, all on MFMA.
// The diagonal-block inverses live in the worker's scratch (`tmp`, at least
// ceil(n / 32) * 32 * 32 elements).  All routines end with a barrier.
#pragma once

#include "dev_gemm.h"

namespace pp {

constexpr int NB = 32;  // panel width and diagonal block size

// Column c of the inverse of the nb x nb triangular block held in LDS (pitch
// NB + 1): unit lower (UPPER = false) or non-unit upper.  Rows >= nb are 0.
template <typename T, bool UPPER>
__device__ __forceinline__ void tri_inverse_column(const PP_LDS T* s, int nb, int c, T y[NB]) {
  // The empty asm statements (which consume each row's result and clobber
  // memory) keep the uniform LDS reads of one row next to their use; hoisted
  // across rows they would need ~NB^2/2 live registers.
  if (!UPPER) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      T acc = (i == c) ? T(1) : T(0);
#pragma unroll
      for (int j = 0; j < i; ++j) acc -= s[i * (NB + 1) + j] * y[j];
      asm volatile("" : "+v"(acc)::"memory");
      y[i] = (i < nb) ? acc : T(0);
    }
  } else {
    // column NB of each row holds 1 / u_ii
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      T acc = (i == c) ? T(1) : T(0);
#pragma unroll
      for (int j = i + 1; j < NB; ++j) acc -= s[i * (NB + 1) + j] * y[j];
      asm volatile("" : "+v"(acc)::"memory");
      y[i] = (i < nb) ? acc * s[i * (NB + 1) + NB] : T(0);
    }
  }
}

// Inverses of the NB x NB diagonal blocks of the n x n triangular matrix F
// (pitch ldf); block b is written row-major (pitch NB) at out + b NB NB.
// Two blocks per wave (lanes 0-31 and 32-63, one inverse column per lane);
// uses NB (NB + 1) elements of LDS per block, the last column holding the
// reciprocals of the diagonal (upper case).
template <typename T, bool UPPER>
__device__ __forceinline__ void diag_inverses(const T* F, int ldf, int n, T* out) {
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id();
  const int half = l >> 5, lc = l & 31;
  const int nblk = (n + NB - 1) / NB;
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const bool fvec = (n % NB) == 0 && (ldf % V) == 0 && (reinterpret_cast<uintptr_t>(F) & 15) == 0;
  PP_LDS T* s = (PP_LDS T*)lds_base() + (2 * w + half) * NB * (NB + 1);
  for (int b0 = 0; b0 < nblk; b0 += 2 * nw) {
    const int b = b0 + 2 * w + half;
    const int r0 = b * NB;
    const int nb = b < nblk ? ((n - r0) < NB ? (n - r0) : NB) : 0;
    if (b < nblk && fvec) {
      // lane lc loads row lc of the block as whole 16-byte vectors, all at
      // once; the reciprocal of the diagonal is taken from LDS
      T v[NB];
#pragma unroll
      for (int q = 0; q < NB / V; ++q) {
        const vec_t x = gload<vec_t>(F + (int64_t)(r0 + lc) * ldf + r0 + q * V);
#pragma unroll
        for (int e = 0; e < V; ++e) v[q * V + e] = x[e];
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) s[lc * (NB + 1) + j] = v[j];
      if (UPPER) s[lc * (NB + 1) + NB] = T(1) / (T)s[lc * (NB + 1) + lc];
    } else if (b < nblk) {
      // eight loads in flight per lane, through the global address space
      // with clamped addresses (see narrow_solve's staging)
      for (int e0 = lc; e0 < NB * NB; e0 += 8 * 32) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + u * 32, i = e / NB, j = e - i * NB;
          const bool in = i < nb && j < nb;
          v[u] = gload<T>(F + (in ? (int64_t)(r0 + i) * ldf + r0 + j : (int64_t)r0 * ldf + r0));
          if (!in) v[u] = T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + u * 32, i = e / NB, j = e - i * NB;
          s[i * (NB + 1) + j] = v[u];
        }
      }
      if (UPPER && lc < nb) s[lc * (NB + 1) + NB] = T(1) / F[(int64_t)(r0 + lc) * ldf + r0 + lc];
    }
    __syncthreads();
    if (b < nblk && lc < nb) {
      T y[NB];
      tri_inverse_column<T, UPPER>(s, nb, lc, y);
      T* o = out + (int64_t)b * NB * NB;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < nb) o[i * NB + lc] = y[i];
    }
    __syncthreads();
  }
}

// diag_inverses with the inverses left in LDS: block b at D + b NB (NB + 1)
// (row-major, pitch NB + 1, in place of its loaded copy; column NB holds the
// reciprocals of the diagonal in the upper case).  n a multiple of NB, at
// most 2 NB blocks per wave (one pass); ends with a barrier.  In two halves:
// diag_blocks_load (the blocks into LDS, no barrier) and diag_blocks_invert.
template <typename T, bool UPPER>
__device__ __forceinline__ void diag_blocks_load(const T* F, int ldf, int n, PP_LDS T* D) {
  const int w = wave_id(), l = lane_id();
  const int half = l >> 5, lc = l & 31;
  const int nblk = n / NB;
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const bool fvec = (ldf % V) == 0 && (reinterpret_cast<uintptr_t>(F) & 15) == 0;
  const int b = 2 * w + half, r0 = b * NB;
  PP_LDS T* s = D + b * NB * (NB + 1);
  if (b < nblk) {
    // lane lc loads row lc of the block (whole vectors when aligned)
    T v[NB];
    if (fvec) {
#pragma unroll
      for (int q = 0; q < NB / V; ++q) {
        const vec_t x = gload<vec_t>(F + (int64_t)(r0 + lc) * ldf + r0 + q * V);
#pragma unroll
        for (int e = 0; e < V; ++e) v[q * V + e] = x[e];
      }
    } else {
#pragma unroll
      for (int j = 0; j < NB; ++j) v[j] = gload<T>(F + (int64_t)(r0 + lc) * ldf + r0 + j);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) s[lc * (NB + 1) + j] = v[j];
    if (UPPER) s[lc * (NB + 1) + NB] = T(1) / (T)s[lc * (NB + 1) + lc];
  }
}
template <typename T, bool UPPER>
__device__ __forceinline__ void diag_blocks_invert(int n, PP_LDS T* D) {
  const int w = wave_id(), l = lane_id();
  const int half = l >> 5, lc = l & 31;
  const int nblk = n / NB;
  const int b = 2 * w + half;
  PP_LDS T* s = D + b * NB * (NB + 1);
  __syncthreads();
  if (b < nblk) {
    T y[NB];
    tri_inverse_column<T, UPPER>(s, NB, lc, y);
    // the half-wave has read the whole block (in order before these writes)
#pragma unroll
    for (int i = 0; i < NB; ++i) s[i * (NB + 1) + lc] = y[i];
  }
  __syncthreads();
}
template <typename T, bool UPPER>
__device__ __forceinline__ void diag_inverses_lds(const T* F, int ldf, int n, PP_LDS T* D) {
  diag_blocks_load<T, UPPER>(F, ldf, n, D);
  diag_blocks_invert<T, UPPER>(n, D);
}

// ---------------------------------------------------------------- panel update
// Pitch of the LDS operand tiles of lu_update_lds: rows of NB elements padded
// so that the 16-byte fragment reads of eight consecutive lanes fall in
// distinct banks.
template <typename T>
constexpr int lu_lds_pitch() { return NB + (sizeof(T) == 8 ? 2 : 4); }

// Whether lu_update_lds holds L21 and U12 of an rm x rn trailing block in LDS.
template <typename T>
__device__ __forceinline__ bool lu_update_fits(int rm, int rn) {
  const int rows = ((rm + 31) & ~31) + ((rn + 31) & ~31);
  return (int64_t)rows * lu_lds_pitch<T>() * sizeof(T) <= kLdsBytes - 1024;
}

// U12 = L11^{-1} A12 and A22 -= L21 U12 with both operands of the trailing
// product in LDS (rm, rn > 0; lu_update_fits):
//   1. the workgroup copies L21 (rows padded to whole tiles, columns past pw
//      and rows past rm zero) into LDS with coalesced loads, while each wave
//      forms U12 tiles (32 columns each; stored to A12 and, transposed, to
//      LDS with rows past pw zero);
//   2. after one barrier the rm x rn tiles of A22 are dealt round-robin to
//      the waves (better balanced than whole columns); a wave keeps the C
//      values of two tiles in flight and reads both fragments from LDS.
// Same products and summation order per element as lu_update's.  At 256 x
// 256 the trailing products run at the FP64 MFMA's issue rate (about 147
// cycles per 16x16x4 instruction measured on one CU, so about 2 us per 32 x
// 32 x 32 tile).
template <typename T, int MODE>
__device__ __forceinline__ void lu_update_lds(T* M, int ld, int p, int pw, int nx, int ny, const T* linv) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW, PA = lu_lds_pitch<T>();
  static_assert(EPL * sizeof(T) % 16 == 0 && NB == 32, "fragment reads are whole 16-byte vectors");
  const int rn = nx - p - pw, rm = ny - p - pw;
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id(), g = l >> 4, li = l & 15;
  const int ctn = (rn + 31) >> 5, rtn = (rm + 31) >> 5;
  T* A12 = M + (int64_t)p * ld + p + pw;
  const T* L21 = M + (int64_t)(p + pw) * ld + p;
  T* A22 = M + (int64_t)(p + pw) * ld + p + pw;
  PP_LDS T* La = (PP_LDS T*)lds_base();  // [rtn * 32][PA]: L21
  PP_LDS T* Ub = La + rtn * 32 * PA;      // [ctn * 32][PA]: U12 transposed
  const int nt = rtn * ctn;               // trailing tiles t = ct * rtn + rt
  auto load_c = [&](int t, T (&cv)[2][2][4]) {
    const int r0 = (t % rtn) * 32, c0 = (t / rtn) * 32;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + a * 16 + P::orow(l, r), col = c0 + j * 16 + li;
          cv[a][j][r] = gload<T>(A22 + (int64_t)(row < rm ? row : rm - 1) * ld + (col < rn ? col : rn - 1));
        }
  };
  // 1a. L21 -> LDS: 16-byte loads, NB / V threads per row, all in flight
  {
    constexpr int V = 16 / sizeof(T), TPR = NB / V;
    typedef T vec_t __attribute__((ext_vector_type(V)));
    const bool vec = pw == NB && (ld % V) == 0 && ((reinterpret_cast<uintptr_t>(M) & 15) == 0) && (p % V) == 0;
    const int k0 = (threadIdx.x % TPR) * V, rstep = blockDim.x / TPR;
#pragma unroll 16
    for (int r = threadIdx.x / TPR; r < rtn * 32; r += rstep) {
      vec_t x;
      if (vec) {
        x = gload<vec_t>(L21 + (int64_t)(r < rm ? r : rm - 1) * ld + k0);
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
          x[e] = gload<T>(L21 + (int64_t)(r < rm ? r : rm - 1) * ld + (k0 + e < pw ? k0 + e : pw - 1));
      }
#pragma unroll
      for (int e = 0; e < V; ++e)
        if (r >= rm || k0 + e >= pw) x[e] = T(0);
      *(PP_LDS vec_t*)(La + r * PA + k0) = x;
    }
  }
  // the C values of the wave's first tile are requested before the U12
  // products (phase 1 does not touch A22); past the last tile the last one
  // is requested again
  T cv0[2][2][4], cv1[2][2][4];
  const int tl = nt - 1;
  load_c(w < nt ? w : tl, cv0);
  // 1b. U12 tiles
  for (int ct = w; ct < ctn; ct += nw) {
    const int c0 = ct * 32;
    OperandLane<T, EPL> la[2], lb[2];
    T fa[CH][2][EPL], fb[CH][2][EPL];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      la[h].init(linv, NB, true, 16 * h + li, pw);
      lb[h].init(A12, ld, false, c0 + 16 * h + li, rn);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        la[h].template load<false>(fa[c][h], c * KW + EPL * g, pw);
        lb[h].template load<false>(fb[c][h], c * KW + EPL * g, pw);
#pragma unroll
        for (int e = 0; e < EPL; ++e)  // k >= pw: the clamped loads cancel through A
          if (c * KW + EPL * g + e >= pw) fa[c][h][e] = T(0);
      }
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[a][j] = P::mma_chunk(fa[c][a], fb[c][j], acc[a][j]);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = a * 16 + P::orow(l, r), col = j * 16 + li;
          const T v = acc[a][j][r];
          Ub[(c0 + col) * PA + row] = row < pw ? v : T(0);
          if (row < pw && c0 + col < rn) gstore<T>(A12 + (int64_t)row * ld + c0 + col, v);
        }
  }
  __syncthreads();
  load_c(w + nw < nt ? w + nw : tl, cv1);
  // 2. trailing tiles
  auto apply = [&](int t, const T (&cv)[2][2][4]) {
    const int r0 = (t % rtn) * 32, c0 = (t / rtn) * 32;
    typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int V = 16 / sizeof(T);
    T fa[CH][2][EPL], fb[CH][2][EPL];
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < EPL / V; ++v) {
          const int k = c * KW + EPL * g + v * V;
          const vec_t x = *(const PP_LDS vec_t*)(La + (r0 + 16 * h + li) * PA + k);
          const vec_t y = *(const PP_LDS vec_t*)(Ub + (c0 + 16 * h + li) * PA + k);
#pragma unroll
          for (int e = 0; e < V; ++e) {
            fa[c][h][v * V + e] = x[e];
            fb[c][h][v * V + e] = y[e];
          }
        }
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[a][j] = P::mma_chunk(fa[c][a], fb[c][j], acc[a][j]);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + a * 16 + P::orow(l, r), col = c0 + j * 16 + li;
          if (row < rm && col < rn) gstore<T>(A22 + (int64_t)row * ld + col, cv[a][j][r] - acc[a][j][r]);
        }
  };
  // two tiles in flight: a buffer is refilled right after its tile's stores
  // are issued, so the wait for the next tile's values never includes the
  // stores before it (the memory counter retires in issue order)
  for (int t = w; t < nt; t += 2 * nw) {
    apply(t, cv0);
    load_c(t + 2 * nw < nt ? t + 2 * nw : tl, cv0);
    if (t + nw < nt) {
      apply(t + nw, cv1);
      load_c(t + 3 * nw < nt ? t + 3 * nw : tl, cv1);
    }
  }
  __syncthreads();
}

// After the panel at p (width pw, multipliers L21 below, L11^{-1} in linv):
// U12 = L11^{-1} A12 and A22 -= L21 U12 in one step.  Each wave owns column
// tiles of U12: it forms the 32-column tile (stored to A12), keeps it as the
// B operand in registers (through a wave-private LDS tile) and applies it to
// every row tile of A22 in that column, so the two products need no barrier
// or global round trip between them.  A22's L21 fragments and C tile are
// requested one row tile ahead.
template <typename T, int MODE>
__device__ __forceinline__ void lu_update(T* M, int ld, int p, int pw, int nx, int ny, const T* linv) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW, V = 16 / sizeof(T), UT = NB + 2;
  const int rn = nx - p - pw, rm = ny - p - pw;
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id(), g = l >> 4, li = l & 15;
  const int ctn = (rn + 31) >> 5, rtn = (rm + 31) >> 5;
  T* A12 = M + (int64_t)p * ld + p + pw;
  const T* L21 = M + (int64_t)(p + pw) * ld + p;
  T* A22 = M + (int64_t)(p + pw) * ld + p + pw;
  // FP64 only: in FP32 the register-resident version below is as fast and
  // the LDS version's registers make the factor kernel spill
  if constexpr (sizeof(T) == 8) {
    if (rm > 0 && rn > 0 && lu_update_fits<T>(rm, rn)) {
      lu_update_lds<T, MODE>(M, ld, p, pw, nx, ny, linv);
      return;
    }
  }
  PP_LDS T* ut = (PP_LDS T*)lds_base() + w * NB * UT;
  const bool lvec = pw == NB && (ld % V) == 0 && ((reinterpret_cast<uintptr_t>(M) & 15) == 0) && (p % V) == 0;
  for (int ct = w; ct < ctn; ct += nw) {
    const int c0 = ct * 32;
    // U12 tile
    {
      OperandLane<T, EPL> la[2], lb[2];
      T fa[CH][2][EPL], fb[CH][2][EPL];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        la[h].init(linv, NB, true, 16 * h + li, pw);
        lb[h].init(A12, ld, false, c0 + 16 * h + li, rn);
      }
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          la[h].template load<false>(fa[c][h], c * KW + EPL * g, pw);
          lb[h].template load<false>(fb[c][h], c * KW + EPL * g, pw);
#pragma unroll
          for (int e = 0; e < EPL; ++e)  // k >= pw: the clamped loads cancel through A
            if (c * KW + EPL * g + e >= pw) fa[c][h][e] = T(0);
        }
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[a][j] = P::mma_chunk(fa[c][a], fb[c][j], acc[a][j]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = a * 16 + P::orow(l, r), col = j * 16 + li;
            const T v = acc[a][j][r];
            ut[row * UT + col] = row < pw ? v : T(0);
            if (row < pw && c0 + col < rn) gstore<T>(A12 + (int64_t)row * ld + c0 + col, v);
          }
    }
    if (rtn == 0) continue;
    // the U12 tile as B fragments (the wave's own LDS writes, in order)
    T ub[CH][2][EPL];
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < EPL; ++e) ub[c][h][e] = ut[(c * KW + EPL * g + e) * UT + 16 * h + li];
    // C values of a tile (clamped addresses; the stores are masked)
    auto load_c = [&](int rt, T (&cv)[2][2][4]) {
      const int r0 = rt * 32;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + a * 16 + P::orow(l, r), col = c0 + j * 16 + li;
            cv[a][j][r] = gload<T>(A22 + (int64_t)(row < rm ? row : rm - 1) * ld + (col < rn ? col : rn - 1));
          }
    };
    // FP32: the next tile's C values travel with its L21 fragments (two
    // buffers); FP64: two C buffers would not fit the register file next to
    // the fragments (the kernel spilled), so a tile's C values are requested
    // when its products start and consumed after them
    constexpr bool kPrefetchC = sizeof(T) == 4;
    auto load_tile = [&](int rt, T (&fa)[CH][2][EPL], T (&cv)[2][2][4]) {
      const int r0 = rt * 32;
      OperandLane<T, EPL> la[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) la[h].init(L21, ld, true, r0 + 16 * h + li, rm);
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (lvec) la[h].template load<true>(fa[c][h], c * KW + EPL * g, pw);
          else la[h].template load<false>(fa[c][h], c * KW + EPL * g, pw);
        }
      if constexpr (kPrefetchC) load_c(rt, cv);
    };
    auto apply_tile = [&](int rt, const T (&fa)[CH][2][EPL], T (&cv)[2][2][4]) {
      const int r0 = rt * 32;
      if constexpr (!kPrefetchC) load_c(rt, cv);
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[a][j] = P::mma_chunk(fa[c][a], ub[c][j], acc[a][j]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + a * 16 + P::orow(l, r), col = c0 + j * 16 + li;
            if (row < rm && col < rn) gstore<T>(A22 + (int64_t)row * ld + col, cv[a][j][r] - acc[a][j][r]);
          }
    };
    // (in FP64 cv0 and cv1 are each live only inside one apply_tile)
    T fa0[CH][2][EPL], fa1[CH][2][EPL], cv0[2][2][4], cv1[2][2][4];
    if constexpr (!kPrefetchC) {
      for (int rt = 0; rt < rtn; ++rt) {
        load_tile(rt, fa0, cv0);
        apply_tile(rt, fa0, cv0);
      }
    } else {
    load_tile(0, fa0, cv0);
    for (int rt = 0; rt < rtn; rt += 2) {
      // past the last tile the last one is requested again, so every
      // iteration issues the same loads
      load_tile(rt + 1 < rtn ? rt + 1 : rt, fa1, cv1);
      apply_tile(rt, fa0, cv0);
      if (rt + 1 < rtn) {
        load_tile(rt + 2 < rtn ? rt + 2 : rt + 1, fa0, cv0);
        apply_tile(rt + 1, fa1, cv1);
      }
    }
    }
  }
  __syncthreads();
}

// LDS ordering between waves (panel LU, narrow solve).  The LDS unit executes the operations of one
// wave in issue order, so (a) lanes of a wave see the wave's earlier LDS
// writes, and (b) a wave that reads a counter written after some data by
// another wave also sees that data.  Only compiler reordering has to be
// prevented, which a signal fence does without waiting on memory counters.
__device__ __forceinline__ void lds_order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
}

// Waits until the LDS counter reaches `target`.  Every such wait ends: the
// wave that writes a counter waits for nothing, or only for counters whose
// writers wait for nothing.
__device__ __forceinline__ void lds_wait_ge(volatile PP_LDS int* c, int target) {
  while (*c < target) __builtin_amdgcn_s_sleep(1);
  lds_order();
}

// ---------------------------------------------------------------- narrow solve
// In-place triangular solve with few right-hand-side columns, the right-hand
// side resident in LDS for the whole solve:
//   forward:  for i = 0, 1, ...:  X_i = D_i W_i;  W_{>i} -= T_{>i,i} X_i
//   backward: for i = last .. 0:  X_i = D_i W_i;  W_{<i} -= T_{<i,i} X_i
// W is m x w (w <= 64), W(r, c) = Wg[r ldw + c] or, when w_trans,
// Wg[c ldw + r].  T(r, q) = F[r ldf + q] or, when f_trans, F[q ldf + r];
// D_i (inverse of T's i-th 32 x 32 diagonal block) is dinv[i NB NB + r NB + q]
// or, when f_trans, its transpose.  One step's tiles are 32 x 32 MFMA tiles,
// T and D fragments loaded from global memory, W fragments from LDS.

// A fragments of a 32-row tile (rows by la, k < kmax, both 16-wide chunks)
template <typename T, int MODE>
__device__ __forceinline__ void load_a_frags(const OperandLane<T, Mma<T, MODE>::EPL> (&la)[2], int kmax,
                                             T (&fa)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL]) {
  using P = Mma<T, MODE>;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW;
  const int g = lane_id() >> 4;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int h = 0; h < 2; ++h) la[h].template load<false>(fa[c][h], c * KW + EPL * g, kmax);
}

// acc += A (fragments fa) x W[k][c0 .. c0 + 31] for k < 32, W rows from LDS
template <typename T, int MODE>
__device__ __forceinline__ void tile_from_lds(const T (&fa)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL],
                                              const PP_LDS T* Wk, int wp, int c0,
                                              typename Mma<T, MODE>::acc_t (&acc)[2][2]) {
  using P = Mma<T, MODE>;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW;
  const int l = lane_id(), g = l >> 4, li = l & 15;
  T fb[CH][2][EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int kg = c * KW + EPL * g;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < EPL; ++e) fb[c][h][e] = Wk[(kg + e) * wp + c0 + 16 * h + li];
  }
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = P::mma_chunk(fa[c][i], fb[c][j], acc[i][j]);
}

// Chain form for FP32 solves of two to eight whole blocks with one 32-wide
// column tile.  Wave 0 forms the blocks in solve order,
//   X_b = D_b (W_b - S_b - T_{b,p} X_p)    (p: the block solved just before b),
// keeping X_p in registers: in FP32 MFMA the accumulator layout of a 32-row
// tile (lane l, register r: row 4 (l / 16) + r) is the B-operand layout of the
// next product (k = 4 (l / 16) + e), so the two products of a block follow
// each other with no LDS round trip.  Waves 1-3 form S_b, the sum of
// T_{b,j} X_j over the blocks j solved before p (each X_j read from LDS after
// wave 0 publishes it), and hand it to wave 0 through LDS (Ss, flag sready[b])
// while wave 0 works on p.  Wave 0 waits only for S_b, whose producer waits
// only for blocks wave 0 has already published, so every wait ends.  Not
// inlined (own register allocation).
template <typename T, int MODE>
__device__ __attribute__((noinline)) void narrow_chain(PP_LDS T* Ws, int wp, PP_LDS T* Ss, const PP_LDS T* Ts,
                                                       volatile PP_LDS int* ready, volatile PP_LDS int* sready, int m,
                                                       int nblk, const T* F, int ldf, bool f_trans,
                                                       const PP_LDS T* Dl, bool forward) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW;
  static_assert(CH == 2 && EPL == 4 && KW == 16, "chain form: 16x16x4 MFMA, 32-row blocks");
  const int l = lane_id(), wv = wave_id(), li = l & 15, g = l >> 4;
  auto blk = [&](int o) { return forward ? o : nblk - 1 - o; };
  // A fragments of T_{b,j} (rows of block b, k over the rows of block j)
  auto load_t = [&](int b, int j, T(&fa)[CH][2][EPL]) {
    OperandLane<T, EPL> la[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (f_trans) la[h].init(F + (int64_t)j * NB * ldf, ldf, false, b * NB + 16 * h + li, m);
      else la[h].init(F + j * NB, ldf, true, b * NB + 16 * h + li, m);
    }
    load_a_frags<T, MODE>(la, NB, fa);
  };
  // A fragments of D_b (inverses in LDS, element (r, q) at r (NB + 1) + q)
  auto load_d = [&](int b, T(&fd)[CH][2][EPL]) {
    const PP_LDS T* Db = Dl + b * NB * (NB + 1);
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
          const int r = 16 * h + li, q = c * KW + EPL * g + e;
          fd[c][h][e] = f_trans ? (T)Db[q * (NB + 1) + r] : (T)Db[r * (NB + 1) + q];
        }
  };
  // A fragments of the chain block T_{blk(o), blk(o - 1)} staged in LDS
  auto load_ts = [&](int o, T(&fa)[CH][2][EPL]) {
    const PP_LDS T* tb = Ts + (o - 1) * NB * (NB + 1);
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < EPL; ++e) fa[c][h][e] = tb[(16 * h + li) * (NB + 1) + c * KW + EPL * g + e];
  };
  // out[h][jj] = A (fragments fa) x B, B given in accumulator layout:
  // chunk c of B is the 16-row tile c (b[c][jj][e] = row 16 c + 4 g + e)
  auto product = [&](const T(&fa)[CH][2][EPL], const acc_t(&b)[2][2], acc_t(&out)[2][2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        out[h][jj] = P::zero();
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          T fb[EPL];
#pragma unroll
          for (int e = 0; e < EPL; ++e) fb[e] = b[c][jj][e];
          out[h][jj] = P::mma_chunk(fa[c][h], fb, out[h][jj]);
        }
      }
  };
  if (wv == 0) {
    acc_t xp[2][2];
    T fdn[CH][2][EPL], ftn[CH][2][EPL];
    load_d(blk(0), fdn);
    load_ts(1, ftn);
    for (int ob = 0; ob < nblk; ++ob) {
      const int b = blk(ob), rb = b * NB;
      // fd = D_b, ft = T_{b,p} (requested one block ahead)
      T fd[CH][2][EPL], ft[CH][2][EPL];
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int e = 0; e < EPL; ++e) {
            fd[c][h][e] = fdn[c][h][e];
            ft[c][h][e] = ftn[c][h][e];
          }
      if (ob + 1 < nblk) {
        load_d(blk(ob + 1), fdn);
        if (ob >= 1) load_ts(ob + 1, ftn);
      }
      acc_t w[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
          for (int r = 0; r < 4; ++r) w[a][jj][r] = Ws[(rb + 16 * a + P::orow(l, r)) * wp + 16 * jj + li];
      if (ob >= 1) {
        acc_t pr[2][2];
        product(ft, xp, pr);
        if (ob >= 2) {
          lds_wait_ge(sready + b, 1);
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
              for (int r = 0; r < 4; ++r) w[a][jj][r] -= Ss[((b * 4 + a * 2 + jj) * 64 + l) * 4 + r];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) w[a][jj][r] -= pr[a][jj][r];
      }
      product(fd, w, xp);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
          for (int r = 0; r < 4; ++r) Ws[(rb + 16 * a + P::orow(l, r)) * wp + 16 * jj + li] = xp[a][jj][r];
      lds_order();
      if (l == 0) ready[b] = 1;
    }
  } else {
    for (int ob = wv + 1; ob < nblk; ob += 3) {
      const int b = blk(ob);
      acc_t s[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) s[a][0] = s[a][1] = P::zero();
      T ft[CH][2][EPL];
      load_t(b, blk(0), ft);
      for (int oj = 0; oj + 1 < ob; ++oj) {
        const int j = blk(oj);
        T ftn[CH][2][EPL];
        if (oj + 2 < ob) load_t(b, blk(oj + 1), ftn);
        lds_wait_ge(ready + j, 1);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              T fb[EPL];
#pragma unroll
              for (int e = 0; e < EPL; ++e) fb[e] = Ws[(j * NB + c * KW + EPL * g + e) * wp + 16 * jj + li];
              s[h][jj] = P::mma_chunk(ft[c][h], fb, s[h][jj]);
            }
        if (oj + 2 < ob) {
#pragma unroll
          for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
              for (int e = 0; e < EPL; ++e) ft[c][h][e] = ftn[c][h][e];
        }
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
          for (int r = 0; r < 4; ++r) Ss[((b * 4 + a * 2 + jj) * 64 + l) * 4 + r] = s[a][jj][r];
      lds_order();
      if (l == 0) sready[b] = 1;
    }
  }
  __syncthreads();
}

// LDS bytes of narrow_solve for m rows and w columns.
template <typename T>
__device__ __forceinline__ int narrow_solve_bytes(int m, int w) {
  const int ntc = (w + 31) >> 5, wp = ntc * 32 + 16 / (int)sizeof(T), nblk = (m + NB - 1) / NB;
  return nblk * NB * wp * (int)sizeof(T) + nblk * (int)sizeof(int);  // W, then one ready flag per block
}

// LDS offset (bytes) of the diagonal inverses a narrow solve keeps in LDS
// (after W and its ready flags), and whether it does: m a multiple of NB,
// at most eight blocks, everything in LDS.
template <typename T>
__device__ __forceinline__ int narrow_dl_offset(int m, int w) {
  return (narrow_solve_bytes<T>(m, w) + 15) & ~15;
}
template <typename T>
__device__ __forceinline__ bool narrow_dl_ok(int m, int w) {
  // FP32 only: in FP64 the fused form measured slower (larger register
  // frames, the FP64 factorization 0.9 ms longer)
  return sizeof(T) == 4 && w <= 64 && m % NB == 0 && m / NB <= 8 &&
         narrow_dl_offset<T>(m, w) + (m / NB) * NB * (NB + 1) * (int)sizeof(T) <= kLdsBytes - 1024;
}

// dl = 0: the inverses of T's diagonal blocks come from dinv (global memory);
// dl = 1 (unit lower T) or 2 (upper T): they are formed here into LDS
// (diag_inverses_lds at narrow_dl_offset, which narrow_dl_ok must admit)
// before W is staged, and read from there.
template <typename T, int MODE>
__device__ __forceinline__ void narrow_solve(T* Wg, int ldw, bool w_trans, int m, int w, const T* F, int ldf,
                                             bool f_trans, const T* dinv, bool forward, int dl = 0) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  const int ntc = (w + 31) >> 5, wc = ntc * 32, wp = wc + 16 / (int)sizeof(T);
  const int nblk = (m + NB - 1) / NB, mr = nblk * NB;
  PP_LDS T* Ws = (PP_LDS T*)lds_base();
  const int tid = threadIdx.x, nt = blockDim.x, wv = wave_id(), nw = nt >> 6, l = lane_id(), li = l & 15;
  // (FP32 only: narrow_dl_ok admits no FP64 solve, and the FP64 kernels are
  // compiled without this path)
  PP_LDS T* Dl = nullptr;
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const bool wvec = !w_trans && (w % V) == 0 && (ldw % V) == 0 && (reinterpret_cast<uintptr_t>(Wg) & 15) == 0;
  // the chain form (FP32 MFMA, two to eight whole blocks, one column tile,
  // diagonal inverses formed here, W and F 16-byte rows): after W and the
  // inverses, S (nblk blocks, C layout) and its flags, then the blocks
  // T_{b,p} of the chain (pitch NB + 1)
  bool chain = false;
  int ss_off = 0, ts_off = 0;
  if constexpr (sizeof(T) == 4 && MODE == MODE_FP32) {
    ss_off = narrow_dl_offset<T>(m, w) + nblk * NB * (NB + 1) * (int)sizeof(T);
    ts_off = (ss_off + nblk * NB * NB * (int)sizeof(T) + nblk * (int)sizeof(int) + 15) & ~15;
    chain = dl != 0 && wvec && ntc == 1 && m % NB == 0 && nblk >= 2 && nblk <= 8 && nw == 4 &&
            (ldf % V) == 0 && (reinterpret_cast<uintptr_t>(F) & 15) == 0 &&
            ts_off + (nblk - 1) * NB * (NB + 1) * (int)sizeof(T) <= kLdsBytes - 1024;
  }
  if constexpr (sizeof(T) == 4) {
    if (dl != 0) Dl = (PP_LDS T*)(lds_base() + narrow_dl_offset<T>(m, w));
    if (chain) {
      // W, the chain blocks and the diagonal blocks requested together (one
      // memory round trip), then stored, then the inverses formed
      const int vpr = wc / V;  // 16-byte vectors per W row (all of them real: wvec)
      vec_t wq[8], tq[7];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = tid + u * nt, r = e / vpr, c = (e - r * vpr) * V;
        const bool in = e < mr * vpr && r < m && c < w;
        wq[u] = gload<vec_t>(Wg + (in ? (int64_t)r * ldw + c : 0));
        if (!in) wq[u] = vec_t{};
      }
#pragma unroll
      for (int o = 1; o < 8; ++o) {
        if (o < nblk) {
          const int b = forward ? o : nblk - 1 - o, p = forward ? o - 1 : nblk - o;
          const int i = tid >> 3, q = (tid & 7) * V;  // row (or, transposed, column) and first of V
          tq[o - 1] = f_trans ? gload<vec_t>(F + (int64_t)(p * NB + i) * ldf + b * NB + q)
                              : gload<vec_t>(F + (int64_t)(b * NB + i) * ldf + p * NB + q);
        }
      }
      if (dl == 1) diag_blocks_load<T, false>(F, ldf, m, Dl);
      else diag_blocks_load<T, true>(F, ldf, m, Dl);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = tid + u * nt, r = e / vpr, c = (e - r * vpr) * V;
        if (e < mr * vpr) *(PP_LDS vec_t*)(Ws + r * wp + c) = wq[u];
      }
      PP_LDS T* Ts = (PP_LDS T*)(lds_base() + ts_off);
#pragma unroll
      for (int o = 1; o < 8; ++o) {
        if (o < nblk) {
          PP_LDS T* tb = Ts + (o - 1) * NB * (NB + 1);
          const int i = tid >> 3, q = (tid & 7) * V;
#pragma unroll
          for (int e = 0; e < V; ++e) {
            if (f_trans) tb[(q + e) * (NB + 1) + i] = tq[o - 1][e];
            else tb[i * (NB + 1) + q + e] = tq[o - 1][e];
          }
        }
      }
      if (dl == 1) diag_blocks_invert<T, false>(m, Dl);
      else diag_blocks_invert<T, true>(m, Dl);
    } else if (dl == 1) {
      diag_inverses_lds<T, false>(F, ldf, m, Dl);
    } else if (dl == 2) {
      diag_inverses_lds<T, true>(F, ldf, m, Dl);
    }
  }
  // stage W (in the order that keeps global accesses coalesced); rows >= m
  // and columns >= w are zero and stay zero.  The loads go through the
  // global address space with clamped addresses (no branch), so the compiler
  // may issue them ahead of the LDS stores of the same batch (a flat load
  // could alias LDS and would wait for each store before it).
  // W row-major with whole 16-byte vectors: every vector of the staged block
  // is requested at once (one memory round trip for up to 16 vectors per
  // thread); rows >= m and columns >= w read a clamped address and are zeroed
  if (chain) {
    // staged above
  } else if (wvec) {
    constexpr int U = 16;
    const int vpr = wc / V, total = mr * vpr;
    for (int e0 = tid; e0 < total; e0 += U * nt) {
      vec_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * nt, r = e / vpr, c = (e - r * vpr) * V;
        const bool in = e < total && r < m && c < w;
        v[u] = gload<vec_t>(Wg + (in ? (int64_t)r * ldw + c : 0));
        if (!in) v[u] = vec_t{};
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * nt, r = e / vpr, c = (e - r * vpr) * V;
        if (e < total) *(PP_LDS vec_t*)(Ws + r * wp + c) = v[u];
      }
    }
  } else {
    // element loads, eight in flight per thread
    constexpr int U = 8;
    const int total = mr * wc;
    for (int e0 = tid; e0 < total; e0 += U * nt) {
      T v[U];
      int at[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * nt;
        int r, c;
        if (w_trans) {
          c = e / mr;
          r = e - c * mr;
        } else {
          r = e / wc;
          c = e - r * wc;
        }
        const bool in = e < total && r < m && c < w;
        const int64_t g = in ? (w_trans ? (int64_t)c * ldw + r : (int64_t)r * ldw + c) : 0;
        v[u] = gload<T>(Wg + g);
        if (!in) v[u] = T(0);
        at[u] = e < total ? r * wp + c : -1;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (at[u] >= 0) Ws[at[u]] = v[u];
    }
  }
  volatile PP_LDS int* ready = (volatile PP_LDS int*)(Ws + mr * wp);
  for (int e = tid; e < nblk; e += nt) ready[e] = 0;
  constexpr int CH = NB / P::KW;
  if constexpr (sizeof(T) == 4 && MODE == MODE_FP32) {
    if (chain) {
      PP_LDS T* Ss = (PP_LDS T*)(lds_base() + ss_off);
      volatile PP_LDS int* sready = (volatile PP_LDS int*)(Ss + nblk * NB * NB);
      for (int e = tid; e < nblk; e += nt) sready[e] = 0;
      __syncthreads();
      narrow_chain<T, MODE>(Ws, wp, Ss, (const PP_LDS T*)(lds_base() + ts_off), ready, sready, m, nblk, F, ldf,
                            f_trans, Dl, forward);
    }
  }
  if (!chain) {
    __syncthreads();
    // Row-block ownership: wave wv owns the blocks at distance wv, wv + nw,
    // ... from the first block of the solve order (increasing for forward,
    // decreasing for backward) and takes them in that order.  For its block
    // b it applies W_b -= T_{b,j} X_j for the blocks j solved before b, in
    // solve order, each as soon as X_j is published (the T fragments are
    // requested before the wait), then forms X_b = D_b W_b in place and
    // publishes it (ready[b]).  Every element sees the same operations in
    // the same order as in the step-synchronous form; the first block waits
    // for nothing and a block only waits for blocks solved before it, so
    // every wait ends.
    for (int ob = wv; ob < nblk; ob += nw) {
      const int b = forward ? ob : nblk - 1 - ob;
      const int rb = b * NB, bsb = (m - rb) < NB ? (m - rb) : NB;
      // D_b's fragments are requested first (they wait for nothing)
      T fd[CH][2][P::EPL];
      if (sizeof(T) == 4 && Dl != nullptr) {
        // element (r, q) of D_b at Db[r (NB + 1) + q]; fragment rows 16 h + li
        const PP_LDS T* Db = Dl + b * NB * (NB + 1);
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < P::EPL; ++e) {
              const int r = 16 * h + li, q = c * P::KW + P::EPL * (l >> 4) + e;
              fd[c][h][e] = f_trans ? (T)Db[q * (NB + 1) + r] : (T)Db[r * (NB + 1) + q];
            }
      } else {
        OperandLane<T, P::EPL> ld_[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) ld_[h].init(dinv + (int64_t)b * NB * NB, NB, !f_trans, 16 * h + li, bsb);
        load_a_frags<T, MODE>(ld_, bsb, fd);
      }
      for (int oj = 0; oj < ob; ++oj) {
        const int j = forward ? oj : nblk - 1 - oj;
        const int r0 = j * NB, bsj = (m - r0) < NB ? (m - r0) : NB;
        OperandLane<T, P::EPL> la[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (f_trans) la[h].init(F + (int64_t)r0 * ldf, ldf, false, rb + 16 * h + li, m);
          else la[h].init(F + r0, ldf, true, rb + 16 * h + li, m);
        }
        T fa[CH][2][P::EPL];
        load_a_frags<T, MODE>(la, bsj, fa);
        lds_wait_ge(ready + j, 1);
        for (int ct = 0; ct < ntc; ++ct) {
          const int c0 = ct * 32;
          acc_t acc[2][2];
#pragma unroll
          for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
          tile_from_lds<T, MODE>(fa, Ws + r0 * wp, wp, c0, acc);
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int row = rb + a * 16 + P::orow(l, r);
                PP_LDS T* wpx = Ws + row * wp + c0 + jj * 16 + li;
                if (row < m) *wpx = *wpx - acc[a][jj][r];
              }
        }
      }
      for (int ct = 0; ct < ntc; ++ct) {
        const int c0 = ct * 32;
        acc_t acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
        tile_from_lds<T, MODE>(fd, Ws + rb * wp, wp, c0, acc);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = a * 16 + P::orow(l, r);
              Ws[(rb + row) * wp + c0 + jj * 16 + li] = row < bsb ? acc[a][jj][r] : T(0);
            }
      }
      lds_order();
      if (l == 0) ready[b] = 1;
    }
    __syncthreads();
  }
  // write W back
  if (wvec) {
    const int vpr = w / V, total = m * vpr;
    for (int e = tid; e < total; e += nt) {
      const int r = e / vpr, c = (e - r * vpr) * V;
      gstore<vec_t>(Wg + (int64_t)r * ldw + c, *(const PP_LDS vec_t*)(Ws + r * wp + c));
    }
    __syncthreads();
    return;
  }
  for (int e = tid; e < m * w; e += nt) {
    if (w_trans) {
      const int c = e / m, r = e - c * m;
      Wg[(int64_t)c * ldw + r] = Ws[r * wp + c];
    } else {
      const int r = e / w, c = e - r * w;
      Wg[(int64_t)r * ldw + c] = Ws[r * wp + c];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------- getrf
// Value v of lane `lane` on every lane (v_readlane into scalar registers,
// which an FMA can take as an operand).
template <typename T>
__device__ __forceinline__ T bcast_lane(T v, int lane) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  }
}

// Stores rows [0, nr) x columns [0, pw) of a wave's LDS tile S (pitch NB + 1)
// to G (pitch ld).  Lanes 0-31 and 32-63 each write one row of consecutive
// elements, so a store instruction covers two contiguous row segments instead
// of 64 elements in 64 different rows (one row per lane, as the elimination
// holds them).  Called by one wave; S was written by the same wave.
template <typename T>
__device__ __forceinline__ void wave_store_rows(T* G, int64_t ld, const PP_LDS T* S, int nr, int pw) {
  lds_order();
  const int l = lane_id(), c = l & 31;
#pragma unroll 4
  for (int i = l >> 5; i < nr; i += 2)
    if (c < pw) G[i * ld + c] = S[i * (NB + 1) + c];
  lds_order();
}

// 1 / x for the panel's pivot chain: the hardware reciprocal refined by
// Newton steps r <- r + r (1 - x r) (two in FP64, one in FP32; each doubles
// the correct bits of the estimate, and the result is within an ulp of
// 1 / x), a shorter dependent sequence than the correctly rounded division.
// The estimate is only refined where it and 1 / x are normal numbers: for
// |x| near the format's maximum 1 / x is subnormal and the estimate may be
// flushed to 0 (which the Newton steps keep), and for infinite x the steps
// turn rcp(inf) = 0 into NaN.  Pivots outside [2^-1000, 2^1000] (FP64) or
// [2^-120, 2^120] (FP32), zero, infinite and NaN pivots therefore take the
// division, so every pivot gives the multipliers 1 / x gives.
template <typename T>
__device__ __forceinline__ T pivot_recip(T x) {
  if constexpr (sizeof(T) == 8) {
    const T ax = __builtin_fabs(x);
    if (!(ax >= 0x1p-1000 && ax <= 0x1p1000)) return T(1) / x;
    T r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, T(1)), r);
    return __builtin_fma(r, __builtin_fma(-x, r, T(1)), r);
  } else {
    const T ax = __builtin_fabsf(x);
    if (!(ax >= 0x1p-120f && ax <= 0x1p120f)) return T(1) / x;
    const T r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, T(1)), r);
  }
}

// Unblocked LU of the panel M[p:ny, p:p+pw] (pw <= NB) in place, run by three
// roles concurrently (no workgroup barrier inside):
//   wave 0     factors the pw x pw diagonal block.  Lane i owns row i in
//              registers; the pivot row of step j is read straight from lane
//              j's registers (v_readlane), so the sequential chain touches no
//              memory.  Multiplier column j goes to LDS after step j
//              (counter `cols`), and final rows go to LDS in groups of GRP (8 in
//              FP64, 4 in FP32)
//              together with their pivot reciprocals (counter `rows`).
//   wave 1     forms L11^{-1} by the same row operations applied to the
//              identity, following the multiplier columns, and writes it
//              row-major (pitch NB) to linv.
//   waves 2-3  eliminate the rows below the block against each group of
//              pivot rows as soon as it is published.
// Multipliers are formed with the reciprocal of the pivot, which the owner of
// the next pivot row computes right after its pivot element is updated.
// Results leave through a per-wave LDS tile (wave_store_rows).
template <typename T>
__device__ __forceinline__ void panel_lu(T* M, int ld, int p, int pw, int ny, T* linv) {
  constexpr int UP = NB + 2;              // pitch of U: 16-byte aligned rows
  constexpr int GRP = sizeof(T) == 4 ? 4 : 8;  // rows per publication (measured best per type)
  constexpr int SP = NB + 1;              // pitch of the output tiles
  PP_LDS T* U = (PP_LDS T*)lds_base();    // [NB][UP]: final rows, column NB = 1 / u_jj
  PP_LDS T* Lc = U + NB * UP;              // [NB][NB]: multiplier column j at Lc[j * NB + i]
  volatile PP_LDS int* cols = (volatile PP_LDS int*)(Lc + NB * NB);
  volatile PP_LDS int* rows = cols + 1;
  // output tiles: waves 0 and 1 32 rows each, waves 2 and 3 128 rows each
  PP_LDS T* tiles = (PP_LDS T*)(lds_base() + 17 * 1024);
  static_assert((NB * UP + NB * NB) * sizeof(T) + 8 <= 17 * 1024, "panel LDS layout");
  static_assert(17 * 1024 + (2 * NB + 256) * SP * sizeof(T) <= kLdsBytes - 16, "panel LDS layout");
  const int tid = threadIdx.x, l = lane_id(), w = wave_id();
  const int R = ny - p;
  if (tid == 0) {
    *cols = 0;
    *rows = 0;
  }
  __syncthreads();
  if (w == 0) {
    const bool own = l < pw;
    T* row = M + (int64_t)(p + (own ? l : 0)) * ld + p;
    T x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = (own && c < pw) ? row[c] : T(0);
    // reciprocal of the next pivot, formed from the broadcast pivot so that
    // every lane takes the same branch of pivot_recip
    T rinv = pivot_recip(bcast_lane(x[0], 0));
    T my_rinv = rinv;      // reciprocal of this lane's own pivot
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < pw) {
        const T rj = rinv;
        const bool act = own && l > j;
        // inactive lanes use a zero multiplier: x - 0 u = x for the finite
        // pivot rows, which avoids a select per element
        const T lv = act ? x[j] * rj : T(0);
        if (act) {
          x[j] = lv;
          Lc[j * NB + l] = lv;
        }
        if (j + 1 < NB) {
          const T u = bcast_lane(x[j + 1], j);
          x[j + 1] -= lv * u;
          rinv = pivot_recip(bcast_lane(x[j + 1], j + 1));
          if (l == j + 1) my_rinv = rinv;
        }
#pragma unroll
        for (int c = j + 2; c < NB; ++c) x[c] -= lv * bcast_lane(x[c], j);
        lds_order();
        if (l == 0) *cols = j + 1;
        if ((j % GRP) == GRP - 1 || j == pw - 1) {
          if (own && l <= j && l > j - GRP) {
#pragma unroll
            for (int c = 0; c < NB; ++c) U[l * UP + c] = x[c];
            U[l * UP + NB] = my_rinv;
          }
          lds_order();
          if (l == 0) *rows = j + 1;
        }
      }
    }
    PP_LDS T* S = tiles;
    if (own) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < pw) S[l * SP + c] = x[c];
    }
    wave_store_rows(M + (int64_t)p * ld + p, ld, S, pw, pw);
  } else if (w == 1) {
    const bool own = l < pw;
    T e[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) e[c] = (c == l) ? T(1) : T(0);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < pw) {
        lds_wait_ge(cols, j + 1);  // multiplier column j is in LDS
        const bool act = own && l > j;
        const T lij = act ? (T)Lc[j * NB + l] : T(0);
#pragma unroll
        for (int c = 0; c <= j; ++c) e[c] -= lij * bcast_lane(e[c], j);  // row j of L11^{-1} is final
      }
    }
    PP_LDS T* S = tiles + NB * SP;
    if (own) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < pw) S[l * SP + c] = e[c];
    }
    wave_store_rows(linv, NB, S, pw, pw);
  } else {
    const int nb_t = blockDim.x - 128, bt = tid - 128;
    for (int r0 = pw; r0 < R; r0 += 2 * nb_t) {
      // two rows per thread: r0 + bt and r0 + nb_t + bt
      const int ra = r0 + bt, rb = r0 + nb_t + bt;
      const bool oa = ra < R, ob = rb < R;
      T* pa = M + (int64_t)(p + (oa ? ra : 0)) * ld + p;
      T* pb = M + (int64_t)(p + (ob ? rb : 0)) * ld + p;
      T xa[NB], xb[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        xa[c] = (oa && c < pw) ? pa[c] : T(0);
        xb[c] = (ob && c < pw) ? pb[c] : T(0);
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < pw) {
          // the first row group waits for each published group of pivots
          if (r0 == pw && (j % GRP) == 0) lds_wait_ge(rows, (j + GRP < pw) ? j + GRP : pw);
          // the whole pivot row is requested before any of it is used (one
          // LDS round trip per step)
          T pr[NB + 1];
#pragma unroll
          for (int c = j + 1; c <= NB; ++c) pr[c] = U[j * UP + c];
          __builtin_amdgcn_sched_barrier(0);
          T la = xa[j] * pr[NB], lb = xb[j] * pr[NB];
          // keeps later steps' LDS reads from being hoisted above this one
          asm volatile("" : "+v"(la), "+v"(lb)::"memory");
          xa[j] = la;
          xb[j] = lb;
#pragma unroll
          for (int c = j + 1; c < NB; ++c) {
            xa[c] -= la * pr[c];
            xb[c] -= lb * pr[c];
          }
        }
      }
      // this wave's rows: r0 + off + [0, 64) and r0 + nb_t + off + [0, 64)
      PP_LDS T* S = tiles + (2 * NB + (w - 2) * 128) * SP;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < pw) {
          S[l * SP + c] = xa[c];
          S[(64 + l) * SP + c] = xb[c];
        }
      }
      const int a0 = r0 + (bt - l), b0 = a0 + nb_t;
      wave_store_rows(M + (int64_t)(p + a0) * ld + p, ld, S, min(64, R - a0), pw);
      wave_store_rows(M + (int64_t)(p + b0) * ld + p, ld, S + 64 * SP, min(64, R - b0), pw);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------- Cholesky
// Control word in the LDS arena (past every routine's working area): set to 1
// by chol_upper when the factorisation failed even with the shift.
__device__ __forceinline__ volatile PP_LDS int* chol_fail_flag() {
  return (volatile PP_LDS int*)(lds_base() + kLdsBytes - 16);
}

// Relative pivot below which a factorisation after the first drops a column
// in mixed precision (chol_upper): after the first (shifted) pass the
// sketch's independent columns have squared norms near 1 and the dependent
// ones near the rounding level, which FP32 resolves only to about k u.
template <typename T>
__device__ __forceinline__ constexpr T chol_drop_tol() { return sizeof(T) == 8 ? T(1e-12) : T(1e-4); }

// G (k x k, pitch k, symmetric positive semi-definite; k <= 64) <- R, the
// upper Cholesky factor of G.  first: the first factorisation of an accum;
// when its plain factorisation meets a non-positive pivot, G + s I is
// factorised instead, s = 11 (rows k + k (k + 1)) u tr(G) (the shift of
// shifted CholeskyQR3, Fukaya et al., SIAM J. Sci. Comput. 42 (2020)); the
// failure flag is reset there.  Later factorisations are not shifted (a
// failure sets the flag and the caller orthonormalises by Householder) unless
// `retry_all` (mixed precision): then a later factorisation never fails on a
// small pivot.  A pivot at or below chol_drop_tol times the largest diagonal
// entry marks a column that depends on the earlier ones to within rounding:
// its row of R is set to zero with an infinite diagonal (the inverses use
// 1 / r_jj = 0), so that column of Q is zero and Q Q^T projects onto the
// independent columns.  Shifting every pass instead (as the first) kept the
// independent columns of rank-deficient sketches short of unit length, and
// Q Q^T A came out about 10 % small; the Householder fallback costs several
// times the accum, and the factorization is a preconditioner for FP64
// refinement built from BF16 products anyway.
// Entries below the diagonal are set to zero.
template <typename T>
__device__ __forceinline__ void chol_upper(T* G, int k, int rows, bool first, bool retry_all = false) {
  PP_LDS T* s = (PP_LDS T*)lds_base();  // k x (k + 1)
  const int sp = k + 1, tid = threadIdx.x, nt = blockDim.x;
  const T u = sizeof(T) == 8 ? T(1.1102230246251565e-16) : T(5.9604644775390625e-08);
  const T inf = T(__builtin_inf());
  bool ok = false;
  for (int attempt = 0; attempt < (first ? 2 : 1) && !ok; ++attempt) {
    // eight loads in flight per thread (global address space, clamped
    // addresses: see narrow_solve's staging)
    for (int e0 = tid; e0 < k * k; e0 += 8 * nt) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        v[u] = gload<T>(G + (e < k * k ? e : 0));
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt, i = e / k, j = e - i * k;
        if (e < k * k) s[i * sp + j] = v[u];
      }
    }
    __syncthreads();
    if (attempt == 1) {
      T tr = T(0);
      for (int i = 0; i < k; ++i) tr += s[i * sp + i];
      const T shift = T(11) * (T(rows) * T(k) + T(k) * T(k + 1)) * u * tr;
      __syncthreads();
      for (int i = tid; i < k; i += nt) s[i * sp + i] += shift;
      __syncthreads();
    }
    T drop = T(0);
    if (retry_all && !first) {
      T gmax = T(0);
      for (int i = 0; i < k; ++i) gmax = fmax(gmax, s[i * sp + i]);
      drop = chol_drop_tol<T>() * gmax;
    }
    ok = true;
    for (int j = 0; j < k; ++j) {
      const T d = s[j * sp + j];  // every thread reads the same pivot
      if (!(d < inf) || (drop == T(0) && !(d > T(0)))) {
        ok = false;
        break;
      }
      const bool dep = drop > T(0) && !(d > drop);
      const T r = dep ? inf : sqrt(d), ri = dep ? T(0) : T(1) / r;
      __syncthreads();
      for (int c = j + tid; c < k; c += nt) s[j * sp + c] = c == j ? r : s[j * sp + c] * ri;
      __syncthreads();
      if (dep) continue;  // row j is zero right of the diagonal: no update
      const int m = k - j - 1;  // trailing upper triangle, row-major over (a, b)
      for (int e = tid; e < m * m; e += nt) {
        const int a = j + 1 + e / m, b = j + 1 + e % m;
        if (b >= a) s[a * sp + b] -= s[j * sp + a] * s[j * sp + b];
      }
      __syncthreads();
    }
    __syncthreads();
  }
  for (int e = tid; e < k * k; e += nt) {
    const int i = e / k, j = e - i * k;
    G[e] = j >= i ? s[i * sp + j] : T(0);
  }
  if (tid == 0) {
    if (first) *chol_fail_flag() = ok ? 0 : 1;
    else if (!ok) *chol_fail_flag() = 1;
  }
  __syncthreads();
}

// Scratch elements the routines above need for an n x n factor.
__host__ __device__ inline int64_t tri_scratch(int n) { return (int64_t)((n + NB - 1) / NB) * NB * NB; }

// ---------------------------------------------------------------- reductions
// Block-wide sum of one value per thread (deterministic fixed tree order).
template <typename T>
__device__ T block_sum(T v, PP_LDS T* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const T r = red[0];
  __syncthreads();
  return r;
}

// w[c] = sum_{i=r0}^{ny-1} v[i] * Y[i][c] for c in [c0, kc), v = column j of Y
// with implicit v[r0] = 1.  Partial sums over row groups, fixed-order reduce.
template <typename T>
__device__ void reflector_dots(const T* Y, int ld, int ny, int r0, int jcol, int c0, int kc, T* w, PP_LDS T* part) {
  const int nt = blockDim.x;
  const int ncol = kc - c0;
  if (ncol <= 0) {
    __syncthreads();
    return;
  }
  const int groups = nt / ncol > 0 ? nt / ncol : 1;
  const int g = threadIdx.x / ncol, c = c0 + threadIdx.x % ncol;
  T s = T(0);
  if (g < groups && (int)threadIdx.x < groups * ncol) {
    for (int i = r0 + g; i < ny; i += groups) {
      const T vi = (i == r0) ? T(1) : Y[(int64_t)i * ld + jcol];
      s += vi * Y[(int64_t)i * ld + c];
    }
    part[g * ncol + (c - c0)] = s;
  }
  __syncthreads();
  for (int cc = threadIdx.x; cc < ncol; cc += nt) {
    T t = T(0);
    for (int gg = 0; gg < groups; ++gg) t += part[gg * ncol + cc];
    w[c0 + cc] = t;
  }
  __syncthreads();
}

// Householder QR of Y (ny x k, pitch ld) in place: R in the upper triangle,
// reflector tails below the diagonal (v[j] = 1 implicit), tau[k].
// Then Q (ny x k, pitch ldq) = H_0 ... H_{k-1} [I; 0].
template <typename T>
__device__ __forceinline__ void householder_qr_q(T* Y, int ld, int ny, int k, T* Q, int ldq, T* tau, T* w, PP_LDS T* red) {
  const int nt = blockDim.x;
  PP_LDS T* part = red;  // blockDim.x scratch
  for (int j = 0; j < k && j < ny; ++j) {
    // norm of Y[j+1:, j]
    T s = T(0);
    for (int i = j + 1 + threadIdx.x; i < ny; i += nt) {
      const T y = Y[(int64_t)i * ld + j];
      s += y * y;
    }
    const T sig = block_sum(s, red);
    const T alpha = Y[(int64_t)j * ld + j];
    T t, beta, scale;
    if (sig == T(0)) {
      t = T(0);
      beta = alpha;
      scale = T(0);
    } else {
      const T nrm = sqrt(alpha * alpha + sig);
      beta = alpha <= T(0) ? nrm : -nrm;
      t = (beta - alpha) / beta;
      scale = T(1) / (alpha - beta);
    }
    for (int i = j + 1 + threadIdx.x; i < ny; i += nt) Y[(int64_t)i * ld + j] *= scale;
    __syncthreads();
    if (threadIdx.x == 0) {
      Y[(int64_t)j * ld + j] = beta;
      tau[j] = t;
    }
    // apply H_j = I - t v v^T to Y[j:, j+1:k]
    reflector_dots(Y, ld, ny, j, j, j + 1, k, w, part);
    for (int e = threadIdx.x; e < (ny - j) * (k - j - 1); e += nt) {
      const int i = j + e / (k - j - 1), c = j + 1 + e % (k - j - 1);
      const T vi = (i == j) ? T(1) : Y[(int64_t)i * ld + j];
      Y[(int64_t)i * ld + c] -= t * vi * w[c];
    }
    __syncthreads();
  }
  // explicit Q, applying reflectors from the last to the first
  for (int e = threadIdx.x; e < ny * k; e += nt) {
    const int i = e / k, c = e % k;
    Q[(int64_t)i * ldq + c] = (i == c) ? T(1) : T(0);
  }
  __syncthreads();
  const int kk = k < ny ? k : ny;
  for (int j = kk - 1; j >= 0; --j) {
    const T t = tau[j];
    // w[c] = v^T Q[j:, c], c in [j, k)
    {
      const int ncol = k - j;
      const int groups = nt / ncol > 0 ? nt / ncol : 1;
      const int g = threadIdx.x / ncol, c = j + threadIdx.x % ncol;
      if (g < groups && (int)threadIdx.x < groups * ncol) {
        T s = T(0);
        for (int i = j + g; i < ny; i += groups) {
          const T vi = (i == j) ? T(1) : Y[(int64_t)i * ld + j];
          s += vi * Q[(int64_t)i * ldq + c];
        }
        part[g * ncol + (c - j)] = s;
      }
      __syncthreads();
      for (int cc = threadIdx.x; cc < ncol; cc += nt) {
        T a = T(0);
        for (int gg = 0; gg < groups; ++gg) a += part[gg * ncol + cc];
        w[j + cc] = a;
      }
      __syncthreads();
    }
    for (int e = threadIdx.x; e < (ny - j) * (k - j); e += nt) {
      const int i = j + e / (k - j), c = j + e % (k - j);
      const T vi = (i == j) ? T(1) : Y[(int64_t)i * ld + j];
      Q[(int64_t)i * ldq + c] -= t * vi * w[c];
    }
    __syncthreads();
  }
}

}  // namespace pp
