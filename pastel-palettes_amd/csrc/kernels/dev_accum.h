// Low-rank accumulation of a small block in one workgroup step, the sketch
// resident in LDS (reference blockLowRankAccum, src/kernel.cu:1255-1280):
//   U1 V1^T <- Q Q^T (U1 V1^T + U2 V2^T),  Q an orthonormal basis of
//   Y = (U1 V1^T + U2 V2^T) Omega  (Omega = the nx x k1 sketch, pitch k1).
//
// The step-by-step form (dev_steps.h) runs about twenty global-memory GEMM
// steps and three barrier-bound Cholesky factorisations; for the blocks of a
// HODLR factorization (256-512 rows, rank 32) each step is a global round
// trip and the whole accum costs 250-350 us on one CU.  Here Y never leaves
// LDS:
//   1. S1 = Omega^T V1, S2 = Omega^T V2 (k1 x k1, k1 x k2): row chunks of
//      Omega, V1, V2 staged into LDS with coalesced loads (every load of a
//      chunk issued before the first LDS store), each wave one 16 x 16
//      quadrant of both products over all rows (no cross-wave reduction);
//   2. Y = U1 S1^T + U2 S2^T into LDS (32-row MFMA tiles, U fragments
//      loaded straight from global memory, S fragments held in registers);
//   3. shifted CholeskyQR3, three passes of G = Y^T Y (LDS, quadrant per
//      wave), R = chol(G) by one wave in LDS (lane per column),
//      R^{-1} column per lane,
//      Y <- Y R^{-1} in place (MFMA tiles from LDS); a pass whose G is the
//      identity to 64 u ends the iteration (Y is orthonormal already);
//   4. P1 = Q^T U1, P2 = Q^T U2 (Q = Y; U chunks staged as in 1);
//      V1 <- V1 P1^T + V2 P2^T in place (tiles of V1 and V2 from global
//      memory), U1 <- Q.
// The same products, in the same precisions, as the step form: sketch and
// projection in the update mode (BF16 in mixed precision), Gram matrix and
// Y R^{-1} in the leaf mode, Cholesky in the storage precision.  The shift
// and column-drop rules are chol_upper's.  When a factorisation fails even with the shift the
// failure flag is set and U1, V1 are left untouched, so the caller falls
// back to Householder QR from the start.
#pragma once

#include "dev_lu.h"

namespace pp {

// LDS row pitch of the resident blocks (16-byte aligned rows; the fragment
// reads of a quadrant spread over the banks)
template <typename T>
constexpr int acc_pitch() { return NB + (sizeof(T) == 8 ? 2 : 4); }
// Largest number of rows of U1 (the sketch's rows) held in LDS
template <typename T>
constexpr int acc_max_rows() { return sizeof(T) == 8 ? 256 : 512; }
// Rows per staged chunk: the sketch stages three matrices, the projection
// two; every thread holds 48 FP32 (24 FP64) values of a chunk in flight
template <typename T>
constexpr int acc_chunk_sketch() { return sizeof(T) == 8 ? 64 : 128; }
template <typename T>
constexpr int acc_chunk_proj() { return sizeof(T) == 8 ? 64 : 192; }

// Layout (elements of T): Y [acc_max_rows][P] at 0 (the sketch's staging
// uses it before Y exists), S1, S2, RT [NB][P] (RT = (R^{-1})^T), G
// [NB][NB + 1] (column NB: reciprocals of R's diagonal), the projection's
// staging, then the status word.
template <typename T>
struct AccLds {
  static constexpr int P = acc_pitch<T>();
  static constexpr int Y = 0;
  static constexpr int S1 = acc_max_rows<T>() * P;
  static constexpr int S2 = S1 + NB * P;
  static constexpr int RT = S2 + NB * P;
  static constexpr int G = RT + NB * P;
  static constexpr int STG = G + NB * (NB + 1);
  static constexpr int END = STG + 2 * acc_chunk_proj<T>() * P;
  static constexpr int BYTES = END * (int)sizeof(T) + 16;
  __device__ static __forceinline__ PP_LDS T* at(int off) { return (PP_LDS T*)lds_base() + off; }
  __device__ static __forceinline__ volatile PP_LDS int* status() {
    return (volatile PP_LDS int*)((PP_LDS T*)lds_base() + END);
  }
};
// the accum failure flag sits in the last 16 bytes of the arena
static_assert(AccLds<double>::BYTES <= kLdsBytes - 16, "accum LDS layout (FP64)");
static_assert(AccLds<float>::BYTES <= kLdsBytes - 16, "accum LDS layout (FP32)");
static_assert(3 * acc_chunk_sketch<double>() * AccLds<double>::P <= AccLds<double>::S1, "sketch staging (FP64)");
static_assert(3 * acc_chunk_sketch<float>() * AccLds<float>::P <= AccLds<float>::S1, "sketch staging (FP32)");

// Rows of a chunk of the sketch in LDS when the sketch itself stays in the
// worker scratch (accum_big)
template <typename T>
constexpr int acc_big_chunk() { return sizeof(T) == 8 ? 96 : 256; }

// Layout of accum_big (elements of T): the chunk of Y [acc_big_chunk][P] at
// 0 (the sketch's staging uses [0, 3 acc_chunk_sketch P) before it), S1, S2,
// RT, G as in AccLds, the staging of U1 and U2 chunks, the status word.
template <typename T>
struct AccBig {
  static constexpr int P = acc_pitch<T>(), CH = acc_big_chunk<T>();
  static constexpr int YC = 0;
  static constexpr int S1 = (3 * acc_chunk_sketch<T>() > CH ? 3 * acc_chunk_sketch<T>() : CH) * P;
  static constexpr int S2 = S1 + NB * P;
  static constexpr int RT = S2 + NB * P;
  static constexpr int G = RT + NB * P;
  static constexpr int STG = G + NB * (NB + 1);
  static constexpr int END = STG + 2 * CH * P;
  static constexpr int BYTES = END * (int)sizeof(T) + 16;
  __device__ static __forceinline__ PP_LDS T* at(int off) { return (PP_LDS T*)lds_base() + off; }
  __device__ static __forceinline__ volatile PP_LDS int* status() {
    return (volatile PP_LDS int*)((PP_LDS T*)lds_base() + END);
  }
};
static_assert(AccBig<double>::BYTES <= kLdsBytes - 16, "streamed accum LDS layout (FP64)");
static_assert(AccBig<float>::BYTES <= kLdsBytes - 16, "streamed accum LDS layout (FP32)");

// Whether accum_big handles an accumulation of these dimensions (larger
// sketches than accum_lds holds; the sketch Y is kept in the worker scratch,
// ny x k1 with pitch k1, as the step form keeps it).
template <typename T>
__device__ __forceinline__ bool accum_big_ok(int nx, int ny, int k1, int k2) {
  return blockDim.x == 256 && k1 >= 1 && k1 <= NB && k2 >= 1 && k2 <= NB && ny > acc_max_rows<T>() && nx >= 1;
}

// Whether accum_lds handles an accumulation of these dimensions.
template <typename T>
__device__ __forceinline__ bool accum_lds_ok(int nx, int ny, int k1, int k2) {
  return blockDim.x == 256 && k1 >= 1 && k1 <= NB && k2 >= 1 && k2 <= NB && ny >= k1 && ny <= acc_max_rows<T>() &&
         nx >= 1;
}

// Rows [r0, r0 + CH) of NM row-major matrices (src[m], pitch ld[m], cols[m]
// valid columns, `rows` valid rows) into LDS at dst + m CH P (pitch P); rows
// past `rows` and columns past cols[m] are zero.  Every load of the chunk is
// issued before the first store (clamped addresses, no branches).  All
// threads; ends with a barrier.
template <typename T, int NM, int CH>
__device__ __forceinline__ void acc_stage(PP_LDS T* dst, const T* const (&src)[NM], const int (&ld)[NM],
                                          const int (&cols)[NM], int r0, int rows) {
  constexpr int P = acc_pitch<T>(), PER = CH / 8, E = NM * PER;  // 256 threads, 32 columns
  const int tid = threadIdx.x, col = tid & 31, rsub = tid >> 5;
  T v[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int m = u / PER, row = rsub + 8 * (u % PER), gr = r0 + row;
    const bool in = gr < rows && col < cols[m];
    v[u] = gload<T>(src[m] + (in ? (int64_t)gr * ld[m] + col : 0));
  }
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int m = u / PER, row = rsub + 8 * (u % PER), gr = r0 + row;
    const bool in = gr < rows && col < cols[m];
    dst[m * CH * P + row * P + col] = in ? v[u] : T(0);
  }
  __syncthreads();
}

// acc += X^T Z over rows [0, kn) (kn a multiple of KW) for the 16 x 16
// quadrant (qi, qj): X, Z in LDS with pitch P.  Two accumulators (even and
// odd chunks) keep consecutive MFMAs independent.
template <typename T, int MODE>
__device__ __forceinline__ void acc_quad(const PP_LDS T* X, const PP_LDS T* Z, int kn, int qi, int qj,
                                         typename Mma<T, MODE>::acc_t (&acc)[2]) {
  using Pm = Mma<T, MODE>;
  constexpr int KW = Pm::KW, EPL = Pm::EPL, P = acc_pitch<T>();
  const int l = lane_id(), g = l >> 4, li = l & 15;
  const PP_LDS T* x = X + (EPL * g) * P + 16 * qi + li;
  const PP_LDS T* z = Z + (EPL * g) * P + 16 * qj + li;
  int kb = 0;
  for (; kb + 2 * KW <= kn; kb += 2 * KW) {
    T fa[2][EPL], fb[2][EPL];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        fa[h][e] = x[(kb + h * KW + e) * P];
        fb[h][e] = z[(kb + h * KW + e) * P];
      }
    acc[0] = Pm::mma_chunk(fa[0], fb[0], acc[0]);
    acc[1] = Pm::mma_chunk(fa[1], fb[1], acc[1]);
  }
  if (kb < kn) {
    T fa[EPL], fb[EPL];
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      fa[e] = x[(kb + e) * P];
      fb[e] = z[(kb + e) * P];
    }
    acc[0] = Pm::mma_chunk(fa, fb, acc[0]);
  }
}

// Stores quadrant (qi, qj) of a 32 x 32 product (sum of the two
// accumulators) to LDS at D (pitch dp).
template <typename T, int MODE>
__device__ __forceinline__ void acc_quad_store(PP_LDS T* D, int dp, int qi, int qj,
                                               const typename Mma<T, MODE>::acc_t (&acc)[2]) {
  using Pm = Mma<T, MODE>;
  const int l = lane_id(), li = l & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) D[(16 * qi + Pm::orow(l, r)) * dp + 16 * qj + li] = acc[0][r] + acc[1][r];
}

// B fragments of a product with S^T: B(k, c) = S[c][k], S in LDS (pitch P).
template <typename T, int MODE>
__device__ __forceinline__ void acc_frag_st(const PP_LDS T* S, T (&fb)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL]) {
  using Pm = Mma<T, MODE>;
  constexpr int KW = Pm::KW, EPL = Pm::EPL, CH = NB / KW, P = acc_pitch<T>();
  const int l = lane_id(), g = l >> 4, li = l & 15;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < EPL; ++e) fb[c][j][e] = S[(16 * j + li) * P + c * KW + EPL * g + e];
}

// A fragments of the 32-row tile at r0 of a row-major matrix in global memory
// (pitch ld, `rows` rows, kv valid columns; others read as zero).
template <typename T, int MODE>
__device__ __forceinline__ void acc_frag_rows_g(const T* X, int ld, int rows, int kv, int r0,
                                                T (&fa)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL]) {
  using Pm = Mma<T, MODE>;
  constexpr int KW = Pm::KW, EPL = Pm::EPL, CH = NB / KW, V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const int l = lane_id(), g = l >> 4, li = l & 15;
  // FP32, all NB columns valid with 16-byte aligned rows: whole-vector loads
  // (in FP64 the vector form enlarges the accum functions' register frames
  // more than it saves)
  const bool vec = sizeof(T) == 4 && kv == NB && (ld % V) == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int row = r0 + 16 * i + li, rs = row < rows ? row : rows - 1;
    if (vec) {
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int q = 0; q < EPL / V; ++q) {
          const vec_t x = gload<vec_t>(X + (int64_t)rs * ld + c * KW + EPL * g + q * V);
#pragma unroll
          for (int e = 0; e < V; ++e) fa[c][i][q * V + e] = x[e];
        }
    } else {
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
          const int k = c * KW + EPL * g + e;
          fa[c][i][e] = gload<T>(X + (int64_t)rs * ld + (k < kv ? k : kv - 1));
        }
    }
  }
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < EPL; ++e) fa[c][i][e] = (c * KW + EPL * g + e) < kv ? fa[c][i][e] : T(0);
}

// A fragments of the 32-row tile at r0 of Y (LDS, pitch P).
template <typename T, int MODE>
__device__ __forceinline__ void acc_frag_rows_l(const PP_LDS T* Y, int r0,
                                                T (&fa)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL]) {
  using Pm = Mma<T, MODE>;
  constexpr int KW = Pm::KW, EPL = Pm::EPL, CH = NB / KW, P = acc_pitch<T>();
  const int l = lane_id(), g = l >> 4, li = l & 15;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < EPL; ++e) fa[c][i][e] = Y[(r0 + 16 * i + li) * P + c * KW + EPL * g + e];
}

template <typename T, int MODE>
__device__ __forceinline__ void acc_tile_mma(const T (&fa)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL],
                                             const T (&fb)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL],
                                             typename Mma<T, MODE>::acc_t (&acc)[2][2]) {
  using Pm = Mma<T, MODE>;
  constexpr int CH = NB / Pm::KW;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = Pm::mma_chunk(fa[c][i], fb[c][j], acc[i][j]);
}

// Out (rows x kout, row-major) = X1 S1^T + X2 S2^T with X1, X2 in global
// memory (kv1 / kv2 valid columns) and S1, S2 in LDS.  Out is Y in LDS
// (OUT_LDS, rows past `rows` stored as zero up to the next multiple of 32)
// or a global matrix (pitch ldo, only rows < rows and columns < kout
// stored; it may be X1 itself: a tile's operands are loaded before it is
// stored, tiles belong to one wave each, and the tile a wave prefetches is
// its own next one).
template <typename T, int MODE, bool OUT_LDS>
__device__ __forceinline__ void acc_rows_2(const T* X1, int ld1, int kv1, const T* X2, int ld2, int kv2, int rows,
                                           const PP_LDS T* S1, const PP_LDS T* S2, PP_LDS T* OutL, T* OutG, int ldo,
                                           int kout) {
  using Pm = Mma<T, MODE>;
  using acc_t = typename Pm::acc_t;
  constexpr int KW = Pm::KW, EPL = Pm::EPL, CH = NB / KW, P = acc_pitch<T>();
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id(), li = l & 15;
  T fb1[CH][2][EPL], fb2[CH][2][EPL];
  acc_frag_st<T, MODE>(S1, fb1);
  acc_frag_st<T, MODE>(S2, fb2);
  const int ntile = (rows + 31) >> 5;
  // the fragments of the wave's next tile are requested before the current
  // tile's products (past the last tile, the last tile again: rows clamped)
  T fa1[2][CH][2][EPL], fa2[2][CH][2][EPL];
  auto load = [&](auto b_c, int t) {
    constexpr int bb = decltype(b_c)::value;
    const int r0 = (t < ntile ? t : ntile - 1) * 32;
    acc_frag_rows_g<T, MODE>(X1, ld1, rows, kv1, r0, fa1[bb]);
    acc_frag_rows_g<T, MODE>(X2, ld2, rows, kv2, r0, fa2[bb]);
  };
  auto step = [&](auto b_c, int t, auto pre_c) {
    constexpr int bb = decltype(b_c)::value;
    const int r0 = t * 32;
    if constexpr (decltype(pre_c)::value) {
      load(IC<bb ^ 1>{}, t + nw);
      __builtin_amdgcn_sched_barrier(0);
    }
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) acc[i][0] = acc[i][1] = Pm::zero();
    acc_tile_mma<T, MODE>(fa1[bb], fb1, acc);
    acc_tile_mma<T, MODE>(fa2[bb], fb2, acc);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + 16 * i + Pm::orow(l, r), col = 16 * j + li;
          if constexpr (OUT_LDS) {
            OutL[row * P + col] = row < rows ? acc[i][j][r] : T(0);
          } else {
            if (row < rows && col < kout) gstore<T>(OutG + (int64_t)row * ldo + col, acc[i][j][r]);
          }
        }
  };
  if constexpr (sizeof(T) == 4) {
    if (w < ntile) load(IC<0>{}, w);
    for (int t = w; t < ntile; t += 2 * nw) {
      step(IC<0>{}, t, IC<1>{});
      if (t + nw < ntile) step(IC<1>{}, t + nw, IC<1>{});
    }
  } else {
    // FP64: no prefetch (the second buffer enlarges the register frames of
    // the accum functions more than the overlap saves)
    for (int t = w; t < ntile; t += nw) {
      load(IC<0>{}, t);
      step(IC<0>{}, t, IC<0>{});
    }
  }
}

// Y <- Y R^{-1} in place (rows_pad rows, a multiple of 32; RT = (R^{-1})^T).
template <typename T, int MODE>
__device__ __forceinline__ void acc_rows_rinv(PP_LDS T* Y, int rows_pad, const PP_LDS T* RT) {
  using Pm = Mma<T, MODE>;
  using acc_t = typename Pm::acc_t;
  constexpr int KW = Pm::KW, EPL = Pm::EPL, CH = NB / KW, P = acc_pitch<T>();
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id(), li = l & 15;
  T fb[CH][2][EPL];
  acc_frag_st<T, MODE>(RT, fb);
  for (int r0 = 32 * w; r0 < rows_pad; r0 += 32 * nw) {
    T fa[CH][2][EPL];
    acc_frag_rows_l<T, MODE>(Y, r0, fa);
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) acc[i][0] = acc[i][1] = Pm::zero();
    acc_tile_mma<T, MODE>(fa, fb, acc);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) Y[(r0 + 16 * i + Pm::orow(l, r)) * P + 16 * j + li] = acc[i][j][r];
  }
}

// 1 / sqrt(d) for a positive pivot: the hardware estimate refined by Newton
// steps y <- y (3 - d y^2) / 2 (two in FP64, one in FP32), a shorter
// dependent sequence than a square root followed by a division.  Pivots
// outside [2^-1000, 2^1000] (FP64) or [2^-120, 2^120] (FP32) take the
// library functions.
template <typename T>
__device__ __forceinline__ T chol_rsqrt(T d) {
  if constexpr (sizeof(T) == 8) {
    if (!(d >= 0x1p-1000 && d <= 0x1p1000)) return T(1) / sqrt(d);
    T y = __builtin_amdgcn_rsq(d);
    const T h = T(0.5) * d;
    y = y * __builtin_fma(-h * y, y, T(1.5));
    return y * __builtin_fma(-h * y, y, T(1.5));
  } else {
    if (!(d >= 0x1p-120f && d <= 0x1p120f)) return T(1) / sqrtf(d);
    const T y = __builtin_amdgcn_rsqf(d);
    return y * __builtin_fmaf(-T(0.5) * d * y, y, T(1.5));
  }
}

// Upper Cholesky factor of G (NB x NB in LDS, pitch NB + 1, symmetric; the
// k x k leading block is the Gram matrix, the rest the identity) in place by
// wave 0, lane c holding column c of every other row (the two half-waves
// alternate rows).  Each step reads everything it needs in one LDS round
// trip (the pivot, row j and the lane's trailing rows, as they stand), forms
// 1 / r_jj, row j of R and the updated trailing rows in registers and
// writes them back; row j of R goes over G's row j (zero left of the
// diagonal, 1 / r_jj in column NB).  Entries below the diagonal of the
// trailing rows are updated too but never read.  Then lane c forms column c
// of R^{-1} and writes it as row c of RT.
// drop > 0: a pivot at or below `drop` marks a column that depends on the
// earlier ones to within rounding; its row of R and its reciprocal are set to
// zero, so column j of R^{-1}, hence of Q, is zero (chol_upper's rule).
// Returns false (RT untouched) on a non-finite pivot, or a non-positive one
// when drop is 0.  Wave 0 only; no workgroup barrier.  Not inlined: with its
// per-row branches inlined into accum_lds / accum_big their register demand
// quadrupled those functions' scratch frames.
template <typename T>
__device__ __attribute__((noinline)) bool acc_chol_rinv(PP_LDS T* G, PP_LDS T* RT, T drop) {
  constexpr int GP = NB + 1, P = acc_pitch<T>();
  const int l = lane_id(), c = l & 31, half = l >> 5;
  drop = bcast_lane<T>(drop, 0);  // the same on every lane
  // the lane's rows a = half + 2 t with 2 t + 1 <= j are final and neither
  // read nor written at step j (a branch of the whole wave per row pair; in
  // the pair 2 t = j + 1 - half, row a = j is written back unchanged)
  for (int j = 0; j < NB; ++j) {
    T ga[NB / 2], gr[NB / 2];
    // the pivot is the same on every lane: taken as a scalar, so the tests
    // on it branch the whole wave
    const T d = bcast_lane<T>(G[j * GP + j], 0);
    const T gjc = G[j * GP + c];
#pragma unroll
    for (int t = 0; t < NB / 2; ++t) {
      if (2 * t + 1 <= j) continue;
      const int a = half + 2 * t;
      ga[t] = G[a * GP + c];  // G[a][c]
      gr[t] = G[j * GP + a];  // G[j][a]
    }
    if (!(d < T(__builtin_inf())) || (drop == T(0) && !(d > T(0)))) return false;
    const bool dep = drop > T(0) && !(d > drop);
    const T ri = dep ? T(0) : chol_rsqrt<T>(d);
    const T rj = (c < j || dep) ? T(0) : (c == j ? d * ri : gjc * ri);  // R[j][c]
    lds_order();
    // the trailing rows of the lane (rows a <= j written back unchanged),
    // then row j of R over it
#pragma unroll
    for (int t = 0; t < NB / 2; ++t) {
      if (2 * t + 1 <= j) continue;
      const int a = half + 2 * t;
      G[a * GP + c] = ga[t] - (a > j ? (gr[t] * ri) * rj : T(0));
    }
    G[j * GP + c] = rj;
    if (l == 0) G[j * GP + NB] = ri;
    lds_order();
  }
  T y[NB];
  tri_inverse_column<T, true>(G, NB, c, y);
  if (half == 0) {
#pragma unroll
    for (int b = 0; b < NB; ++b) RT[c * P + b] = y[b];
  }
  lds_order();
  return true;
}

// The whole accumulation (see the header).  retry_all (mixed precision):
// the factorisations after the first drop dependent columns instead of
// failing (chol_upper).
// Returns with the failure flag (chol_fail_flag) set to 0 on success and 1
// when the caller must fall back; ends with a barrier.  Not inlined: runs
// once per accum entry, and inlined into the VM kernel its register demand
// made the kernel spill (FP64 16 -> 61 spilled VGPRs, FP32 70 -> 165) on
// every path; called, the kernel spills fewer than before (9 and 63).
template <typename T, int FM, int GM>
__device__ __attribute__((noinline)) void accum_lds(T* U1, int ld_u1, T* V1, int ld_v1, const T* U2, int ld_u2, const T* V2,
                                          int ld_v2, int nx, int ny, int k1, int k2, const T* rnd, bool retry_all) {
  using L = AccLds<T>;
  constexpr int P = L::P;
  using PG = Mma<T, GM>;
  using PF = Mma<T, FM>;
  const int tid = threadIdx.x, nt = blockDim.x, w = wave_id();
  const int qi = w >> 1, qj = w & 1;  // 4 waves: one quadrant each
  PP_LDS T* Y = L::at(L::Y);
  PP_LDS T* S1 = L::at(L::S1);
  PP_LDS T* S2 = L::at(L::S2);
  PP_LDS T* RT = L::at(L::RT);
  PP_LDS T* G = L::at(L::G);
  volatile PP_LDS int* status = L::status();
  const int ny_pad = (ny + 31) & ~31;

  // 1. sketch products S1 = Omega^T V1, S2 = Omega^T V2
  {
    constexpr int CH = acc_chunk_sketch<T>();
    typename PG::acc_t a1[2] = {PG::zero(), PG::zero()}, a2[2] = {PG::zero(), PG::zero()};
    const T* src[3] = {rnd, V1, V2};
    const int ld[3] = {k1, ld_v1, ld_v2}, cols[3] = {k1, k1, k2};
    for (int r0 = 0; r0 < nx; r0 += CH) {
      acc_stage<T, 3, CH>(Y, src, ld, cols, r0, nx);
      const int n = nx - r0 < CH ? nx - r0 : CH, kn = (n + PG::KW - 1) / PG::KW * PG::KW;
      acc_quad<T, GM>(Y, Y + CH * P, kn, qi, qj, a1);
      acc_quad<T, GM>(Y, Y + 2 * CH * P, kn, qi, qj, a2);
      __syncthreads();
    }
    acc_quad_store<T, GM>(S1, P, qi, qj, a1);
    acc_quad_store<T, GM>(S2, P, qi, qj, a2);
    __syncthreads();
  }
  // 2. Y = U1 S1^T + U2 S2^T (columns >= k1 are zero: so are S1's rows)
  acc_rows_2<T, GM, true>(U1, ld_u1, k1, U2, ld_u2, k2, ny, S1, S2, Y, nullptr, 0, NB);
  __syncthreads();
  // 3. shifted CholeskyQR3
  const T u = sizeof(T) == 8 ? T(1.1102230246251565e-16) : T(5.9604644775390625e-08);
  bool ok = true, orthonormal = false;
  for (int p = 0; p < 3 && ok && !orthonormal; ++p) {
    const int attempts = p == 0 ? 2 : 1;
    bool done = false;
    for (int at = 0; at < attempts && !done; ++at) {
      typename PF::acc_t ag[2] = {PF::zero(), PF::zero()};
      acc_quad<T, FM>(Y, Y, ny_pad, qi, qj, ag);
      acc_quad_store<T, FM>(G, NB + 1, qi, qj, ag);
      __syncthreads();
      if (p > 0) {
        // Y already orthonormal to working precision (|G - I| <= 64 u):
        // the remaining passes would leave it as it is
        int off = 0;
        for (int e = tid; e < NB * NB; e += nt) {
          const int i = e >> 5, j = e & 31;
          if (i < k1 && j < k1) off |= !(__builtin_fabs(G[i * (NB + 1) + j] - (i == j ? T(1) : T(0))) <= T(64) * u);
        }
        if (!__syncthreads_or(off)) {
          orthonormal = true;
          done = true;
          break;
        }
      }
      if (w == 0) {
        if (at == 1) {
          T tr = T(0);
          for (int i = 0; i < k1; ++i) tr += G[i * (NB + 1) + i];
          const T shift = T(11) * (T(ny) * T(k1) + T(k1) * T(k1 + 1)) * u * tr;
          lds_order();
          if (lane_id() < k1) G[lane_id() * (NB + 1) + lane_id()] += shift;
        }
        // the padding columns (>= k1) of Y are zero: identity there
        if (lane_id() >= k1 && lane_id() < NB) G[lane_id() * (NB + 1) + lane_id()] = T(1);
        lds_order();
        T drop = T(0);
        if (retry_all && p > 0) {
          T gmax = T(0);
          for (int i = 0; i < k1; ++i) gmax = fmax(gmax, G[i * (NB + 1) + i]);
          drop = chol_drop_tol<T>() * gmax;
        }
        const bool good = acc_chol_rinv<T>(G, RT, drop);
        if (lane_id() == 0) *status = good ? 1 : 0;
      }
      __syncthreads();
      done = *status != 0;
      __syncthreads();
    }
    if (!done) {
      ok = false;
      break;
    }
    if (orthonormal) break;
    acc_rows_rinv<T, FM>(Y, ny_pad, RT);
    __syncthreads();
  }
  if (!ok) {
    if (tid == 0) *chol_fail_flag() = 1;
    __syncthreads();
    return;
  }
  // 4. projection: P1 = Q^T U1 (into S1), P2 = Q^T U2 (into S2)
  {
    constexpr int CH = acc_chunk_proj<T>();
    PP_LDS T* stg = L::at(L::STG);
    typename PG::acc_t a1[2] = {PG::zero(), PG::zero()}, a2[2] = {PG::zero(), PG::zero()};
    const T* src[2] = {U1, U2};
    const int ld[2] = {ld_u1, ld_u2}, cols[2] = {k1, k2};
    for (int r0 = 0; r0 < ny; r0 += CH) {
      acc_stage<T, 2, CH>(stg, src, ld, cols, r0, ny);
      const int n = ny - r0 < CH ? ny - r0 : CH, kn = (n + PG::KW - 1) / PG::KW * PG::KW;
      acc_quad<T, GM>(Y + r0 * P, stg, kn, qi, qj, a1);
      acc_quad<T, GM>(Y + r0 * P, stg + CH * P, kn, qi, qj, a2);
      __syncthreads();
    }
    acc_quad_store<T, GM>(S1, P, qi, qj, a1);
    acc_quad_store<T, GM>(S2, P, qi, qj, a2);
    __syncthreads();
  }
  // V1 <- V1 P1^T + V2 P2^T (in place), then U1 <- Q
  acc_rows_2<T, GM, false>(V1, ld_v1, k1, V2, ld_v2, k2, nx, S1, S2, nullptr, V1, ld_v1, k1);
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  if ((k1 % V) == 0 && (ld_u1 % V) == 0 && (reinterpret_cast<uintptr_t>(U1) & 15) == 0 && (P % V) == 0) {
    const int vpr = k1 / V;
    for (int e = tid; e < ny * vpr; e += nt) {
      const int i = e / vpr, j = (e - i * vpr) * V;
      gstore<vec_t>(U1 + (int64_t)i * ld_u1 + j, *(const PP_LDS vec_t*)(Y + i * P + j));
    }
  } else {
    for (int e = tid; e < ny * k1; e += nt) {
      const int i = e / k1, j = e - i * k1;
      gstore<T>(U1 + (int64_t)i * ld_u1 + j, Y[i * P + j]);
    }
  }
  if (tid == 0) *chol_fail_flag() = 0;
  __syncthreads();
}

// Y chunk (LDS, rows [0, n) of the chunk at row r0) to the sketch in the
// worker scratch (pitch k1: columns < k1 only).
template <typename T>
__device__ __forceinline__ void acc_chunk_out(T* Yg, int k1, int r0, int n, const PP_LDS T* Yc) {
  constexpr int P = acc_pitch<T>(), V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  if ((k1 % V) == 0 && (reinterpret_cast<uintptr_t>(Yg) & 15) == 0) {
    const int vpr = k1 / V;
    for (int e = threadIdx.x; e < n * vpr; e += blockDim.x) {
      const int i = e / vpr, j = (e - i * vpr) * V;
      gstore<vec_t>(Yg + (int64_t)(r0 + i) * k1 + j, *(const PP_LDS vec_t*)(Yc + i * P + j));
    }
    return;
  }
  for (int e = threadIdx.x; e < n * k1; e += blockDim.x) {
    const int i = e / k1, j = e - i * k1;
    gstore<T>(Yg + (int64_t)(r0 + i) * k1 + j, Yc[i * P + j]);
  }
}

// The accumulation for sketches too tall for LDS: the same steps, products
// and precisions as accum_lds, with Y (ny x k1) in the worker scratch and
// streamed through LDS in chunks of acc_big_chunk rows.  Each chunk is
// formed once (Y = U1 S1^T + U2 S2^T, stored, its Gram contribution added),
// and each later CholeskyQR pass reads it once (Y <- Y R^{-1}, stored, the
// next Gram contribution added); the projection reads the chunks of Q, U1
// and U2 together and writes Q over U1's rows right after their products
// are taken.  Failure handling as accum_lds (U1, V1 untouched; the flag
// tells the caller to run the step form).
template <typename T, int FM, int GM>
__device__ __attribute__((noinline)) void accum_big(T* U1, int ld_u1, T* V1, int ld_v1, const T* U2, int ld_u2,
                                                    const T* V2, int ld_v2, int nx, int ny, int k1, int k2,
                                                    const T* rnd, T* Yg, bool retry_all) {
  using L = AccBig<T>;
  constexpr int P = L::P, CH = L::CH;
  using PG = Mma<T, GM>;
  using PF = Mma<T, FM>;
  const int tid = threadIdx.x, nt = blockDim.x, w = wave_id();
  const int qi = w >> 1, qj = w & 1;
  PP_LDS T* Yc = L::at(L::YC);
  PP_LDS T* S1 = L::at(L::S1);
  PP_LDS T* S2 = L::at(L::S2);
  PP_LDS T* RT = L::at(L::RT);
  PP_LDS T* G = L::at(L::G);
  PP_LDS T* stg = L::at(L::STG);
  volatile PP_LDS int* status = L::status();
  const T u = sizeof(T) == 8 ? T(1.1102230246251565e-16) : T(5.9604644775390625e-08);

  // 1. sketch products S1 = Omega^T V1, S2 = Omega^T V2 (as accum_lds)
  {
    constexpr int CS = acc_chunk_sketch<T>();
    typename PG::acc_t a1[2] = {PG::zero(), PG::zero()}, a2[2] = {PG::zero(), PG::zero()};
    const T* src[3] = {rnd, V1, V2};
    const int ld[3] = {k1, ld_v1, ld_v2}, cols[3] = {k1, k1, k2};
    for (int r0 = 0; r0 < nx; r0 += CS) {
      acc_stage<T, 3, CS>(Yc, src, ld, cols, r0, nx);
      const int n = nx - r0 < CS ? nx - r0 : CS, kn = (n + PG::KW - 1) / PG::KW * PG::KW;
      acc_quad<T, GM>(Yc, Yc + CS * P, kn, qi, qj, a1);
      acc_quad<T, GM>(Yc, Yc + 2 * CS * P, kn, qi, qj, a2);
      __syncthreads();
    }
    acc_quad_store<T, GM>(S1, P, qi, qj, a1);
    acc_quad_store<T, GM>(S2, P, qi, qj, a2);
    __syncthreads();
  }
  // 2. Y = U1 S1^T + U2 S2^T chunk by chunk, stored, with G = Y^T Y
  typename PF::acc_t ag[2] = {PF::zero(), PF::zero()};
  for (int r0 = 0; r0 < ny; r0 += CH) {
    const int n = ny - r0 < CH ? ny - r0 : CH, npad = (n + 31) & ~31;
    acc_rows_2<T, GM, true>(U1 + (int64_t)r0 * ld_u1, ld_u1, k1, U2 + (int64_t)r0 * ld_u2, ld_u2, k2, n, S1, S2, Yc,
                            nullptr, 0, NB);
    __syncthreads();
    acc_chunk_out<T>(Yg, k1, r0, n, Yc);
    acc_quad<T, FM>(Yc, Yc, npad, qi, qj, ag);
    __syncthreads();
  }
  // 3. shifted CholeskyQR3 over the stored chunks
  bool ok = true, orthonormal = false;
  for (int p = 0; p < 3; ++p) {
    acc_quad_store<T, FM>(G, NB + 1, qi, qj, ag);
    __syncthreads();
    if (p > 0) {
      int off = 0;
      for (int e = tid; e < NB * NB; e += nt) {
        const int i = e >> 5, j = e & 31;
        if (i < k1 && j < k1) off |= !(__builtin_fabs(G[i * (NB + 1) + j] - (i == j ? T(1) : T(0))) <= T(64) * u);
      }
      if (!__syncthreads_or(off)) {
        orthonormal = true;
        break;
      }
    }
    const int attempts = p == 0 ? 2 : 1;
    bool done = false;
    for (int at = 0; at < attempts && !done; ++at) {
      if (at == 1) {  // the shifted retry starts from the Gram matrix again
        acc_quad_store<T, FM>(G, NB + 1, qi, qj, ag);
        __syncthreads();
      }
      if (w == 0) {
        if (at == 1) {
          T tr = T(0);
          for (int i = 0; i < k1; ++i) tr += G[i * (NB + 1) + i];
          const T shift = T(11) * (T(ny) * T(k1) + T(k1) * T(k1 + 1)) * u * tr;
          lds_order();
          if (lane_id() < k1) G[lane_id() * (NB + 1) + lane_id()] += shift;
        }
        if (lane_id() >= k1 && lane_id() < NB) G[lane_id() * (NB + 1) + lane_id()] = T(1);
        lds_order();
        T drop = T(0);
        if (retry_all && p > 0) {
          T gmax = T(0);
          for (int i = 0; i < k1; ++i) gmax = fmax(gmax, G[i * (NB + 1) + i]);
          drop = chol_drop_tol<T>() * gmax;
        }
        const bool good = acc_chol_rinv<T>(G, RT, drop);
        if (lane_id() == 0) *status = good ? 1 : 0;
      }
      __syncthreads();
      done = *status != 0;
      __syncthreads();
    }
    if (!done) {
      ok = false;
      break;
    }
    // Y <- Y R^{-1} chunk by chunk, with the next pass's Gram matrix
    for (int i = 0; i < 2; ++i) ag[i] = PF::zero();
    for (int r0 = 0; r0 < ny; r0 += CH) {
      const int n = ny - r0 < CH ? ny - r0 : CH, npad = (n + 31) & ~31;
      const T* src[1] = {Yg};
      const int ld[1] = {k1}, cols[1] = {k1};
      acc_stage<T, 1, CH>(Yc, src, ld, cols, r0, ny);
      acc_rows_rinv<T, FM>(Yc, npad, RT);
      __syncthreads();
      acc_chunk_out<T>(Yg, k1, r0, n, Yc);
      acc_quad<T, FM>(Yc, Yc, npad, qi, qj, ag);
      __syncthreads();
    }
  }
  (void)orthonormal;
  if (!ok) {
    if (tid == 0) *chol_fail_flag() = 1;
    __syncthreads();
    return;
  }
  // 4. projection P1 = Q^T U1, P2 = Q^T U2 chunk by chunk; Q's rows go over
  // U1's once their products are taken
  {
    typename PG::acc_t a1[2] = {PG::zero(), PG::zero()}, a2[2] = {PG::zero(), PG::zero()};
    for (int r0 = 0; r0 < ny; r0 += CH) {
      const int n = ny - r0 < CH ? ny - r0 : CH, kn = (n + PG::KW - 1) / PG::KW * PG::KW;
      {
        const T* src[1] = {Yg};
        const int ld[1] = {k1}, cols[1] = {k1};
        acc_stage<T, 1, CH>(Yc, src, ld, cols, r0, ny);
      }
      {
        const T* src[2] = {U1, U2};
        const int ld[2] = {ld_u1, ld_u2}, cols[2] = {k1, k2};
        acc_stage<T, 2, CH>(stg, src, ld, cols, r0, ny);
      }
      acc_quad<T, GM>(Yc, stg, kn, qi, qj, a1);
      acc_quad<T, GM>(Yc, stg + CH * P, kn, qi, qj, a2);
      for (int e = tid; e < n * k1; e += nt) {
        const int i = e / k1, j = e - i * k1;
        gstore<T>(U1 + (int64_t)(r0 + i) * ld_u1 + j, Yc[i * P + j]);
      }
      __syncthreads();
    }
    acc_quad_store<T, GM>(S1, P, qi, qj, a1);
    acc_quad_store<T, GM>(S2, P, qi, qj, a2);
    __syncthreads();
  }
  // V1 <- V1 P1^T + V2 P2^T (in place)
  acc_rows_2<T, GM, false>(V1, ld_v1, k1, V2, ld_v2, k2, nx, S1, S2, nullptr, V1, ld_v1, k1);
  if (tid == 0) *chol_fail_flag() = 0;
  __syncthreads();
}

}  // namespace pp
