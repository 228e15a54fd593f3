// LU without pivoting of a diagonal leaf held entirely in LDS (reference
// DenseGetrf / blockDenseGetrf, src/kernel.cu:104-147, 820-882: unit-lower
// L, upper U, in place).
//
// The leaf is copied into LDS once, padded to whole 32 x 32 tiles with the
// identity (the padded matrix factors as [L 0; 0 I] [U 0; 0 I]), factored
// there and copied back: no global round trip between the panels, which
// dominates the leaf factorization when every panel and trailing update
// goes through global memory.  Per 32-column panel:
//   1. waves 0 and 1 factor the diagonal block from registers, each holding
//      half of its columns (pivot rows by v_readlane, as panel_lu),
//      publishing multiplier columns and groups of final rows with the pivot
//      reciprocals, then each forms half of the columns of L11^{-1} from the
//      multiplier columns; waves 2-3 eliminate the rows below against each
//      published group (one row per thread);
//   2. U12 = L11^{-1} A12 on MFMA, one 32-column tile per wave, in place;
//   3. A22 -= L21 U12 on MFMA, 32 x 32 tiles dealt round-robin to the
//      waves, both operands and C read from LDS.
// Orders up to lds_lu_max (FP32 160, FP64 128) fit next to the panel's
// work area.
#pragma once

#include "dev_lu.h"

namespace pp {

template <typename T>
constexpr int lds_lu_max() { return sizeof(T) == 8 ? 128 : 160; }

// LDS layout (compile-time, so that every access in the unrolled panel code
// is a base register plus an immediate offset): the matrix at lds_base()
// with pitch SP = MAXN + one 16-byte vector whatever the order (rows are
// read and written as whole vectors), then the panel's work
// area: lc [NB][NB] multiplier columns (column j at lc[j NB + i]), linv
// [NB][NB] L11^{-1} (row-major), rinv [NB] pivot reciprocals, three
// publication counters (16 bytes).
template <typename T, int MAXN = lds_lu_max<T>()>
struct LdsLu {
  static constexpr int SP = MAXN + 16 / (int)sizeof(T);  // 16-byte aligned rows
  static constexpr int MAT = MAXN * SP;  // elements
  static constexpr int BYTES = (MAT + 2 * NB * NB + NB) * (int)sizeof(T) + 16;
  __device__ static __forceinline__ PP_LDS T* mat() { return (PP_LDS T*)lds_base(); }
  __device__ static __forceinline__ PP_LDS T* lc() { return mat() + MAT; }
  __device__ static __forceinline__ PP_LDS T* linv() { return mat() + MAT + NB * NB; }
  __device__ static __forceinline__ PP_LDS T* rinv() { return mat() + MAT + 2 * NB * NB; }
  __device__ static __forceinline__ volatile PP_LDS int* cnt() {
    return (volatile PP_LDS int*)(mat() + MAT + 2 * NB * NB + NB);
  }
};
// the accum Cholesky's failure flag sits in the last 16 bytes of the arena
static_assert(LdsLu<double>::BYTES <= kLdsBytes - 16, "LDS LU layout (FP64)");
static_assert(LdsLu<float>::BYTES <= kLdsBytes - 16, "LDS LU layout (FP32)");

// A row of NB elements in LDS (16-byte aligned) to registers, as whole
// 16-byte vectors (a broadcast when every lane reads the same row).
template <typename T>
__device__ __forceinline__ void lds_row_read(const PP_LDS T* p, T (&v)[NB]) {
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
#pragma unroll
  for (int q = 0; q < NB / V; ++q) {
    const vec_t t = *(const PP_LDS vec_t*)(p + q * V);
#pragma unroll
    for (int e = 0; e < V; ++e) v[q * V + e] = t[e];
  }
}

// Half HALF (columns HALF NB/2 .. of row l, lanes l < NB) of the diagonal
// block of the panel, factored by one wave.  Pivot j belongs to half j / H.
// In its own half the wave forms the multipliers of pivot j (publishing
// multiplier column j, its reciprocal, and cnt[0] = j + 1) and updates its
// later columns; wave 1 applies pivots j < H to its columns with the
// published multipliers.  Each element sees the updates of the one-wave
// elimination in the same order, so the factors are the same.  Rows whose
// half is final are written to D in groups of GRP (cnt[1 + HALF]); wave 0's
// columns are final after pivot H - 1, when it writes all remaining rows.
template <typename T, int MAXN, int HALF>
__device__ __forceinline__ void lds_diag_half(PP_LDS T* D, PP_LDS T* keep_rinv, T (&x)[NB / 2]) {
  using L = LdsLu<T, MAXN>;
  constexpr int GRP = sizeof(T) == 4 ? 4 : 8;  // rows per publication
  constexpr int H = NB / 2, C0 = HALF * H;
  constexpr int sp = L::SP;
  const int l = lane_id();
  PP_LDS T* const lc = L::lc();
  PP_LDS T* const rinvs = L::rinv();
  volatile PP_LDS int* const cnt = L::cnt();
  const bool own = l < NB;
#pragma unroll
  for (int c = 0; c < H; ++c) x[c] = own ? (T)D[l * sp + C0 + c] : T(0);
  // reciprocal of the next pivot, formed from the broadcast pivot so that
  // every lane takes the same branch of pivot_recip
  T rinv = HALF == 0 ? pivot_recip(bcast_lane(x[0], 0)) : T(0);
  int seen = 0;  // multiplier columns known to be published (HALF 1)
#pragma unroll
  for (int j = 0; j < (HALF == 0 ? H : NB); ++j) {
    const bool act = own && l > j;
    if (j / H == HALF) {
      // pivot j is in this half, column jj of x
      const int jj = j - C0;
      const T rj = rinv;
      // inactive lanes use a zero multiplier: x - 0 u = x
      const T lv = act ? x[jj] * rj : T(0);
      x[jj] = act ? lv : x[jj];
      // stores without lane masks: lanes <= j store a zero multiplier that
      // is never read, lanes >= NB one into the next column (written at the
      // next pivot; after column NB - 1, into linv, which wave 0 writes after
      // the last column is published); every lane stores the same reciprocal
      lc[j * NB + l] = lv;
      rinvs[j] = rj;
      if (keep_rinv != nullptr) keep_rinv[j] = rj;
      if (jj + 1 < H) {
        const T u = bcast_lane(x[jj + 1], j);
        x[jj + 1] -= lv * u;
        rinv = pivot_recip(bcast_lane(x[jj + 1], j + 1));
      }
#pragma unroll
      for (int c = jj + 2; c < H; ++c) x[c] -= lv * bcast_lane(x[c], j);
      lds_order();
      cnt[0] = j + 1;
    } else {
      // HALF 1, pivot j < H: the multipliers come from wave 0
      if (seen <= j) {
        while ((seen = cnt[0]) <= j) __builtin_amdgcn_s_sleep(1);
        lds_order();
      }
      const T lcv = lc[j * NB + l];
      const T lv = act ? lcv : T(0);
#pragma unroll
      for (int c = 0; c < H; ++c) x[c] -= lv * bcast_lane(x[c], j);
      if (j == H - 1) rinv = pivot_recip(bcast_lane(x[0], H));
    }
    if ((j % GRP) == GRP - 1) {
      // rows <= j have final values in this half (HALF 0 after pivot H - 1:
      // every row)
      const bool put = HALF == 0 && j == H - 1 ? (own && l > j - GRP) : (own && l <= j && l > j - GRP);
      if (put) {
#pragma unroll
        for (int c = 0; c < H; ++c) D[l * sp + C0 + c] = x[c];
      }
      lds_order();
      cnt[1 + HALF] = HALF == 0 && j == H - 1 ? NB : j + 1;
    }
  }
}

// Columns HALF NB/2 .. of L11^{-1} (row l in lane l) from the multiplier
// columns, by the wave that factored that half of the diagonal block (x:
// its half of the factored rows, the multipliers of its own pivots below
// the diagonal).  Row j of the inverse is final after multiplier columns
// < j and column j updates columns <= j, so wave 0 takes every multiplier
// column (those of wave 1 from LDS once published) and wave 1 only its own
// (the earlier ones leave its columns of the inverse at zero).  Each
// element sees the updates of the one-wave formation in the same order.
template <typename T, int MAXN, int HALF>
__device__ __forceinline__ void lds_linv_half(PP_LDS T* keep_linv, const T (&x)[NB / 2]) {
  using L = LdsLu<T, MAXN>;
  constexpr int H = NB / 2, C0 = HALF * H;
  const int l = lane_id();
  PP_LDS T* const lc = L::lc();
  PP_LDS T* const linv = L::linv();
  volatile PP_LDS int* const cnt = L::cnt();
  const bool own = l < NB;
  T e[H];
#pragma unroll
  for (int c = 0; c < H; ++c) e[c] = (C0 + c == l) ? T(1) : T(0);
#pragma unroll
  for (int j = C0; j < NB; ++j) {
    const bool act = own && l > j;
    T lij;
    if (j / H == HALF) {
      lij = act ? x[j - C0] : T(0);
    } else {
      lds_wait_ge(&cnt[0], j + 1);  // multiplier column j is in LDS
      lij = act ? (T)lc[j * NB + l] : T(0);
    }
#pragma unroll
    for (int c = 0; c < H; ++c)
      if (C0 + c <= j) e[c] -= lij * bcast_lane(e[c], j);
  }
  if (own) {
#pragma unroll
    for (int c = 0; c < H; ++c) linv[l * NB + C0 + c] = e[c];
    if (keep_linv != nullptr) {
#pragma unroll
      for (int c = 0; c < H; ++c) keep_linv[l * NB + C0 + c] = e[c];
    }
  }
}

// Step 1 for the panel at k0 of the np x np matrix.
// keep_linv / keep_rinv (optional): L11^{-1} of the diagonal block
// (row-major, pitch NB) and its pivot reciprocals are also left there.
// Waves 0 and 1 factor the diagonal block (lds_diag_half) and then form the
// two column halves of L11^{-1} from the multiplier columns; waves 2-3
// eliminate the rows below against each published group of pivot rows.
template <typename T, int MAXN>
__device__ __forceinline__ void lds_panel(int k0, int np, PP_LDS T* keep_linv = nullptr, PP_LDS T* keep_rinv = nullptr) {
  using L = LdsLu<T, MAXN>;
  constexpr int GRP = sizeof(T) == 4 ? 4 : 8;  // rows per publication
  constexpr int H = NB / 2;
  constexpr int sp = L::SP;
  const int tid = threadIdx.x, l = lane_id(), w = wave_id();
  PP_LDS T* D = L::mat() + k0 * sp + k0;
  PP_LDS T* const lc = L::lc();
  PP_LDS T* const rinvs = L::rinv();
  volatile PP_LDS int* const cnt = L::cnt();
  if (tid == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
    cnt[2] = 0;
  }
  __syncthreads();
  if (w < 2) {
    // L11^{-1} serves the U12 product of this panel, so the last panel
    // forms it only when the caller keeps it
    const bool need_linv = k0 + NB < np || keep_linv != nullptr;
    T x[H];
    if (w == 0) {
      lds_diag_half<T, MAXN, 0>(D, keep_rinv, x);
      if (need_linv) lds_linv_half<T, MAXN, 0>(keep_linv, x);
    } else {
      lds_diag_half<T, MAXN, 1>(D, keep_rinv, x);
      if (need_linv) lds_linv_half<T, MAXN, 1>(keep_linv, x);
    }
  } else {
    const int bt = tid - 128, below = np - k0 - NB, pw = np - k0 < NB ? np - k0 : NB;
    const bool on = bt < below;
    PP_LDS T* row = L::mat() + (k0 + NB + (on ? bt : 0)) * sp + k0;
    T xa[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) xa[c] = on ? (T)row[c] : T(0);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      // pw == NB, but not known to the compiler: the unrolled steps are then
      // kept apart instead of interleaved (which spills xa)
      if (j >= pw) continue;
      if ((j % GRP) == 0) {
        // both halves of pivot rows j .. j + GRP - 1 (the left half of rows
        // >= H is not read: only columns > j are)
        if (j < H) lds_wait_ge(&cnt[1], j + GRP);
        lds_wait_ge(&cnt[2], j + GRP);
      }
      // the whole pivot row is requested before any of it is used (one LDS
      // round trip per step)
      T pr[NB];
      lds_row_read<T>(D + j * sp, pr);  // broadcast 16-byte reads of the whole row
      const T rr = rinvs[j];
      __builtin_amdgcn_sched_barrier(0);
      T la = xa[j] * rr;
      // keeps later steps' LDS reads from being hoisted above this one
      // (their values would all be live at once)
      asm volatile("" : "+v"(la)::"memory");
      xa[j] = la;
#pragma unroll
      for (int c = j + 1; c < NB; ++c) xa[c] -= la * pr[c];
    }
    if (on) {
#pragma unroll
      for (int c = 0; c < NB; ++c) row[c] = xa[c];
    }
  }
  __syncthreads();
}

// acc += A(r, k) B(k, c) over k < NB for one 32 x 32 tile, A at a (pitch
// sa, row-major), B at b (pitch sb), both in LDS.
template <typename T, int MODE>
__device__ __forceinline__ void lds_tile_product(const PP_LDS T* a, int sa, const PP_LDS T* b, int sb,
                                                 typename Mma<T, MODE>::acc_t (&acc)[2][2]) {
  using P = Mma<T, MODE>;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW;
  const int l = lane_id(), g = l >> 4, li = l & 15;
  T fa[CH][2][EPL], fb[CH][2][EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int k = c * KW + EPL * g + e;
        fa[c][h][e] = a[(16 * h + li) * sa + k];
        fb[c][h][e] = b[k * sb + 16 * h + li];
      }
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = P::mma_chunk(fa[c][i], fb[c][j], acc[i][j]);
}

// In-place LU of the np x np matrix (np a multiple of NB, np <= MAXN) in LDS.
// keep_linv / keep_rinv (optional): panel k leaves its L11^{-1} at
// keep_linv + k NB NB and its pivot reciprocals at keep_rinv + k NB.
template <typename T, int MODE, int MAXN = lds_lu_max<T>()>
__device__ __forceinline__ void lds_lu(int np, PP_LDS T* keep_linv = nullptr, PP_LDS T* keep_rinv = nullptr) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  using L = LdsLu<T, MAXN>;
  constexpr int sp = L::SP;
  PP_LDS T* const S = L::mat();
  const int w = wave_id(), nw = blockDim.x >> 6, l = lane_id(), li = l & 15;
  for (int k0 = 0; k0 < np; k0 += NB) {
    lds_panel<T, MAXN>(k0, np, keep_linv ? keep_linv + k0 * NB : nullptr, keep_rinv ? keep_rinv + k0 : nullptr);
    const int ctn = (np - k0 - NB) / NB;
    if (ctn == 0) break;
    // U12 = L11^{-1} A12, one column tile per wave (read fully before written)
    for (int ct = w; ct < ctn; ct += nw) {
      PP_LDS T* A12 = S + k0 * sp + k0 + NB + ct * NB;
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
      lds_tile_product<T, MODE>(L::linv(), NB, A12, sp, acc);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) A12[(16 * a + P::orow(l, r)) * sp + 16 * j + li] = acc[a][j][r];
    }
    __syncthreads();
    // A22 -= L21 U12
    const int nt = ctn * ctn;
    for (int t = w; t < nt; t += nw) {
      const int r0 = k0 + NB + (t % ctn) * NB, c0 = k0 + NB + (t / ctn) * NB;
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = P::zero();
      lds_tile_product<T, MODE>(S + r0 * sp + k0, sp, S + k0 * sp + c0, sp, acc);
      PP_LDS T* C = S + r0 * sp + c0;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            PP_LDS T* cp = C + (16 * a + P::orow(l, r)) * sp + 16 * j + li;
            *cp = *cp - acc[a][j][r];
          }
    }
    __syncthreads();
  }
}

// Copies the n x n block M (global, pitch ld) into S (pitch sp), padded to
// np with the identity.  When the rows are whole 16-byte vectors (n == np,
// 16-byte aligned rows) every thread requests all its vectors of a batch of
// up to 16 before writing any of them, so one memory latency covers the
// batch; otherwise element by element, eight in flight per thread.
template <typename T>
__device__ __forceinline__ void lds_load_padded(PP_LDS T* S, int sp, int np, const T* M, int ld, int n) {
  constexpr int V = 16 / (int)sizeof(T), KB = 16;
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const int nt = blockDim.x;
  if (n == np && (ld % V) == 0 && (reinterpret_cast<uintptr_t>(M) & 15) == 0) {
    const int nv = np / V, total = np * nv;
    for (int q0 = threadIdx.x; q0 < total; q0 += KB * nt) {
      vec_t v[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) {
        const int q = q0 + u * nt < total ? q0 + u * nt : q0, i = q / nv, c = (q - i * nv) * V;
        v[u] = gload<vec_t>(M + (int64_t)i * ld + c);
      }
#pragma unroll
      for (int u = 0; u < KB; ++u) {
        const int q = q0 + u * nt, i = q / nv, c = (q - i * nv) * V;
        if (q < total) {
#pragma unroll
          for (int e = 0; e < V; ++e) S[i * sp + c + e] = v[u][e];
        }
      }
    }
    __syncthreads();
    return;
  }
  const int total = np * np;
  for (int e0 = threadIdx.x; e0 < total; e0 += 8 * nt) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt < total ? e0 + u * nt : e0, i = e / np, c = e - i * np;
      v[u] = gload<T>(M + (int64_t)(i < n ? i : 0) * ld + (c < n ? c : 0));
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt, i = e / np, c = e - i * np;
      if (e < total) S[i * sp + c] = (i < n && c < n) ? v[u] : (i == c ? T(1) : T(0));
    }
  }
  __syncthreads();
}

// The n x n block of S (pitch sp) to M (pitch ld): 16-byte stores when the
// rows are whole aligned vectors.
template <typename T>
__device__ __forceinline__ void lds_store(T* M, int ld, int n, const PP_LDS T* S, int sp) {
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const int nt = blockDim.x;
  if ((n % V) == 0 && (ld % V) == 0 && (reinterpret_cast<uintptr_t>(M) & 15) == 0) {
    const int nv = n / V, total = n * nv;
    for (int q = threadIdx.x; q < total; q += nt) {
      const int i = q / nv, c = (q - i * nv) * V;
      vec_t v;
#pragma unroll
      for (int e = 0; e < V; ++e) v[e] = S[i * sp + c + e];
      gstore<vec_t>(M + (int64_t)i * ld + c, v);
    }
    __syncthreads();
    return;
  }
  const int total = n * n;
  for (int e = threadIdx.x; e < total; e += nt) {
    const int i = e / n, c = e - i * n;
    gstore<T>(M + (int64_t)i * ld + c, (T)S[i * sp + c]);
  }
  __syncthreads();
}

// The whole getrf of a square leaf of order n <= lds_lu_max<T>().
template <typename T, int MODE>
__device__ __forceinline__ void getrf_lds(T* M, int ld, int n) {
  using L = LdsLu<T>;
  const int np = (n + NB - 1) / NB * NB;
  lds_load_padded<T>(L::mat(), L::SP, np, M, ld, n);
  lds_lu<T, MODE>(np);
  lds_store<T>(M, ld, n, L::mat(), L::SP);
}

// ---------------------------------------------------------------- two-block LU
// FP32 leaves of order 128 < n <= 256 (a 256 x 256 FP32 leaf does not fit
// in LDS): with A = [A11 A12; A21 A22], A11 of order 128,
//   1. A11 = L11 U11 in LDS (lds_lu), stored; the inverses D of the 32 x 32
//      diagonal blocks of L11 and U11 are formed in LDS;
//   2. each wave w < nt2 (nt2 = the tiles of the trailing order) solves,
//      with no barrier, column tile w of U12 = L11^{-1} A12 and row tile w of
//      L21 = A21 U11^{-1} by blocked substitution on MFMA
//        X_i = D_i (B_i - sum_{k<i} L_ik X_k),  X_j = (B_j - sum_{k<j} X_k U_kj) D_j
//      with the right-hand side and the solved tiles in registers (an acc
//      tile is turned into operand fragments through a wave-private LDS
//      tile), L11 / U11 read from LDS; both are stored;
//   3. U12 goes to LDS, and wave w forms row tile w of A22 - L21 U12 (its L21
//      tile still in registers) straight into LDS;
//   4. A22 (the Schur complement) is factored in LDS and stored.
template <typename T>
constexpr int lds2_max() { return sizeof(T) == 4 ? 256 : 0; }

template <typename T, int MODE>
struct Lds2 {
  using L = LdsLu<T, 128>;
  using P = Mma<T, MODE>;
  static constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW;
  static constexpr int CP = NB + 1;                 // pitch of the conversion tiles
  static constexpr int UP = 128 + 1;                // pitch of U12 in LDS
  static constexpr int R0 = (L::BYTES + 15) / 16 * 16;  // byte offset of the region after the LU area
  // region: during the solves D_L[4], D_U[4] (row-major 32 x 32) and one
  // conversion tile per wave; afterwards U12 (128 x UP)
  __device__ static __forceinline__ PP_LDS T* reg() { return (PP_LDS T*)(lds_base() + R0); }
  __device__ static __forceinline__ PP_LDS T* dl(int i) { return reg() + i * NB * NB; }
  __device__ static __forceinline__ PP_LDS T* du(int i) { return reg() + (4 + i) * NB * NB; }
  __device__ static __forceinline__ PP_LDS T* ct(int w) { return reg() + 8 * NB * NB + w * NB * CP; }
  // pivot reciprocals of A11 (128), after the conversion tiles
  __device__ static __forceinline__ PP_LDS T* rin() { return reg() + 8 * NB * NB + 4 * NB * CP; }
  __device__ static __forceinline__ PP_LDS T* u12() { return reg(); }
  static constexpr int SOLVE = 8 * NB * NB + 4 * NB * CP + 128;  // elements of the region during the solves
  static constexpr int BYTES = R0 + (int)sizeof(T) * (128 * UP > SOLVE ? 128 * UP : SOLVE);
};
static_assert(Lds2<float, MODE_FP32>::BYTES <= kLdsBytes - 16, "two-block LU layout (FP32)");

// operand fragments (A: rows of the tile, B: columns) of a 32 x 32 tile held
// in accumulator layout, through the wave's conversion tile
template <typename T, int MODE>
__device__ __forceinline__ void acc_to_frags(const typename Mma<T, MODE>::acc_t (&acc)[2][2], PP_LDS T* tile,
                                             bool as_a, T (&f)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL]) {
  using P = Mma<T, MODE>;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW, CP = NB + 1;
  const int l = lane_id(), g = l >> 4, li = l & 15;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) tile[(16 * a + P::orow(l, r)) * CP + 16 * j + li] = acc[a][j][r];
  lds_order();
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int k = c * KW + EPL * g + e;
        f[c][h][e] = as_a ? (T)tile[(16 * h + li) * CP + k] : (T)tile[k * CP + 16 * h + li];
      }
  lds_order();
}

// fragments of a 32 x 32 block in LDS (pitch p): A fragments (rows) or B
// fragments (columns), optionally negated
template <typename T, int MODE>
__device__ __forceinline__ void lds_frags(const PP_LDS T* b, int p, bool as_a, bool neg,
                                          T (&f)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL]) {
  using P = Mma<T, MODE>;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW;
  const int l = lane_id(), g = l >> 4, li = l & 15;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int k = c * KW + EPL * g + e;
        const T v = as_a ? (T)b[(16 * h + li) * p + k] : (T)b[k * p + 16 * h + li];
        f[c][h][e] = neg ? -v : v;
      }
}

template <typename T, int MODE>
__device__ __forceinline__ void frag_mma(const T (&fa)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL],
                                         const T (&fb)[NB / Mma<T, MODE>::KW][2][Mma<T, MODE>::EPL],
                                         typename Mma<T, MODE>::acc_t (&acc)[2][2]) {
  using P = Mma<T, MODE>;
  constexpr int CH = NB / P::KW;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[a][j] = P::mma_chunk(fa[c][a], fb[c][j], acc[a][j]);
}

// Inverse of the upper 32 x 32 diagonal block at b (LDS, pitch p) into out
// (row-major, pitch NB), with the pivot reciprocals rinv of its diagonal:
// lane c < 32 forms column c, y = U^{-1} e_c, column-oriented (for r from
// the last row up: y_r *= 1 / u_rr, then y_k -= u_kr y_r for k < r), so the
// updates of one step are independent and the dependent chain is 32 steps
// long.  Rows below the diagonal come out zero.
template <typename T>
__device__ __forceinline__ void lds_upper_inverse(const PP_LDS T* b, int p, const PP_LDS T* rinv, PP_LDS T* out) {
  const int c = lane_id();
  if (c >= NB) return;
  T y[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) y[r] = (r == c) ? T(1) : T(0);
#pragma unroll
  for (int r = NB - 1; r >= 0; --r) {
    const T yr = y[r] * (T)rinv[r];
    y[r] = yr;
#pragma unroll
    for (int k = 0; k < r; ++k) y[k] -= (T)b[k * p + r] * yr;
  }
#pragma unroll
  for (int r = 0; r < NB; ++r) out[r * NB + c] = y[r];
}

template <typename T, int MODE>
__device__ __forceinline__ void getrf_lds2(T* M, int ld, int n) {
  using G = Lds2<T, MODE>;
  using L = typename G::L;
  using P = typename G::P;
  using acc_t = typename P::acc_t;
  constexpr int KW = G::KW, EPL = G::EPL, CH = G::CH, sp = L::SP;
  const int w = wave_id(), l = lane_id(), li = l & 15;
  const int n2 = n - 128, np2 = (n2 + NB - 1) / NB * NB, nt2 = np2 / NB;
  PP_LDS T* S = L::mat();
  // 1. A11
  lds_load_padded<T>(S, sp, 128, M, ld, 128);
  // the panels leave the inverses of L11's diagonal blocks and U11's pivot
  // reciprocals; wave w inverts U11's diagonal block w
  lds_lu<T, MODE, 128>(128, G::dl(0), G::rin());
  lds_upper_inverse<T>(S + (NB * w) * sp + NB * w, sp, G::rin() + NB * w, G::du(w));
  lds_store<T>(M, ld, 128, S, sp);
  T* A12 = M + 128;
  T* A21 = M + (int64_t)128 * ld;
  T* A22 = M + (int64_t)128 * ld + 128;
  PP_LDS T* tile = G::ct(w);
  T fbU[4][CH][2][EPL];  // this wave's column tile of U12 as B fragments (rows 32 i ..)
  T faL[4][CH][2][EPL];  // this wave's row tile of L21 as A fragments (columns 32 j ..)
  if (w < nt2) {
    const int c0 = NB * w;  // column tile of A12 / row tile of A21
    // 2a. U12 column tile: X_i = D_Li (B_i - sum_{k<i} L_ik X_k); the
    // right-hand side tiles are loaded one ahead (fbU grows meanwhile)
    acc_t bn[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * a + P::orow(l, r), col = c0 + 16 * j + li;
          bn[a][j][r] = gload<T>(A12 + (int64_t)row * ld + (col < n2 ? col : n2 - 1));
        }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[a][j] = bn[a][j];
      if (i + 1 < 4) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = NB * (i + 1) + 16 * a + P::orow(l, r), col = c0 + 16 * j + li;
              bn[a][j][r] = gload<T>(A12 + (int64_t)row * ld + (col < n2 ? col : n2 - 1));
            }
      }
#pragma unroll
      for (int k = 0; k < i; ++k) {
        T fa[CH][2][EPL];
        lds_frags<T, MODE>(S + (NB * i) * sp + NB * k, sp, true, true, fa);  // -L_ik
        frag_mma<T, MODE>(fa, fbU[k], acc);
      }
      T fb[CH][2][EPL], fa[CH][2][EPL];
      acc_to_frags<T, MODE>(acc, tile, false, fb);
      lds_frags<T, MODE>(G::dl(i), NB, true, false, fa);
      acc_t x[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) x[a][0] = x[a][1] = P::zero();
      frag_mma<T, MODE>(fa, fb, x);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = NB * i + 16 * a + P::orow(l, r), col = c0 + 16 * j + li;
            if (col < n2) gstore<T>(A12 + (int64_t)row * ld + col, x[a][j][r]);
          }
      acc_to_frags<T, MODE>(x, tile, false, fbU[i]);
    }
    // 2b. L21 row tile: X_j = (B_j - sum_{k<j} X_k U_kj) D_Uj (fbU and the
    // growing faL live; right-hand sides one ahead)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = c0 + 16 * a + P::orow(l, r), col = 16 * h + li;
          bn[a][h][r] = gload<T>(A21 + (int64_t)(row < n2 ? row : n2 - 1) * ld + col);
        }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int h = 0; h < 2; ++h) acc[a][h] = bn[a][h];
      if (j + 1 < 4) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = c0 + 16 * a + P::orow(l, r), col = NB * (j + 1) + 16 * h + li;
              bn[a][h][r] = gload<T>(A21 + (int64_t)(row < n2 ? row : n2 - 1) * ld + col);
            }
      }
#pragma unroll
      for (int k = 0; k < j; ++k) {
        T fb[CH][2][EPL];
        lds_frags<T, MODE>(S + (NB * k) * sp + NB * j, sp, false, true, fb);  // -U_kj
        frag_mma<T, MODE>(faL[k], fb, acc);
      }
      T fa[CH][2][EPL], fb[CH][2][EPL];
      acc_to_frags<T, MODE>(acc, tile, true, fa);
      lds_frags<T, MODE>(G::du(j), NB, false, false, fb);
      acc_t x[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) x[a][0] = x[a][1] = P::zero();
      frag_mma<T, MODE>(fa, fb, x);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = c0 + 16 * a + P::orow(l, r), col = NB * j + 16 * h + li;
            if (row < n2) gstore<T>(A21 + (int64_t)row * ld + col, x[a][h][r]);
          }
      acc_to_frags<T, MODE>(x, tile, true, faL[j]);
    }
  }
  __syncthreads();
  // 3. U12 into LDS (the column tile of each wave, from its B fragments)
  if (w < nt2) {
    const int g = l >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int e = 0; e < EPL; ++e)
            G::u12()[(NB * i + c * KW + EPL * g + e) * G::UP + NB * w + 16 * h + li] = fbU[i][c][h][e];
  }
  __syncthreads();
  // row tile w of A22 - L21 U12 into the LU area (padding: identity)
  if (w < nt2) {
    const int r0 = NB * w;
    for (int ct = 0; ct < nt2; ++ct) {
      const int c0 = NB * ct;
      acc_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + 16 * a + P::orow(l, r), col = c0 + 16 * j + li;
            acc[a][j][r] =
                gload<T>(A22 + (int64_t)(row < n2 ? row : n2 - 1) * ld + (col < n2 ? col : n2 - 1));
          }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        T fb[CH][2][EPL];
        lds_frags<T, MODE>(G::u12() + (NB * k) * G::UP + c0, G::UP, false, true, fb);  // -U12
        frag_mma<T, MODE>(faL[k], fb, acc);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + 16 * a + P::orow(l, r), col = c0 + 16 * j + li;
            S[row * sp + col] = (row < n2 && col < n2) ? acc[a][j][r] : (row == col ? T(1) : T(0));
          }
    }
  }
  __syncthreads();
  // 4. the Schur complement
  lds_lu<T, MODE, 128>(np2);
  lds_store<T>(A22, ld, n2, S, sp);
}

}  // namespace pp
