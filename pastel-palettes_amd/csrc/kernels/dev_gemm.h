// Workgroup-level dense kernels for the persistent VM (CDNA4 / gfx950).
//
// Every routine here is executed by ALL threads of one workgroup (one VM
// worker) and ends with a workgroup barrier.  Matrices are row-major with an
// explicit row pitch; op(X) = X or X^T selected by a flag.  Arithmetic:
//   * FP64 storage: v_mfma_f64_16x16x4f64;
//   * FP32 storage: v_mfma_f32_16x16x4f32, or in BF16 mode the operands are
//     rounded to bfloat16 in registers (v_cvt_pk_bf16_f32) and fed to
//     v_mfma_f32_16x16x32_bf16 (FP32 accumulation).
//
// GEMM organisation (gemm_block).  Every wave owns whole 32 x 32 output tiles
// (2 x 2 MFMA tiles) and runs independently: the operand fragments are loaded
// straight from global memory (the operands of the H-LU kernels are L2
// resident) into registers, one K chunk ahead of the MFMAs that consume the
// current chunk, over a flattened (tile, chunk) loop, so load latency hides
// behind the MFMAs of a chunk and there is no barrier inside the product.  The
// accumulator starts from the C tile.  When the output has fewer tiles than
// the workgroup has waves, gk waves share a tile and split its K chunks; their
// partial sums are added through LDS in a fixed order (deterministic results).
//
// K permutation: inside a wave's chunk of KW values, MFMA step s of lane group
// g = lane / 16 uses k = EPL g + s, so the EPL operand values of one lane are
// contiguous along k (one or two 16-byte loads when op(X) is stored
// k-contiguous).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s8_t __attribute__((ext_vector_type(8)));

enum : int { MODE_FP64 = 0, MODE_FP32 = 1, MODE_BF16 = 2 };

// LDS arena shared by all routines of one worker (bytes).  It is declared at
// namespace scope so that every routine addresses it as LDS (ds_* instructions)
// rather than through generic pointers.
constexpr int kLdsBytes = 152 * 1024;
static __shared__ __attribute__((aligned(16))) char pp_lds[kLdsBytes];
// Pointers into the arena carry the LDS address space explicitly (PP_LDS), so
// accesses compile to ds_* instructions even through selects and loops.
#define PP_LDS __attribute__((address_space(3)))
__device__ __forceinline__ PP_LDS char* lds_base() { return (PP_LDS char*)pp_lds; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Compile-time integer, used to unroll loops whose register buffers must be
// indexed statically.
template <int N>
struct IC {
  static constexpr int value = N;
};

// Round-to-nearest-even float -> bfloat16 bits.
__device__ __forceinline__ short f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// Global loads / stores through the global address space (no flat-address
// coupling with the LDS counter).
template <typename V>
__device__ __forceinline__ V gload(const void* p) {
  typedef const __attribute__((address_space(1))) V* gptr_t;
  return *(gptr_t)(p);
}
template <typename V>
__device__ __forceinline__ void gstore(void* p, V v) {
  typedef __attribute__((address_space(1))) V* gptr_t;
  *(gptr_t)(p) = v;
}

// MFMA policy.  KSTEP / KPL / mma describe one instruction (lane l supplies
// A[l % 16][KPL (l / 16) + e] and B[KPL (l / 16) + e][l % 16]); KW / EPL /
// mma_chunk describe the K chunk a wave consumes per fragment load (EPL
// contiguous k values per lane).
template <typename T, int MODE>
struct Mma;

template <>
struct Mma<double, MODE_FP64> {
  using acc_t = d4_t;
  static constexpr int KSTEP = 4, KPL = 1, KW = 16, EPL = 4;
  // accumulator r of lane l holds row (l / 16) + 4 r, column l % 16 of the tile
  __device__ static __forceinline__ int orow(int l, int r) { return (l >> 4) + 4 * r; }
  __device__ static __forceinline__ acc_t zero() { return acc_t{0., 0., 0., 0.}; }
  __device__ static __forceinline__ acc_t mma(const double* a, const double* b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
  }
  __device__ static __forceinline__ acc_t mma_chunk(const double (&a)[EPL], const double (&b)[EPL], acc_t c) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[2], b[2], c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a[3], b[3], c, 0, 0, 0);
  }
};

template <>
struct Mma<float, MODE_FP32> {
  using acc_t = f4_t;
  static constexpr int KSTEP = 4, KPL = 1, KW = 16, EPL = 4;
  // accumulator r of lane l holds row 4 (l / 16) + r, column l % 16 of the tile
  __device__ static __forceinline__ int orow(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ static __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static __forceinline__ acc_t mma(const float* a, const float* b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
  }
  __device__ static __forceinline__ acc_t mma_chunk(const float (&a)[EPL], const float (&b)[EPL], acc_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], c, 0, 0, 0);
  }
};

template <>
struct Mma<float, MODE_BF16> {
  using acc_t = f4_t;
  static constexpr int KSTEP = 32, KPL = 8, KW = 32, EPL = 8;
  __device__ static __forceinline__ int orow(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ static __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static __forceinline__ s8_t to_bf16(const float* a) {
    s8_t v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = __builtin_bit_cast(short, (__bf16)a[e]);  // RNE
    return v;
  }
  __device__ static __forceinline__ acc_t mma(const float* a, const float* b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(to_bf16(a), to_bf16(b), c, 0, 0, 0);
  }
  __device__ static __forceinline__ acc_t mma_chunk(const float (&a)[EPL], const float (&b)[EPL], acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(to_bf16(a), to_bf16(b), c, 0, 0, 0);
  }
};

// ---------------------------------------------------------------- operand fragments
// One lane's view of a 16-wide operand group: lane (g = l / 16, i = l % 16)
// reads op(X)(x0 + i, k0 + EPL g + e), e < EPL, where op(X)(r, k) is
// X[r ld + k] (along_k) or X[k ld + r].  Loads are never predicated and their
// values are not inspected until the MFMAs that use them (branch-free loads
// keep the compiler's memory-counter waits exact, so the prefetched chunk
// stays in flight while the current one is multiplied):
//   * rows outside op(X) read row 0; they only feed output rows / columns
//     that are never stored;
//   * k at or beyond kmax reads k = kmax - 1 (finite data); the A operand of
//     such k is zeroed just before its MFMA, which cancels the product.
template <typename T, int EPL>
struct OperandLane {
  const T* p;    // &op(X)(clamped x0 + i, 0)
  int64_t step;  // element distance between consecutive k

  __device__ __forceinline__ void init(const T* X, int ld, bool along_k, int x, int xmax) {
    const int xs = x < xmax ? x : 0;
    step = along_k ? 1 : ld;
    p = along_k ? X + (int64_t)xs * ld : X + xs;
  }

  // VEC: op(X) is k-contiguous with 16-byte aligned rows and k is a whole
  // number of chunks.  Otherwise element loads with the k clamp.
  template <bool VEC>
  __device__ __forceinline__ void load(T (&f)[EPL], int kg, int kmax) const {
    if constexpr (VEC) {
      constexpr int V = 16 / sizeof(T);
      typedef T vec_t __attribute__((ext_vector_type(V)));
#pragma unroll
      for (int h = 0; h < EPL / V; ++h) {
        const vec_t v = gload<vec_t>(p + kg + h * V);
#pragma unroll
        for (int e = 0; e < V; ++e) f[h * V + e] = v[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int kk = kg + e;
        f[e] = gload<T>(p + (int64_t)(kk < kmax ? kk : kmax - 1) * step);
      }
    }
  }
};

// ---------------------------------------------------------------- gemm
// Per-wave tile loop of gemm_block for one operand-load variant.
template <typename T, int MODE, bool AV, bool BV>
__device__ __forceinline__ void gemm_tiles(int m, int n, int k, T alpha, T beta, T* C, int ldc, const T* A, int lda,
                                           bool a_along, const T* B, int ldb, bool b_along, int tiles_n, int grp,
                                           int groups, int my_tiles, int ks, int gk, int my_chunks,
                                           typename Mma<T, MODE>::acc_t (&acc)[2][2], int& r0, int& c0) {
  using P = Mma<T, MODE>;
  constexpr int KW = P::KW, EPL = P::EPL;
  const int l = lane_id(), g = l >> 4, li = l & 15;
  T fa[2][2][EPL], fb[2][2][EPL];  // [buffer][16-wide group][element]
  int kgb[2] = {0, 0};              // first k of this lane in each buffer
  OperandLane<T, EPL> la[2], lb[2];
  // alpha is +-1 at every call site: C_new = alpha (alpha beta C + A B), so
  // the accumulator starts at alpha beta C and the store multiplies by alpha
  const T ab = alpha * beta;
  auto setup_tile = [&](int t) {
    r0 = (t / tiles_n) * 32;
    c0 = (t % tiles_n) * 32;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      la[h].init(A, lda, a_along, r0 + 16 * h + li, m);
      lb[h].init(B, ldb, b_along, c0 + 16 * h + li, n);
    }
  };
  auto load_chunk = [&](auto buf_c, int c) {
    constexpr int buf = decltype(buf_c)::value;
    const int kg = c * KW + EPL * g;
    kgb[buf] = kg;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      la[h].template load<AV>(fa[buf][h], kg, k);
      lb[h].template load<BV>(fb[buf][h], kg, k);
    }
  };
  // every active wave has at least one chunk (k > 0 and gk <= chunks)
  const int chunks_eff = my_chunks;
  const int total = my_tiles * chunks_eff;
  if (total == 0) return;
  int ti = 0, ci = 0;
  setup_tile(grp);
  load_chunk(IC<0>{}, ks);
  auto step = [&](auto cur_c) {
    constexpr int cur = decltype(cur_c)::value;
    if (ci == 0) {
      if (ks == 0 && beta != T(0)) {
        // the whole C tile is requested before any value is used (one
        // memory round trip per tile); addresses are clamped as in
        // OperandLane.  With beta == 0 (or for the other K slices) C is
        // not read at all.
        T cv[2][2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = r0 + i * 16 + P::orow(l, r), col = c0 + j * 16 + li;
              cv[i][j][r] = gload<T>(C + (int64_t)(row < m ? row : m - 1) * ldc + (col < n ? col : n - 1));
            }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = r0 + i * 16 + P::orow(l, r), col = c0 + j * 16 + li;
              acc[i][j][r] = (row < m && col < n) ? ab * cv[i][j][r] : T(0);
            }
      } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = P::zero();
      }
    }
    int nti = ti, nci = ci + 1;
    if (nci == chunks_eff) {
      nci = 0;
      ++nti;
    }
    const int cr0 = r0, cc0 = c0;
    // prefetch the next (tile, chunk); past the last one, reload the last
    // chunk so that every step issues the same loads
    if (nti < my_tiles && nti != ti) setup_tile(grp + nti * groups);
    load_chunk(IC<cur ^ 1>{}, ks + (nti < my_tiles ? nci : ci) * gk);
    // every prefetch load is issued before the MFMAs of the current chunk
    __builtin_amdgcn_sched_barrier(0);
    if (!(AV && BV)) {  // k tail: cancel k >= kmax through A
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < EPL; ++e) fa[cur][h][e] = kgb[cur] + e < k ? fa[cur][h][e] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = P::mma_chunk(fa[cur][i], fb[cur][j], acc[i][j]);
    if (ci == chunks_eff - 1 && gk == 1) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = cr0 + i * 16 + P::orow(l, r), col = cc0 + j * 16 + li;
            if (row < m && col < n) gstore<T>(C + (int64_t)row * ldc + col, alpha * acc[i][j][r]);
          }
    }
    ti = nti;
    ci = nci;
  };
  for (int it = 0; it < total; it += 2) {
    step(IC<0>{});
    if (it + 1 < total) step(IC<1>{});
  }
}

// ---------------------------------------------------------------- staged gemm
// Updates whose operands are both k-contiguous and whose output tiles are all
// whole (m, n multiples of 32, at least one tile per wave).  op(A) (m x k)
// and op(B)^T (n x k) are copied into LDS once by LDS DMA
// (global_load_lds_dwordx4: 64 lanes x 16 bytes per instruction, every copy
// of the workgroup in flight at once), so the fragments of every output tile
// come from LDS instead of a global round trip per tile; the C tiles are
// requested through a ring of register buffers RING tiles ahead.  Chunk c of
// staged row r holds the row's 16-byte chunk c ^ (r mod cpr), so the rows
// read by one fragment load spread over the banks.
// FP32 storage forms the transposed tile (op(B)^T op(A)^T on the MFMA): each
// lane's four accumulators are then four consecutive columns of C, which
// moves as 16-byte vectors.  FP64 keeps the product's orientation (sixteen
// lanes cover a 128-byte row segment).  Every element sees the same MFMA
// sequence as in gemm_tiles (C scaled first, K chunks in order), so results
// are identical.
constexpr int kStagedBytes = 144 * 1024;  // the failure flag of accum keeps the last 16 bytes

template <typename T, int MODE>
__device__ __forceinline__ bool gemm_staged_ok(int m, int n, int k, const T* A, int lda, bool a_along, const T* B,
                                               int ldb, bool b_along, const T* C, int ldc, int nw) {
  constexpr int KW = Mma<T, MODE>::KW, ES = (int)sizeof(T);
  const int cpr = k * ES / 16;  // 16-byte chunks per staged row
  const int64_t bytes = (int64_t)(m + n) * k * ES;
  return a_along && b_along && m % 32 == 0 && n % 32 == 0 && (m >> 5) * (n >> 5) >= nw && k % KW == 0 &&
         cpr >= 2 && (cpr & (cpr - 1)) == 0 && bytes <= kStagedBytes && bytes % 1024 == 0 &&
         (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0 &&
         (lda * ES) % 16 == 0 && (ldb * ES) % 16 == 0 &&
         (ES == 8 || ((reinterpret_cast<uintptr_t>(C) & 15) == 0 && ldc % 4 == 0));
}

// A_DMA false: rows 0 .. m-1 of the staged area (op(A)) are written by
// fill_a(sa) (in the swizzled layout, after the copies of op(B)^T are
// issued, any barriers inside it reached by every thread) instead of copied
// from A.
template <typename T, int MODE, bool BETA, bool A_DMA, typename FillA>
__device__ __forceinline__ void gemm_staged_run(int m, int n, int k, T alpha, T beta, T* C, int ldc, const T* A,
                                                int lda, const T* B, int ldb, FillA&& fill_a) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  constexpr bool TR = sizeof(T) == 4;
  constexpr int KW = P::KW, EPL = P::EPL, V = 16 / (int)sizeof(T), RING = 2;
  typedef T vec_t __attribute__((ext_vector_type(V)));
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id(), g = l >> 4, li = l & 15;
  const int tiles_n = n >> 5, ntiles = (m >> 5) * tiles_n;
  const int my = (ntiles - w + nw - 1) / nw;  // >= 1: every wave has a tile
  const int cpr = k / V, lg = __builtin_ctz(cpr), nch = k / KW;
  const T ab = alpha * beta;
  PP_LDS T* sa = (PP_LDS T*)lds_base();  // op(A) rows 0 .. m-1, op(B)^T rows m .. m+n-1, pitch k
  acc_t cb[RING][2][2];
  auto tile_rc = [&](int q, int& r0, int& c0) {
    const int t = w + (q < my ? q : my - 1) * nw, tr = t / tiles_n;
    r0 = tr * 32;
    c0 = (t - tr * tiles_n) * 32;
  };
  // element r of accumulator (i, j): FP32 C[r0 + 16 j + li][c0 + 16 i + 4 g + r],
  // FP64 C[r0 + 16 i + orow(l, r)][c0 + 16 j + li]
  auto cptr = [&](int r0, int c0, int i, int j, int r) -> T* {
    if constexpr (TR) return C + (int64_t)(r0 + 16 * j + li) * ldc + c0 + 16 * i + 4 * g + r;
    else return C + (int64_t)(r0 + 16 * i + P::orow(l, r)) * ldc + c0 + 16 * j + li;
  };
  auto load_c = [&](auto s_c, int q) {
    constexpr int s = decltype(s_c)::value;
    int r0, c0;
    tile_rc(q, r0, c0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if constexpr (TR) {
          cb[s][i][j] = gload<f4_t>(cptr(r0, c0, i, j, 0));
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) cb[s][i][j][r] = gload<T>(cptr(r0, c0, i, j, r));
        }
      }
  };
  // the first C tiles are requested before the copies; the copies are issued
  // after them, so once the copies have landed (vmcnt 0) so have these
  if constexpr (BETA) {
    load_c(IC<0>{}, 0);
    load_c(IC<1>{}, 1);
  }
  {
    const int i0 = A_DMA ? 0 : m * cpr / 64, ninst = (m + n) * cpr / 64;
    for (int i = i0 + w; i < ninst; i += nw) {
      const int q = 64 * i + l, row = q >> lg, pc = q & (cpr - 1), lc = pc ^ (row & (cpr - 1));
      const T* src = row < m ? A + (int64_t)row * lda + V * lc : B + (int64_t)(row - m) * ldb + V * lc;
      __builtin_amdgcn_global_load_lds((const void*)src, (PP_LDS void*)(sa + 64 * V * i), 16, 0, 0);
    }
  }
  if constexpr (!A_DMA) fill_a(sa);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  auto body = [&](auto s_c, int q) {
    constexpr int s = decltype(s_c)::value;
    int r0, c0;
    tile_rc(q, r0, c0);
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = BETA ? ab * cb[s][i][j] : P::zero();
    if constexpr (BETA) load_c(IC<s>{}, q + RING);
    for (int c = 0; c < nch; ++c) {
      T fa[2][EPL], fb[2][EPL];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int ra = r0 + 16 * h + li, rb = m + c0 + 16 * h + li;
#pragma unroll
        for (int v = 0; v < EPL / V; ++v) {
          const int ch = (c * KW + EPL * g) / V + v;
          const vec_t x = *(const PP_LDS vec_t*)(sa + ra * k + V * (ch ^ (ra & (cpr - 1))));
          const vec_t y = *(const PP_LDS vec_t*)(sa + rb * k + V * (ch ^ (rb & (cpr - 1))));
#pragma unroll
          for (int e = 0; e < V; ++e) {
            fa[h][V * v + e] = x[e];
            fb[h][V * v + e] = y[e];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if constexpr (TR) acc[i][j] = P::mma_chunk(fb[i], fa[j], acc[i][j]);
          else acc[i][j] = P::mma_chunk(fa[i], fb[j], acc[i][j]);
        }
    }
    if (q < my) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if constexpr (TR) {
            gstore<f4_t>(cptr(r0, c0, i, j, 0), alpha * acc[i][j]);
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) gstore<T>(cptr(r0, c0, i, j, r), alpha * acc[i][j][r]);
          }
        }
    }
  };
  if constexpr (MODE == MODE_FP32) {
    if (nch == 2) {
      // FP32 MFMA with k = 32 (two chunks): the next tile's fragments are
      // read from LDS before the current tile's MFMAs (the FP32 products are
      // long enough to hide the LDS latency behind them)
      T fr[2][2][2][2][EPL];  // [buffer][chunk][a/b][group][element]
      auto load_fr = [&](auto b_c, int q) {
        constexpr int bb = decltype(b_c)::value;
        int r0, c0;
        tile_rc(q, r0, c0);
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int ra = r0 + 16 * h + li, rb = m + c0 + 16 * h + li;
#pragma unroll
            for (int v = 0; v < EPL / V; ++v) {
              const int ch = (c * KW + EPL * g) / V + v;
              const vec_t x = *(const PP_LDS vec_t*)(sa + ra * k + V * (ch ^ (ra & (cpr - 1))));
              const vec_t y = *(const PP_LDS vec_t*)(sa + rb * k + V * (ch ^ (rb & (cpr - 1))));
#pragma unroll
              for (int e = 0; e < V; ++e) {
                fr[bb][c][0][h][V * v + e] = x[e];
                fr[bb][c][1][h][V * v + e] = y[e];
              }
            }
          }
      };
      auto pbody = [&](auto s_c, int q) {
        constexpr int s = decltype(s_c)::value;
        constexpr int bb = s & 1;
        int r0, c0;
        tile_rc(q, r0, c0);
        load_fr(IC<bb ^ 1>{}, q + 1);
        acc_t acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = BETA ? ab * cb[s][i][j] : P::zero();
        if constexpr (BETA) load_c(IC<s>{}, q + RING);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = P::mma_chunk(fr[bb][c][1][i], fr[bb][c][0][j], acc[i][j]);
        if (q < my) {
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) gstore<f4_t>(cptr(r0, c0, i, j, 0), alpha * acc[i][j]);
        }
      };
      load_fr(IC<0>{}, 0);
      for (int q0 = 0; q0 < my; q0 += RING) {
        pbody(IC<0>{}, q0);
        pbody(IC<1>{}, q0 + 1);
      }
      __syncthreads();
      return;
    }
  }
  for (int q0 = 0; q0 < my; q0 += RING) {
    body(IC<0>{}, q0);
    body(IC<1>{}, q0 + 1);
  }
  __syncthreads();
}

template <typename T, int MODE, bool BETA>
__device__ __forceinline__ void gemm_staged(int m, int n, int k, T alpha, T beta, T* C, int ldc, const T* A, int lda,
                                            const T* B, int ldb) {
  gemm_staged_run<T, MODE, BETA, true>(m, n, k, alpha, beta, C, ldc, A, lda, B, ldb, [](PP_LDS T*) {});
}

// C = beta * C + alpha * op(A) op(B); C m x n, op(A) m x k, op(B) k x n;
// alpha must be +1 or -1.  With beta == 0 the old C is not read.  Uses LDS
// for the staged operands (gemm_staged) or the split-K reduction.  Ends with
// a barrier.
template <typename T, int MODE>
__device__ __forceinline__ void gemm_block(int m, int n, int k, T alpha, T beta, T* C, int ldc, const T* A, int lda,
                                           bool aT, const T* B, int ldb, bool bT) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  constexpr int KW = P::KW, V = 16 / sizeof(T);
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id();
  const int li = l & 15;
  if (m <= 0 || n <= 0) {
    __syncthreads();
    return;
  }
  const int tiles_n = (n + 31) >> 5, ntiles = ((m + 31) >> 5) * tiles_n;
  if (k <= 0) {  // C = beta C
    for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
      const int i = e / n, j = e - i * n;
      T* c = C + (int64_t)i * ldc + j;
      gstore<T>(c, beta == T(0) ? T(0) : beta * gload<T>(c));
    }
    __syncthreads();
    return;
  }
  const int nch = (k + KW - 1) / KW;
  int gk = 1;
  while (2 * gk * ntiles <= nw && 2 * gk <= nch) gk *= 2;
  const int groups = nw / gk, grp = w / gk, ks = w - grp * gk;
  // this wave: tiles grp, grp + groups, ...; chunks ks, ks + gk, ...
  const int my_tiles = grp < ntiles ? (ntiles - grp + groups - 1) / groups : 0;
  const int my_chunks = ks < nch ? (nch - ks + gk - 1) / gk : 0;
  const bool a_along = !aT, b_along = bT, full = (k % KW) == 0;
  if (gemm_staged_ok<T, MODE>(m, n, k, A, lda, a_along, B, ldb, b_along, C, ldc, nw)) {
    if (beta != T(0)) gemm_staged<T, MODE, true>(m, n, k, alpha, beta, C, ldc, A, lda, B, ldb);
    else gemm_staged<T, MODE, false>(m, n, k, alpha, beta, C, ldc, A, lda, B, ldb);
    return;
  }
  const bool a_vec = full && a_along && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % V) == 0;
  const bool b_vec = full && b_along && ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (ldb % V) == 0;
  acc_t acc[2][2];
  int r0 = 0, c0 = 0;
#define PP_GEMM_TILES(AV, BV)                                                                                      \
  gemm_tiles<T, MODE, AV, BV>(m, n, k, alpha, beta, C, ldc, A, lda, a_along, B, ldb, b_along, tiles_n, grp, groups, \
                              my_tiles, ks, gk, my_chunks, acc, r0, c0)
  if (a_vec) {
    if (b_vec) PP_GEMM_TILES(true, true);
    else PP_GEMM_TILES(true, false);
  } else {
    if (b_vec) PP_GEMM_TILES(false, true);
    else PP_GEMM_TILES(false, false);
  }
#undef PP_GEMM_TILES
  if (gk > 1) {
    // each wave holds at most one tile: partial sums of its K slice, added
    // in wave order by the first wave of the group
    PP_LDS T* red = (PP_LDS T*)lds_base();
    if (my_tiles > 0) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[((size_t)w * 64 + l) * 16 + (i * 2 + j) * 4 + r] = acc[i][j][r];
    }
    __syncthreads();
    if (my_tiles > 0 && ks == 0) {
      for (int q = 1; q < gk; ++q) {
        const PP_LDS T* src = red + ((size_t)(w + q) * 64 + l) * 16;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] += src[(i * 2 + j) * 4 + r];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + i * 16 + P::orow(l, r), col = c0 + j * 16 + li;
            if (row < m && col < n) gstore<T>(C + (int64_t)row * ldc + col, alpha * acc[i][j][r]);
          }
    }
  }
  __syncthreads();
}

}  // namespace pp
