// Workgroup-level dense kernels for the persistent VM (CDNA4 / gfx950).
//
// Every routine here is executed by ALL threads of one workgroup (one VM
// worker) and ends with a workgroup barrier.  Matrices are row-major with an
// explicit row pitch; op(X) = X or X^T selected by a flag.  Arithmetic:
//   * FP64 storage: v_mfma_f64_16x16x4f64;
//   * FP32 storage: v_mfma_f32_16x16x4f32, or in BF16 mode the operands are
//     rounded to bfloat16 and fed to v_mfma_f32_16x16x32_bf16 (FP32
//     accumulation).
// A wavefront owns a 32x32 output tile (2x2 MFMA tiles of 16x16); the
// workgroup's waves tile the output (or split K for small outputs).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s8_t __attribute__((ext_vector_type(8)));

enum : int { MODE_FP64 = 0, MODE_FP32 = 1, MODE_BF16 = 2 };

// LDS arena shared by all routines of one worker (bytes).
constexpr int kLdsBytes = 144 * 1024;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Round-to-nearest-even float -> bfloat16 bits.
__device__ __forceinline__ short f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// MFMA policy: K per instruction, operand elements per lane, accumulator.
template <typename T, int MODE>
struct Mma;

template <>
struct Mma<double, MODE_FP64> {
  using acc_t = d4_t;
  static constexpr int KSTEP = 4, KPL = 1;
  // accumulator r of lane l holds row (l / 16) + 4 r, column l % 16 of the tile
  __device__ static __forceinline__ int orow(int l, int r) { return (l >> 4) + 4 * r; }
  __device__ static __forceinline__ acc_t zero() { return acc_t{0., 0., 0., 0.}; }
  __device__ static __forceinline__ acc_t mma(const double* a, const double* b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
  }
};

template <>
struct Mma<float, MODE_FP32> {
  using acc_t = f4_t;
  static constexpr int KSTEP = 4, KPL = 1;
  // accumulator r of lane l holds row 4 (l / 16) + r, column l % 16 of the tile
  __device__ static __forceinline__ int orow(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ static __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static __forceinline__ acc_t mma(const float* a, const float* b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
  }
};

template <>
struct Mma<float, MODE_BF16> {
  using acc_t = f4_t;
  static constexpr int KSTEP = 32, KPL = 8;
  __device__ static __forceinline__ int orow(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ static __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static __forceinline__ acc_t mma(const float* a, const float* b, acc_t c) {
    s8_t av, bv;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      av[e] = f2bf(a[e]);
      bv[e] = f2bf(b[e]);
    }
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
  }
};

// Element (i, k) of op(A) where A is row-major with pitch ld.
template <typename T>
__device__ __forceinline__ T opget(const T* A, int ld, bool t, int i, int k) {
  return t ? A[(int64_t)k * ld + i] : A[(int64_t)i * ld + k];
}

// One wave accumulates a 32x32 tile C[r0.., c0..] over k in [k_begin, k_end)
// stepping by k_stride (multiple of KSTEP).  acc[i][j] covers rows
// r0+16i.., cols c0+16j...
template <typename T, int MODE>
__device__ __forceinline__ void wave_tile(typename Mma<T, MODE>::acc_t acc[2][2], int m, int n, int r0, int c0,
                                          int k_begin, int k_end, int k_stride, const T* A, int lda, bool aT,
                                          const T* B, int ldb, bool bT) {
  using P = Mma<T, MODE>;
  const int l = lane_id();
  const int li = l & 15, lk = (l >> 4) * P::KPL;
  const int ra0 = r0 + li, ra1 = r0 + 16 + li;
  const int cb0 = c0 + li, cb1 = c0 + 16 + li;
  const bool va0 = ra0 < m, va1 = ra1 < m, vb0 = cb0 < n, vb1 = cb1 < n;
  // Full chunks of U k-steps: every fragment load of the chunk is issued
  // before the first MFMA, so the loads overlap instead of paying one memory
  // round trip per k-step.
  constexpr int U = (MODE == MODE_BF16) ? 2 : 4;
  int kk = k_begin;
  for (; kk + (U - 1) * k_stride + P::KSTEP <= k_end; kk += U * k_stride) {
    T a0[U][P::KPL], a1[U][P::KPL], b0[U][P::KPL], b1[U][P::KPL];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < P::KPL; ++e) {
        const int kx = kk + u * k_stride + lk + e;
        a0[u][e] = va0 ? opget(A, lda, aT, ra0, kx) : T(0);
        a1[u][e] = va1 ? opget(A, lda, aT, ra1, kx) : T(0);
        b0[u][e] = vb0 ? opget(B, ldb, !bT, cb0, kx) : T(0);
        b1[u][e] = vb1 ? opget(B, ldb, !bT, cb1, kx) : T(0);
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      acc[0][0] = P::mma(a0[u], b0[u], acc[0][0]);
      acc[0][1] = P::mma(a0[u], b1[u], acc[0][1]);
      acc[1][0] = P::mma(a1[u], b0[u], acc[1][0]);
      acc[1][1] = P::mma(a1[u], b1[u], acc[1][1]);
    }
  }
  for (; kk < k_end; kk += k_stride) {
    T a0[P::KPL], a1[P::KPL], b0[P::KPL], b1[P::KPL];
    if (kk + P::KSTEP <= k_end) {
#pragma unroll
      for (int e = 0; e < P::KPL; ++e) {
        const int kx = kk + lk + e;
        a0[e] = va0 ? opget(A, lda, aT, ra0, kx) : T(0);
        a1[e] = va1 ? opget(A, lda, aT, ra1, kx) : T(0);
        b0[e] = vb0 ? opget(B, ldb, !bT, cb0, kx) : T(0);
        b1[e] = vb1 ? opget(B, ldb, !bT, cb1, kx) : T(0);
      }
    } else {
#pragma unroll
      for (int e = 0; e < P::KPL; ++e) {
        const int kx = kk + lk + e;
        const bool vk = kx < k_end;
        a0[e] = (va0 && vk) ? opget(A, lda, aT, ra0, kx) : T(0);
        a1[e] = (va1 && vk) ? opget(A, lda, aT, ra1, kx) : T(0);
        b0[e] = (vb0 && vk) ? opget(B, ldb, !bT, cb0, kx) : T(0);
        b1[e] = (vb1 && vk) ? opget(B, ldb, !bT, cb1, kx) : T(0);
      }
    }
    acc[0][0] = P::mma(a0, b0, acc[0][0]);
    acc[0][1] = P::mma(a0, b1, acc[0][1]);
    acc[1][0] = P::mma(a1, b0, acc[1][0]);
    acc[1][1] = P::mma(a1, b1, acc[1][1]);
  }
}

// C = beta * C + alpha * op(A) op(B); C m x n.  beta == 0 never reads C.
// `lds` must hold kLdsBytes.  Ends with a barrier.
template <typename T, int MODE>
__device__ void gemm_block(int m, int n, int k, T alpha, T beta, T* C, int ldc, const T* A, int lda, bool aT,
                           const T* B, int ldb, bool bT, void* lds) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id();
  if (m <= 0 || n <= 0) {
    __syncthreads();
    return;
  }
  // k == 0: C = beta * C
  const bool small = (m <= 32 && n <= 32 && k >= 8 * P::KSTEP * nw);
  if (small) {
    // split K over waves, deterministic reduction through LDS
    acc_t acc[2][2] = {{P::zero(), P::zero()}, {P::zero(), P::zero()}};
    wave_tile<T, MODE>(acc, m, n, 0, 0, w * P::KSTEP, k, nw * P::KSTEP, A, lda, aT, B, ldb, bT);
    T* red = reinterpret_cast<T*>(lds);  // [4 slots][64 lanes][16]
    const int slot = w & 3;
    for (int round = 0; round < (nw + 3) / 4; ++round) {
      if ((w >> 2) == round) {
        T* dst = red + ((size_t)slot * 64 + l) * 16;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int q = (i * 2 + j) * 4 + r;
              dst[q] = (round == 0) ? acc[i][j][r] : dst[q] + acc[i][j][r];
            }
      }
      __syncthreads();
    }
    // 1024 outputs, each the fixed-order sum of up to 4 slots
    const int slots = nw < 4 ? nw : 4;
    for (int idx = threadIdx.x; idx < 64 * 16; idx += blockDim.x) {
      const int ln = idx >> 4, q = idx & 15;
      const int i = q >> 3, j = (q >> 2) & 1, r = q & 3;
      const int row = i * 16 + P::orow(ln, r), col = j * 16 + (ln & 15);
      if (row < m && col < n) {
        T s = red[(size_t)ln * 16 + q];
        for (int sl = 1; sl < slots; ++sl) s += red[((size_t)sl * 64 + ln) * 16 + q];
        T* c = &C[(int64_t)row * ldc + col];
        *c = (beta == T(0)) ? alpha * s : beta * *c + alpha * s;
      }
    }
    __syncthreads();
    return;
  }

  // wave grid: tall outputs stack waves vertically, wide ones horizontally
  int wm, wn;
  if (n <= 32) { wm = nw; wn = 1; }
  else if (m <= 32) { wm = 1; wn = nw; }
  else if (nw >= 16) { wm = 4; wn = nw / 4; }
  else if (nw >= 4) { wm = 2; wn = nw / 2; }
  else { wm = nw; wn = 1; }
  const int tm = wm * 32, tn = wn * 32;
  const int tiles_m = (m + tm - 1) / tm, tiles_n = (n + tn - 1) / tn;
  const int my_m = w / wn, my_n = w % wn;
  for (int t = 0; t < tiles_m * tiles_n; ++t) {
    const int r0 = (t / tiles_n) * tm + my_m * 32, c0 = (t % tiles_n) * tn + my_n * 32;
    if (r0 >= m || c0 >= n) continue;
    // the old C tile is fetched before the K loop so its latency overlaps it
    T cold[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + i * 16 + P::orow(l, r), col = c0 + j * 16 + (l & 15);
          cold[i][j][r] = (beta != T(0) && row < m && col < n) ? C[(int64_t)row * ldc + col] : T(0);
        }
    acc_t acc[2][2] = {{P::zero(), P::zero()}, {P::zero(), P::zero()}};
    wave_tile<T, MODE>(acc, m, n, r0, c0, 0, k, P::KSTEP, A, lda, aT, B, ldb, bT);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + i * 16 + P::orow(l, r), col = c0 + j * 16 + (l & 15);
          if (row < m && col < n)
            C[(int64_t)row * ldc + col] = beta * cold[i][j][r] + alpha * acc[i][j][r];
        }
  }
  __syncthreads();
}

}  // namespace pp
