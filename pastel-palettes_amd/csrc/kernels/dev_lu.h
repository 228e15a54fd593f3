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
//   * getrf: right-looking.  The 32-column panel is eliminated with one row
//     per thread (pivot row published through double-buffered LDS, one
//     barrier per column); the same row operations applied to the identity
//     give L11^{-1} for free.  U12 = L11^{-1} A12 and the trailing update
//     A22 -= L21 U12 run on MFMA through gemm_block.
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
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      T acc = (i == c) ? T(1) : T(0);
#pragma unroll
      for (int j = i + 1; j < NB; ++j) acc -= s[i * (NB + 1) + j] * y[j];
      asm volatile("" : "+v"(acc)::"memory");
      y[i] = (i < nb) ? acc / s[i * (NB + 1) + i] : T(0);
    }
  }
}

// Inverses of the NB x NB diagonal blocks of the n x n triangular matrix F
// (pitch ldf); block b is written row-major (pitch NB) at out + b NB NB.
// One wave per block; uses NB (NB + 1) elements of LDS per wave.
template <typename T, bool UPPER>
__device__ __forceinline__ void diag_inverses(const T* F, int ldf, int n, T* out) {
  const int nw = blockDim.x >> 6, w = wave_id(), l = lane_id();
  const int nblk = (n + NB - 1) / NB;
  PP_LDS T* s = (PP_LDS T*)lds_base() + w * NB * (NB + 1);
  for (int b0 = 0; b0 < nblk; b0 += nw) {
    const int b = b0 + w;
    const int r0 = b * NB;
    const int nb = b < nblk ? ((n - r0) < NB ? (n - r0) : NB) : 0;
    if (b < nblk) {
      for (int e = l; e < NB * NB; e += 64) {
        const int i = e / NB, j = e - i * NB;
        s[i * (NB + 1) + j] = (i < nb && j < nb) ? F[(int64_t)(r0 + i) * ldf + r0 + j] : T(0);
      }
    }
    __syncthreads();
    if (b < nblk && l < nb) {
      T y[NB];
      tri_inverse_column<T, UPPER>(s, nb, l, y);
      T* o = out + (int64_t)b * NB * NB;
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < nb) o[i * NB + l] = y[i];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- getrf
// Unblocked LU of the panel M[p:ny, p:p+pw] (pw <= NB) in place.  Thread t
// owns row p + t; the pivot row of step j is published by its owner into one
// of two LDS buffers, so one barrier per column suffices (the buffer written
// at step j + 2 was last read before every thread passed barrier j + 1).  The
// top pw rows also carry the identity through the elimination, which leaves
// L11^{-1} there; it is written row-major (pitch NB) to linv.  Rows beyond
// blockDim.x are eliminated afterwards against the final pivot rows.
template <typename T>
__device__ __forceinline__ void panel_lu(T* M, int ld, int p, int pw, int ny, T* linv) {
  PP_LDS T* piv = (PP_LDS T*)lds_base();  // [2][2 NB]
  const int tid = threadIdx.x, nt = blockDim.x;
  const int R = ny - p;
  const bool own = tid < R;
  T* row = M + (int64_t)(p + tid) * ld + p;
  T x[NB], e[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    x[c] = (own && c < pw) ? row[c] : T(0);
    e[c] = (c == tid) ? T(1) : T(0);
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < pw) {
      PP_LDS T* pv = piv + (j & 1) * 2 * NB;
      if (tid == j) {
#pragma unroll
        for (int c = j; c < NB; ++c) pv[c] = x[c];
#pragma unroll
        for (int c = 0; c <= j; ++c) pv[NB + c] = e[c];
      }
      __syncthreads();
      if (own && tid > j) {
        const T lv = x[j] / pv[j];
        x[j] = lv;
#pragma unroll
        for (int c = j + 1; c < NB; ++c) x[c] -= lv * pv[c];
        if (tid < pw) {
#pragma unroll
          for (int c = 0; c <= j; ++c) e[c] -= lv * pv[NB + c];
        }
      }
    }
  }
  if (own) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < pw) row[c] = x[c];
  }
  if (tid < pw) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < pw) linv[tid * NB + c] = e[c];
  }
  __syncthreads();
  for (int r = nt + tid; r < R; r += nt) {
    T* rr = M + (int64_t)(p + r) * ld + p;
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = c < pw ? rr[c] : T(0);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < pw) {
        const T* pr = M + (int64_t)(p + j) * ld + p;
        const T lv = x[j] / pr[j];
        x[j] = lv;
#pragma unroll
        for (int c = j + 1; c < NB; ++c)
          if (c < pw) x[c] -= lv * pr[c];
      }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < pw) rr[c] = x[c];
  }
  __syncthreads();
}

// ---------------------------------------------------------------- Cholesky
// Control word in the LDS arena (past every routine's working area): set to 1
// by chol_upper when the factorisation failed even with the shift.
__device__ __forceinline__ volatile PP_LDS int* chol_fail_flag() {
  return (volatile PP_LDS int*)(lds_base() + kLdsBytes - 16);
}

// G (k x k, pitch k, symmetric positive semi-definite; k <= 64) <- R, the
// upper Cholesky factor of G.  first: the first factorisation of an accum;
// when its plain factorisation meets a non-positive pivot, G + s I is
// factorised instead, s = 11 (rows k + k (k + 1)) u tr(G) (the shift of
// shifted CholeskyQR3, Fukaya et al., SIAM J. Sci. Comput. 42 (2020)); the
// failure flag is reset there.  Later factorisations are never shifted: a
// failure sets the flag (the caller then orthonormalises by Householder).
// Entries below the diagonal are set to zero.
template <typename T>
__device__ __forceinline__ void chol_upper(T* G, int k, int rows, bool first) {
  PP_LDS T* s = (PP_LDS T*)lds_base();  // k x (k + 1)
  const int sp = k + 1, tid = threadIdx.x, nt = blockDim.x;
  const T u = sizeof(T) == 8 ? T(1.1102230246251565e-16) : T(5.9604644775390625e-08);
  bool ok = false;
  for (int attempt = 0; attempt < (first ? 2 : 1) && !ok; ++attempt) {
    for (int e = tid; e < k * k; e += nt) {
      const int i = e / k, j = e - i * k;
      s[i * sp + j] = G[e];
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
    ok = true;
    for (int j = 0; j < k; ++j) {
      const T d = s[j * sp + j];  // every thread reads the same pivot
      if (!(d > T(0)) || !(d < T(__builtin_inf()))) {
        ok = false;
        break;
      }
      const T r = sqrt(d), ri = T(1) / r;
      __syncthreads();
      for (int c = j + tid; c < k; c += nt) s[j * sp + c] = c == j ? r : s[j * sp + c] * ri;
      __syncthreads();
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
