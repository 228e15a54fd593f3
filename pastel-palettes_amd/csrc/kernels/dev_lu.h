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
// Blocking is by panels of PB columns; the panel's updates run on MFMA via
// gemm_block.  All routines end with a workgroup barrier.
#pragma once

#include "dev_gemm.h"
#include "dev_leaf.h"

namespace pp {

constexpr int PB = 16;  // panel width

// ---------------------------------------------------------------- getrf
// M is ny x nx, pitch ld; requires ny <= 2 * blockDim.x.
template <typename T, int FMODE>
__device__ void getrf_block(T* M, int nx, int ny, int ld, void* lds) {
  if (getrf_lds<T>(M, nx, ny, ld, lds, kLdsBytes)) return;  // whole block fits in LDS
  T* piv = reinterpret_cast<T*>(lds);            // [PB] current pivot row
  T* l11 = piv + PB;                             // [PB][PB] diagonal block
  void* glds = reinterpret_cast<char*>(lds) + 8 * 1024;  // gemm scratch after 8 KB
  const int tid = threadIdx.x, nt = blockDim.x;
  const int mn = nx < ny ? nx : ny;
  for (int p = 0; p < mn; p += PB) {
    const int pw = (mn - p) < PB ? (mn - p) : PB;
    // ---- panel M[p:ny, p:p+pw]: thread owns rows p + tid and p + tid + nt
    T row0[PB], row1[PB];
    const int i0 = p + tid, i1 = p + tid + nt;
    const bool own0 = i0 < ny, own1 = i1 < ny;
#pragma unroll
    for (int j = 0; j < PB; ++j) {
      row0[j] = (own0 && j < pw) ? M[(int64_t)i0 * ld + p + j] : T(0);
      row1[j] = (own1 && j < pw) ? M[(int64_t)i1 * ld + p + j] : T(0);
    }
    // fully unrolled so that row0/row1 stay in registers
#pragma unroll
    for (int j = 0; j < PB; ++j) {
      if (j < pw) {
        // owner of pivot row p + j publishes it (rows < nt + p are in row0)
        if (tid == j) {
#pragma unroll
          for (int c = 0; c < PB; ++c) piv[c] = row0[c];
        }
        __syncthreads();
        const T d = piv[j];
        if (own0 && i0 > p + j) {
          const T lv = row0[j] / d;
          row0[j] = lv;
#pragma unroll
          for (int c = j + 1; c < PB; ++c) row0[c] -= lv * piv[c];
        }
        if (own1) {
          const T lv = row1[j] / d;
          row1[j] = lv;
#pragma unroll
          for (int c = j + 1; c < PB; ++c) row1[c] -= lv * piv[c];
        }
        __syncthreads();
      }
    }
#pragma unroll
    for (int j = 0; j < PB; ++j) {
      if (j < pw) {
        if (own0) M[(int64_t)i0 * ld + p + j] = row0[j];
        if (own1) M[(int64_t)i1 * ld + p + j] = row1[j];
      }
    }
    if (tid < pw) {
#pragma unroll
      for (int c = 0; c < PB; ++c) l11[tid * PB + c] = row0[c];
    }
    __syncthreads();
    // ---- U12 = L11^{-1} M[p:p+pw, p+pw:nx]
    for (int c = p + pw + tid; c < nx; c += nt) {
      T x[PB];
#pragma unroll
      for (int i = 0; i < PB; ++i) x[i] = (i < pw) ? M[(int64_t)(p + i) * ld + c] : T(0);
#pragma unroll
      for (int i = 0; i < PB; ++i)
#pragma unroll
        for (int j = 0; j < PB; ++j)
          if (j < i) x[i] -= l11[i * PB + j] * x[j];
#pragma unroll
      for (int i = 0; i < PB; ++i)
        if (i < pw) M[(int64_t)(p + i) * ld + c] = x[i];
    }
    __syncthreads();
    // ---- trailing update
    const int rm = ny - p - pw, rn = nx - p - pw;
    if (rm > 0 && rn > 0)
      gemm_block<T, FMODE>(rm, rn, pw, T(-1), T(1), &M[(int64_t)(p + pw) * ld + p + pw], ld,
                           &M[(int64_t)(p + pw) * ld + p], ld, false, &M[(int64_t)p * ld + p + pw], ld, false, glds);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- triangular solves
// Stage the PB x PB block T[r0.., c0..] (pitch ld) into lds (row-major PB).
template <typename T>
__device__ __forceinline__ void stage_block(T* dst, const T* src, int ld, int r0, int c0, int pr, int pc) {
  for (int e = threadIdx.x; e < PB * PB; e += blockDim.x) {
    const int i = e / PB, j = e % PB;
    dst[e] = (i < pr && j < pc) ? src[(int64_t)(r0 + i) * ld + c0 + j] : T(0);
  }
}

// B (ny_b x nx_b logical) <- L^{-1} B, L unit lower ny_b x ny_b.
template <typename T, int FMODE>
__device__ void trsml_block(T* B, const T* L, int nx_b, int ny_b, int ld_b, int ld_l, bool b_T, void* lds) {
  if (!b_T && trsm_rows<T, 0>(B, L, ny_b, nx_b, ld_b, ld_l, lds, kLdsBytes)) return;
  T* lb = reinterpret_cast<T*>(lds);
  void* glds = reinterpret_cast<char*>(lds) + 8 * 1024;
  for (int p = 0; p < ny_b; p += PB) {
    const int pw = (ny_b - p) < PB ? (ny_b - p) : PB;
    stage_block(lb, L, ld_l, p, p, pw, pw);
    __syncthreads();
    // columns of the logical B are independent
    for (int c = threadIdx.x; c < nx_b; c += blockDim.x) {
      T x[PB];
#pragma unroll
      for (int i = 0; i < PB; ++i)
        x[i] = (i < pw) ? (b_T ? B[(int64_t)c * ld_b + p + i] : B[(int64_t)(p + i) * ld_b + c]) : T(0);
#pragma unroll
      for (int i = 0; i < PB; ++i)
#pragma unroll
        for (int j = 0; j < PB; ++j)
          if (j < i) x[i] -= lb[i * PB + j] * x[j];
#pragma unroll
      for (int i = 0; i < PB; ++i)
        if (i < pw) {
          if (b_T) B[(int64_t)c * ld_b + p + i] = x[i];
          else B[(int64_t)(p + i) * ld_b + c] = x[i];
        }
    }
    __syncthreads();
    const int rm = ny_b - p - pw;
    if (rm > 0) {
      if (!b_T)  // B[p+pw:, :] -= L[p+pw:, p:p+pw] B[p:p+pw, :]
        gemm_block<T, FMODE>(rm, nx_b, pw, T(-1), T(1), &B[(int64_t)(p + pw) * ld_b], ld_b,
                             &L[(int64_t)(p + pw) * ld_l + p], ld_l, false, &B[(int64_t)p * ld_b], ld_b, false, glds);
      else  // Bs[:, p+pw:] -= Bs[:, p:p+pw] L[p+pw:, p:p+pw]^T
        gemm_block<T, FMODE>(nx_b, rm, pw, T(-1), T(1), &B[p + pw], ld_b, &B[p], ld_b, false,
                             &L[(int64_t)(p + pw) * ld_l + p], ld_l, true, glds);
    }
  }
  __syncthreads();
}

// B (ny_b x nx_b logical) <- B U^{-1}, U upper non-unit nx_b x nx_b.
template <typename T, int FMODE>
__device__ void trsmr_block(T* B, const T* U, int nx_b, int ny_b, int ld_b, int ld_u, bool b_T, void* lds) {
  // transposed B (the V factor of a low-rank block): U^T Bs_new = Bs
  if (b_T && trsm_rows<T, 1>(B, U, nx_b, ny_b, ld_b, ld_u, lds, kLdsBytes)) return;
  T* ub = reinterpret_cast<T*>(lds);
  void* glds = reinterpret_cast<char*>(lds) + 8 * 1024;
  for (int p = 0; p < nx_b; p += PB) {
    const int pw = (nx_b - p) < PB ? (nx_b - p) : PB;
    stage_block(ub, U, ld_u, p, p, pw, pw);
    __syncthreads();
    // rows of the logical B are independent: x U11 = b
    for (int r = threadIdx.x; r < ny_b; r += blockDim.x) {
      T x[PB];
#pragma unroll
      for (int j = 0; j < PB; ++j)
        x[j] = (j < pw) ? (b_T ? B[(int64_t)(p + j) * ld_b + r] : B[(int64_t)r * ld_b + p + j]) : T(1);
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        if (j < pw) {
#pragma unroll
          for (int i = 0; i < PB; ++i)
            if (i < j) x[j] -= x[i] * ub[i * PB + j];
          x[j] = x[j] / ub[j * PB + j];
        }
      }
#pragma unroll
      for (int j = 0; j < PB; ++j)
        if (j < pw) {
          if (b_T) B[(int64_t)(p + j) * ld_b + r] = x[j];
          else B[(int64_t)r * ld_b + p + j] = x[j];
        }
    }
    __syncthreads();
    const int rn = nx_b - p - pw;
    if (rn > 0) {
      if (!b_T)  // B[:, p+pw:] -= B[:, p:p+pw] U[p:p+pw, p+pw:]
        gemm_block<T, FMODE>(ny_b, rn, pw, T(-1), T(1), &B[p + pw], ld_b, &B[p], ld_b, false,
                             &U[(int64_t)p * ld_u + p + pw], ld_u, false, glds);
      else  // Bs[p+pw:, :] -= U[p:p+pw, p+pw:]^T Bs[p:p+pw, :]
        gemm_block<T, FMODE>(rn, ny_b, pw, T(-1), T(1), &B[(int64_t)(p + pw) * ld_b], ld_b,
                             &U[(int64_t)p * ld_u + p + pw], ld_u, true, &B[(int64_t)p * ld_b], ld_b, false, glds);
    }
  }
  __syncthreads();
}

// B (ny_b x nx_b) <- U^{-1} B, U upper non-unit ny_b x ny_b (backward).
template <typename T, int FMODE>
__device__ void trsmu_block(T* B, const T* U, int nx_b, int ny_b, int ld_b, int ld_u, void* lds) {
  if (trsm_rows<T, 2>(B, U, ny_b, nx_b, ld_b, ld_u, lds, kLdsBytes)) return;
  T* ub = reinterpret_cast<T*>(lds);
  void* glds = reinterpret_cast<char*>(lds) + 8 * 1024;
  const int nblk = (ny_b + PB - 1) / PB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int p = bi * PB;
    const int pw = (ny_b - p) < PB ? (ny_b - p) : PB;
    stage_block(ub, U, ld_u, p, p, pw, pw);
    __syncthreads();
    for (int c = threadIdx.x; c < nx_b; c += blockDim.x) {
      T x[PB];
#pragma unroll
      for (int i = 0; i < PB; ++i) x[i] = (i < pw) ? B[(int64_t)(p + i) * ld_b + c] : T(0);
#pragma unroll
      for (int ii = PB - 1; ii >= 0; --ii) {
        if (ii < pw) {
#pragma unroll
          for (int j = 0; j < PB; ++j)
            if (j > ii && j < pw) x[ii] -= ub[ii * PB + j] * x[j];
          x[ii] = x[ii] / ub[ii * PB + ii];
        }
      }
#pragma unroll
      for (int i = 0; i < PB; ++i)
        if (i < pw) B[(int64_t)(p + i) * ld_b + c] = x[i];
    }
    __syncthreads();
    if (p > 0)  // B[0:p, :] -= U[0:p, p:p+pw] B[p:p+pw, :]
      gemm_block<T, FMODE>(p, nx_b, pw, T(-1), T(1), B, ld_b, &U[p], ld_u, false, &B[(int64_t)p * ld_b], ld_b, false,
                           glds);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- reductions
// Block-wide sum of one value per thread (deterministic fixed tree order).
template <typename T>
__device__ T block_sum(T v, T* red) {
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
__device__ void reflector_dots(const T* Y, int ld, int ny, int r0, int jcol, int c0, int kc, T* w, T* part) {
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
__device__ void householder_qr_q(T* Y, int ld, int ny, int k, T* Q, int ldq, T* tau, T* w, T* red) {
  const int nt = blockDim.x;
  T* part = red;  // blockDim.x scratch
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

// ---------------------------------------------------------------- accum
// U1 (ny x k1), V1 (nx x k1), U2 (ny x k2), V2 (nx x k2); scratch layout:
// Y [0, off1), Vn [off1, off2), Q [off2, ...) then small panels.
template <typename T, int GMODE>
__device__ void accum_block(T* U1, T* V1, const T* U2, const T* V2, int nx, int ny, int k1, int k2, int ld_u1,
                            int ld_v1, int ld_u2, int ld_v2, int off1, int off2, T* tmp, const T* rnd, void* lds) {
  T* Y = tmp;
  T* Vn = tmp + off1;
  T* Q = tmp + off2;
  const int kmax = k1 > k2 ? k1 : k2;
  T* small = Q + (((int64_t)ny * k1 + 31) / 32) * 32;
  T* S = small;                    // k1 x kmax
  T* tau = small + 4 * k1 * kmax;  // kmax
  T* w = tau + kmax;               // kmax
  T* red = reinterpret_cast<T*>(reinterpret_cast<char*>(lds) + 40 * 1024);  // blockDim.x
  // Y = U1 (Omega^T V1)^T + U2 (Omega^T V2)^T
  gemm_block<T, GMODE>(k1, k1, nx, T(1), T(0), S, k1, rnd, k1, true, V1, ld_v1, false, lds);
  gemm_block<T, GMODE>(ny, k1, k1, T(1), T(0), Y, k1, U1, ld_u1, false, S, k1, true, lds);
  gemm_block<T, GMODE>(k1, k2, nx, T(1), T(0), S, k2, rnd, k1, true, V2, ld_v2, false, lds);
  gemm_block<T, GMODE>(ny, k1, k2, T(1), T(1), Y, k1, U2, ld_u2, false, S, k2, true, lds);
  // orthonormal basis of range(Y)
  householder_qr_q<T>(Y, k1, ny, k1, Q, k1, tau, w, red);
  // Vn = V1 (Q^T U1)^T + V2 (Q^T U2)^T
  gemm_block<T, GMODE>(k1, k1, ny, T(1), T(0), S, k1, Q, k1, true, U1, ld_u1, false, lds);
  gemm_block<T, GMODE>(nx, k1, k1, T(1), T(0), Vn, k1, V1, ld_v1, false, S, k1, true, lds);
  gemm_block<T, GMODE>(k1, k2, ny, T(1), T(0), S, k2, Q, k1, true, U2, ld_u2, false, lds);
  gemm_block<T, GMODE>(nx, k1, k2, T(1), T(1), Vn, k1, V2, ld_v2, false, S, k2, true, lds);
  // U1 <- Q, V1 <- Vn
  for (int e = threadIdx.x; e < ny * k1; e += blockDim.x) {
    const int i = e / k1, c = e % k1;
    U1[(int64_t)i * ld_u1 + c] = Q[e];
  }
  for (int e = threadIdx.x; e < nx * k1; e += blockDim.x) {
    const int i = e / k1, c = e % k1;
    V1[(int64_t)i * ld_v1 + c] = Vn[e];
  }
  __syncthreads();
}

}  // namespace pp
