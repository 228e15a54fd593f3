// LDS-resident leaf routines for the shapes on the H-LU critical path.
//
// Same semantics as the panel-blocked routines of dev_lu.h (no pivoting,
// unit-lower L, non-unit upper U), organised for latency instead:
//   * getrf_lds: the whole ny x nx block lives in LDS (row pitch nx + 1 to
//     spread banks); one thread owns one row; per pivot column one barrier.
//   * trsm_rows_*: the right-hand-side block (<= 64 columns) lives in LDS, a
//     thread owns one row of it, and the triangular factor is streamed in
//     16-column (or 16-row) panels through a double-buffered LDS staging
//     area, so each substitution step costs one barrier and no global load
//     on the dependency chain.
// Every routine returns false without touching memory when the shape does
// not fit (the caller then uses the panel-blocked version).
#pragma once

#include "dev_gemm.h"

namespace pp {

constexpr int kPanel = 16;

template <typename T>
__device__ __forceinline__ bool lds_fits(int64_t elems, int cap_bytes) {
  return elems * (int64_t)sizeof(T) <= cap_bytes;
}

// ---------------------------------------------------------------- getrf
template <typename T>
__device__ bool getrf_lds(T* M, int nx, int ny, int ld, void* lds, int cap_bytes) {
  const int sp = nx + 1;
  if (ny > (int)blockDim.x || !lds_fits<T>((int64_t)ny * sp, cap_bytes)) return false;
  T* s = reinterpret_cast<T*>(lds);
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < ny * nx; e += nt) {
    const int i = e / nx, c = e - i * nx;
    s[i * sp + c] = M[(int64_t)i * ld + c];
  }
  __syncthreads();
  const int mn = nx < ny ? nx : ny;
  T* ri = s + tid * sp;
  constexpr int CH = 16;  // LDS loads of a chunk are issued before its FMAs
  for (int j = 0; j < mn; ++j) {
    if (tid > j && tid < ny) {
      const T* rj = s + j * sp;
      const T l = ri[j] / rj[j];
      ri[j] = l;
      int c = j + 1;
      for (; c + CH <= nx; c += CH) {
        T a[CH], b[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          a[u] = rj[c + u];
          b[u] = ri[c + u];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) ri[c + u] = b[u] - l * a[u];
      }
      for (; c < nx; ++c) ri[c] -= l * rj[c];
    }
    __syncthreads();
  }
  for (int e = tid; e < ny * nx; e += nt) {
    const int i = e / nx, c = e - i * nx;
    M[(int64_t)i * ld + c] = s[i * sp + c];
  }
  __syncthreads();
  return true;
}

// ---------------------------------------------------------------- staging helpers
// Stage the n x kPanel column panel P[0:n, j0:j0+kPanel] (pitch ld) into dst
// (row i at dst[i * kPanel]); columns past jn are left unset.
template <typename T>
__device__ __forceinline__ void stage_cols(T* dst, const T* P, int ld, int n, int j0, int jn) {
  for (int e = threadIdx.x; e < n * kPanel; e += blockDim.x) {
    const int i = e / kPanel, jj = e - i * kPanel;
    if (j0 + jj < jn) dst[e] = P[(int64_t)i * ld + j0 + jj];
  }
}

// Stage the kPanel x n row panel P[j0:j0+kPanel, 0:n] into dst (row jj at
// dst[jj * n]); rows past jn are left unset.
template <typename T>
__device__ __forceinline__ void stage_rows(T* dst, const T* P, int ld, int n, int j0, int jn) {
  for (int e = threadIdx.x; e < n * kPanel; e += blockDim.x) {
    const int jj = e / n, i = e - jj * n;
    if (j0 + jj < jn) dst[e] = P[(int64_t)(j0 + jj) * ld + i];
  }
}

// B (n x nb, pitch ldb) into LDS rows of pitch nb + 1 and back.
template <typename T>
__device__ __forceinline__ void load_rows(T* s, const T* B, int ldb, int n, int nb) {
  const int sp = nb + 1;
  for (int e = threadIdx.x; e < n * nb; e += blockDim.x) {
    const int i = e / nb, c = e - i * nb;
    s[i * sp + c] = B[(int64_t)i * ldb + c];
  }
}

template <typename T>
__device__ __forceinline__ void store_rows(T* B, int ldb, const T* s, int n, int nb) {
  const int sp = nb + 1;
  for (int e = threadIdx.x; e < n * nb; e += blockDim.x) {
    const int i = e / nb, c = e - i * nb;
    B[(int64_t)i * ldb + c] = s[i * sp + c];
  }
}

// Row-owner substitution: a thread owns one row of the right-hand side and
// the n steps run in order, one barrier each.  coefficient(i, j) comes from a
// staged panel of the factor; after the steps every row is divided by its
// diagonal where the factor is non-unit.  Mode:
//   0  B <- L^{-1} B, L unit lower (n x n, pitch ldf):   coef = L[i][j], i > j
//   1  Bs <- U^{-T} Bs, U upper non-unit:                coef = U[j][i] / U[j][j], i > j; then /U[i][i]
//   2  B <- U^{-1} B, U upper non-unit (backward):       coef = U[i][j] / U[j][j], i < j; then /U[i][i]
// Rows of up to kRegRow columns live in registers and the pivot row of each
// step is published through a double-buffered LDS row; wider rows (<= 64)
// live in LDS.
constexpr int kRegRow = 32;

template <typename T, int MODE>
__device__ bool trsm_rows(T* B, const T* F, int n, int nb, int ldb, int ldf, void* lds, int cap_bytes) {
  const bool reg = nb <= kRegRow;
  const int sp = nb + 1;
  const int64_t rows_elems = reg ? 2LL * kRegRow : (int64_t)n * sp;
  if (n > (int)blockDim.x || nb > 64 || nb < 1 || !lds_fits<T>(rows_elems + 2LL * n * kPanel + 2 * kPanel, cap_bytes))
    return false;
  T* s = reinterpret_cast<T*>(lds);  // register mode: 2 published rows; LDS mode: all rows
  T* pan[2] = {s + rows_elems, s + rows_elems + (int64_t)n * kPanel};
  T* dg = pan[1] + (int64_t)n * kPanel;  // 2 x kPanel diagonal entries
  const int tid = threadIdx.x, nt = blockDim.x;
  const int npan = (n + kPanel - 1) / kPanel;
  // A panel covers steps j in [j0, j0 + kPanel).  Its n * kPanel coefficients
  // are prefetched into registers (<= 16 per thread since n <= blockDim.x)
  // one panel ahead and committed to LDS before the last barrier of the
  // current panel, so global latency stays off the substitution chain.
  T pre[kPanel];
  T pre_d = T(0);
  auto panel_j0 = [&](int p) { return (MODE == 2) ? (npan - 1 - p) * kPanel : p * kPanel; };
  auto prefetch = [&](int p) {
    const int j0 = panel_j0(p);
#pragma unroll
    for (int q = 0; q < kPanel; ++q) {
      const int e = tid + q * nt;
      pre[q] = T(0);
      if (e < n * kPanel) {
        if (MODE == 1) {  // row panel: U[j0 + jj][i]
          const int jj = e / n, i = e - jj * n;
          if (j0 + jj < n) pre[q] = F[(int64_t)(j0 + jj) * ldf + i];
        } else {  // column panel: F[i][j0 + jj]
          const int i = e / kPanel, jj = e - i * kPanel;
          if (j0 + jj < n) pre[q] = F[(int64_t)i * ldf + j0 + jj];
        }
      }
    }
    if (MODE != 0 && tid < kPanel && j0 + tid < n) pre_d = F[(int64_t)(j0 + tid) * ldf + j0 + tid];
  };
  auto commit = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kPanel; ++q) {
      const int e = tid + q * nt;
      if (e < n * kPanel) pan[buf][e] = pre[q];
    }
    if (MODE != 0 && tid < kPanel) dg[buf * kPanel + tid] = pre_d;
  };
  // rows
  T r[kRegRow];
  const bool own = tid < n;
  if (reg) {
#pragma unroll
    for (int c = 0; c < kRegRow; ++c) r[c] = (own && c < nb) ? B[(int64_t)tid * ldb + c] : T(0);
    const int first = (MODE == 2) ? n - 1 : 0;
    if (tid == first) {
#pragma unroll
      for (int c = 0; c < kRegRow; ++c) s[c] = r[c];
    }
  } else {
    load_rows(s, B, ldb, n, nb);
  }
  prefetch(0);
  commit(0);
  __syncthreads();
  T* ri = s + tid * sp;
  int cur = 0;
  for (int p = 0; p < npan; ++p) {
    const int buf = p & 1;
    if (p + 1 < npan) prefetch(p + 1);
    const int j0 = panel_j0(p);
    const T* P = pan[buf];
    for (int t = 0; t < kPanel; ++t) {
      if (t == kPanel - 1 && p + 1 < npan) commit(buf ^ 1);
      const int jj = (MODE == 2) ? kPanel - 1 - t : t;
      const int j = j0 + jj;
      if (j < n) {
        const bool act = (MODE == 2) ? (tid < j) : (tid > j && own);
        T coef = T(0);
        if (act) {
          if (MODE == 0) coef = P[tid * kPanel + jj];
          else if (MODE == 1) coef = P[jj * n + tid] / dg[buf * kPanel + jj];
          else coef = P[tid * kPanel + jj] / dg[buf * kPanel + jj];
        }
        if (reg) {
          const T* rj = s + cur * kRegRow;
          if (act) {
#pragma unroll
            for (int c = 0; c < kRegRow; ++c) r[c] -= coef * rj[c];
          }
          const int nxt = (MODE == 2) ? j - 1 : j + 1;
          if (tid == nxt && nxt >= 0 && nxt < n) {
#pragma unroll
            for (int c = 0; c < kRegRow; ++c) s[(cur ^ 1) * kRegRow + c] = r[c];
          }
          cur ^= 1;
        } else if (act) {
          const T* rj = s + j * sp;
          constexpr int CH = 16;
          for (int c0 = 0; c0 < nb; c0 += CH) {
            T a[CH], b[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
              a[u] = (c0 + u < nb) ? rj[c0 + u] : T(0);
              b[u] = (c0 + u < nb) ? ri[c0 + u] : T(0);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u)
              if (c0 + u < nb) ri[c0 + u] = b[u] - coef * a[u];
          }
        }
      }
      __syncthreads();
    }
  }
  if (reg) {
    if (own) {
      const T d = (MODE != 0) ? F[(int64_t)tid * ldf + tid] : T(1);
#pragma unroll
      for (int c = 0; c < kRegRow; ++c)
        if (c < nb) B[(int64_t)tid * ldb + c] = (MODE != 0) ? r[c] / d : r[c];
    }
  } else {
    if (MODE != 0 && own) {
      const T d = F[(int64_t)tid * ldf + tid];
      for (int c = 0; c < nb; ++c) ri[c] /= d;
    }
    __syncthreads();
    store_rows(B, ldb, s, n, nb);
  }
  __syncthreads();
  return true;
}

}  // namespace pp
