// Partial-pivoted dense LU and row permutation of row-major blocks.
//
// The factorization follows the restatement of the reference's pivot helpers
// (src/kernel.cu:958-1051, dead code there): column j's pivot is the row of
// largest |a_ij| at or below the diagonal (smallest row on ties), rows are
// exchanged whole, multipliers are a_ij / u_jj and every trailing element
// takes one fused multiply-add per step.  Those are exactly the operations of
// the CPU restatement, so the results are bit-identical to it.
//
// One workgroup of 256 threads per block (blockIdx.x = block of a batch); a
// step is an argmax over the column (per-thread scan, 64-lane butterfly,
// four-wave merge in LDS), the row exchange, the multiplier column and a
// rank-1 update with a wave per row and lanes along the row (coalesced).
// Blocks up to a few hundred rows: the rank-1 updates make it bandwidth- and
// barrier-bound, one step per column.
#include <hip/hip_runtime.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace pp {
namespace {

constexpr int kPivThreads = 256;

template <typename T>
__device__ __forceinline__ T fused(T a, T b, T c) {
  if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
  else return __builtin_fmaf(a, b, c);
}

template <typename T>
__device__ __forceinline__ T magnitude(T x) {
  if constexpr (sizeof(T) == 8) return __builtin_fabs(x);
  else return __builtin_fabsf(x);
}

// (v, i) beats (bv, bi): larger magnitude, then smaller row
template <typename T>
__device__ __forceinline__ bool better(T v, int i, T bv, int bi) {
  return v > bv || (v == bv && i < bi);
}

template <typename T>
__global__ __launch_bounds__(kPivThreads) void getrf_pivoted_kernel(T* A0, int ld, int ny, int nx, int64_t stride,
                                                                    int32_t* P0, int32_t* info0) {
  T* A = A0 + (int64_t)blockIdx.x * stride;
  int32_t* p = P0 + (int64_t)blockIdx.x * ny;
  __shared__ T s_v[kPivThreads / 64];
  __shared__ int s_i[kPivThreads / 64];
  __shared__ int s_best;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = kPivThreads / 64;
  for (int i = tid; i < ny; i += kPivThreads) p[i] = i;
  int info = 0;
  const int K = ny < nx ? ny : nx;
  for (int j = 0; j < K; ++j) {
    // argmax of |a_ij| over i >= j; threads without a row (or with only NaN)
    // offer the placeholder (-1, ny)
    T bv = T(-1);
    int bi = ny;
    for (int i = j + tid; i < ny; i += kPivThreads) {
      const T v = magnitude(A[(int64_t)i * ld + j]);
      if (v > bv) {  // rows ascend within a thread: the first maximum stays
        bv = v;
        bi = i;
      }
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      const T ov = __shfl_xor(bv, m);
      const int oi = __shfl_xor(bi, m);
      if (better(ov, oi, bv, bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      s_v[w] = bv;
      s_i[w] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      T v = s_v[0];
      int b = s_i[0];
      for (int q = 1; q < nw; ++q)
        if (better(s_v[q], s_i[q], v, b)) {
          v = s_v[q];
          b = s_i[q];
        }
      // NaN never wins a comparison: a NaN diagonal keeps its row (as the
      // sequential scan that starts from it does), and a column that is NaN
      // at and below the diagonal leaves every thread's placeholder row ny,
      // which is not a row -- the diagonal row is taken then as well
      const T d = A[(int64_t)j * ld + j];
      s_best = (b >= ny || d != d) ? j : b;
    }
    __syncthreads();
    const int best = s_best;
    if (best != j) {
      for (int k = tid; k < nx; k += kPivThreads) {
        const T t = A[(int64_t)j * ld + k];
        A[(int64_t)j * ld + k] = A[(int64_t)best * ld + k];
        A[(int64_t)best * ld + k] = t;
      }
      if (tid == 0) {
        const int32_t t = p[j];
        p[j] = p[best];
        p[best] = t;
      }
      __syncthreads();
    }
    const T piv = A[(int64_t)j * ld + j];
    if (piv == T(0)) {  // the column is zero at and below the diagonal
      if (info == 0) info = j + 1;
      continue;
    }
    for (int i = j + 1 + tid; i < ny; i += kPivThreads) A[(int64_t)i * ld + j] = A[(int64_t)i * ld + j] / piv;
    __syncthreads();
    const T* urow = A + (int64_t)j * ld;
    for (int i = j + 1 + w; i < ny; i += nw) {
      T* row = A + (int64_t)i * ld;
      const T l = row[j];
      for (int k = j + 1 + lane; k < nx; k += 64) row[k] = fused(-l, urow[k], row[k]);
    }
    __syncthreads();
  }
  if (tid == 0) info0[blockIdx.x] = info;
}

// Rows of a column chunk through LDS: forward, row r <- row p[r]; recover,
// row p[r] <- row r (the inverse permutation).
template <typename T>
__global__ __launch_bounds__(kPivThreads) void apply_pivot_kernel(T* M, int ld, int nx, int ny, const int32_t* p,
                                                                  int recover, int cw) {
  extern __shared__ __attribute__((aligned(16))) char pivot_smem[];
  T* s = reinterpret_cast<T*>(pivot_smem);
  const int c0 = blockIdx.x * cw;
  const int cols = (nx - c0) < cw ? (nx - c0) : cw;
  for (int e = threadIdx.x; e < ny * cols; e += blockDim.x) {
    const int r = e / cols, c = e - r * cols, q = p[r];
    if (recover) s[q * cols + c] = M[(int64_t)r * ld + c0 + c];
    else s[r * cols + c] = M[(int64_t)q * ld + c0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ny * cols; e += blockDim.x) {
    const int r = e / cols, c = e - r * cols;
    M[(int64_t)r * ld + c0 + c] = s[r * cols + c];
  }
}

}  // namespace

hipError_t launch_getrf_pivoted(int fp64, void* a, int ld, int ny, int nx, int count, int64_t stride,
                                int32_t* pivots, int32_t* info, hipStream_t s) {
  if (fp64)
    hipLaunchKernelGGL(getrf_pivoted_kernel<double>, dim3(count), dim3(kPivThreads), 0, s, (double*)a, ld, ny, nx,
                       stride, pivots, info);
  else
    hipLaunchKernelGGL(getrf_pivoted_kernel<float>, dim3(count), dim3(kPivThreads), 0, s, (float*)a, ld, ny, nx,
                       stride, pivots, info);
  return hipGetLastError();
}

int apply_pivot_chunk(int fp64, int ny) {
  const int es = fp64 ? 8 : 4;
  if (ny < 1 || (int64_t)ny * es > kPivotLdsBytes) return 0;
  const int64_t cw = kPivotLdsBytes / ((int64_t)ny * es);
  return (int)(cw < 64 ? cw : 64);
}

hipError_t launch_apply_pivot(int fp64, void* m, int ld, int nx, int ny, const int32_t* pivots, int recover,
                              hipStream_t s) {
  const int cw = apply_pivot_chunk(fp64, ny);
  if (cw < 1) return hipErrorInvalidValue;
  const int blocks = (nx + cw - 1) / cw;
  const size_t lds = (size_t)ny * cw * (fp64 ? 8 : 4);
  if (fp64)
    hipLaunchKernelGGL(apply_pivot_kernel<double>, dim3(blocks), dim3(kPivThreads), lds, s, (double*)m, ld, nx, ny,
                       pivots, recover, cw);
  else
    hipLaunchKernelGGL(apply_pivot_kernel<float>, dim3(blocks), dim3(kPivThreads), lds, s, (float*)m, ld, nx, ny,
                       pivots, recover, cw);
  return hipGetLastError();
}

}  // namespace pp
