// Launch wrappers of the persistent VM and the small device kernels of
// the solve / refinement path (fills, conversions, H-matrix-vector product,
// norms).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_gemm.h"
#include "kernels.h"
#include "vm.h"
#include "vm_inst.h"

namespace pp {

// Advances the launch epoch and clears the status words (a kernel rather
// than a memset so that the sequence can be captured in a hipGraph).
__global__ void bump_epoch_kernel(uint32_t* e, int32_t* status) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *e = *e + 1u;
  if (threadIdx.x < 8) status[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- launch wrappers
// The interpreter is instantiated once per precision mode in a translation
// unit of its own (vm_fp64.hip, vm_fp32.hip, vm_mixed.hip, through vm_run.h).

hipError_t launch_vm(const VmArgs& a, int precision, int workers, int threads, bool factor, uint32_t* epoch_dev,
                     int32_t* status, hipStream_t s) {
  hipLaunchKernelGGL(bump_epoch_kernel, dim3(1), dim3(64), 0, s, epoch_dev, status);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (precision) {
    case 0: return vm_launch_fp64(a, workers, threads, factor, s);
    case 1: return vm_launch_fp32(a, workers, threads, factor, s);
    case 2: return vm_launch_mixed(a, workers, threads, factor, s);
    default: return hipErrorInvalidValue;
  }
}

int vm_blocks_per_cu(int precision, int threads) {
  if (precision == 0) return vm_blocks_fp64(threads);
  if (precision == 1) return vm_blocks_fp32(threads);
  if (precision == 2) return vm_blocks_mixed(threads);
  return 0;
}

// ---------------------------------------------------------------- small kernels

template <typename T>
__global__ void fill_zero_kernel(T* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

template <typename D, typename S>
__global__ void convert_kernel(D* dst, const S* src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (D)src[i];
}

// x[i, c] += d[i, c] for an n x nrhs block with pitches
template <typename S>
__global__ void add_correction_kernel(double* x, int ldx, const S* d, int ldd, int n, int nrhs) {
  const int64_t total = (int64_t)n * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / nrhs, c = e % nrhs;
    x[i * ldx + c] += (double)d[i * ldd + c];
  }
}

// copy an n x nrhs block between pitched buffers with conversion
template <typename D, typename S>
__global__ void copy2d_kernel(D* dst, int ldd, const S* src, int lds_, int n, int nrhs) {
  const int64_t total = (int64_t)n * nrhs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / nrhs, c = e % nrhs;
    dst[i * ldd + c] = (D)src[i * lds_ + c];
  }
}

// H-matrix-vector product over a block list (nrhs columns), two launches of
// the MFMA workgroup GEMM:
//   out = y + sd * sum(D x) - sum(U (V^T x))
// with sd = -1 and y = b for the FP64 residual r = b - A_H x, and sd = +1,
// y = 0 for applying a hierarchical inverse (whose off-diagonal blocks are
// stored as -U V^T).
// phase 1: one workgroup per low-rank block: t_b = V_b^T x_b (rank x nrhs).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void hmv_lowrank_kernel(const MatvecBlock* blocks, const int* lr_list, const T* A,
                                                           const T* x, int ldx, int nrhs, T* t) {
  const MatvecBlock b = blocks[lr_list[blockIdx.x]];
  gemm_block<T, MODE>(b.rank, nrhs, b.nx, T(1), T(0), t + b.t_off, nrhs, A + b.off_v, b.rank, true,
                      x + (int64_t)b.col0 * ldx, ldx, false);
}

// phase 1 split over row chunks: workgroup (block, s) forms the partial
// product of rows [nx s / S, nx (s + 1) / S) of V_b into partial area s.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void hmv_lowrank_part_kernel(const MatvecBlock* blocks, const int* lr_list,
                                                                const T* A, const T* x, int ldx, int nrhs, T* tp,
                                                                int64_t t_elems) {
  const MatvecBlock b = blocks[lr_list[blockIdx.x]];
  const int s = blockIdx.y, S = gridDim.y;
  const int k0 = (int)((int64_t)b.nx * s / S), k1 = (int)((int64_t)b.nx * (s + 1) / S);
  gemm_block<T, MODE>(b.rank, nrhs, k1 - k0, T(1), T(0), tp + s * t_elems + b.t_off, nrhs,
                      A + b.off_v + (int64_t)k0 * b.rank, b.rank, true, x + (int64_t)(b.col0 + k0) * ldx, ldx, false);
}

// t[e] = sum over s in order of the partial areas
template <typename T>
__global__ void hmv_sum_kernel(T* t, const T* tp, int64_t t_elems, int S) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < t_elems; e += (int64_t)gridDim.x * blockDim.x) {
    T v = tp[e];
    for (int s = 1; s < S; ++s) v += tp[s * t_elems + e];
    t[e] = v;
  }
}

// phase 2: one workgroup per row tile: out = y + contributions of the blocks
// covering the tile, in a fixed order.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void hmv_rows_kernel(const MatvecBlock* blocks, const MatvecTile* tiles,
                                                        const int* contrib, const T* A, const T* x, int ldx, const T* t,
                                                        const T* y, int ldy, T* out, int ldo, int nrhs, T sd) {
  const MatvecTile tl = tiles[blockIdx.x];
  T* o = out + (int64_t)tl.row0 * ldo;
  for (int e = threadIdx.x; e < tl.rows * nrhs; e += blockDim.x) {
    const int i = e / nrhs, c = e % nrhs;
    o[(int64_t)i * ldo + c] = y != nullptr ? y[(int64_t)(tl.row0 + i) * ldy + c] : T(0);
  }
  __syncthreads();
  for (int q = tl.first; q < tl.first + tl.count; ++q) {
    const MatvecBlock b = blocks[contrib[2 * q]];
    const int li = tl.row0 - b.row0;
    if (b.type == 1)
      gemm_block<T, MODE>(tl.rows, nrhs, b.nx, sd, T(1), o, ldo, A + b.off + (int64_t)li * b.nx, b.nx, false,
                          x + (int64_t)b.col0 * ldx, ldx, false);
    else
      gemm_block<T, MODE>(tl.rows, nrhs, b.rank, T(-1), T(1), o, ldo, A + b.off_u + (int64_t)li * b.rank, b.rank,
                          false, t + b.t_off, nrhs, false);
  }
}

// per-column sums of squares of an n x nrhs block (one workgroup per column)
__global__ __launch_bounds__(256) void colnorm2_kernel(const double* r, int ldr, int n, double* out) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double s = 0.;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = r[(int64_t)i * ldr + c];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// per (segment, column) sums of squares; seg holds (row0, rows) pairs
__global__ __launch_bounds__(256) void segnorm2_kernel(const double* r, int ldr, const int* seg, int nrhs,
                                                        double* out) {
  __shared__ double red[256];
  const int g = blockIdx.x / nrhs, c = blockIdx.x % nrhs;
  const int row0 = seg[2 * g], rows = seg[2 * g + 1];
  double s = 0.;
  for (int i = threadIdx.x; i < rows; i += blockDim.x) {
    const double v = r[(int64_t)(row0 + i) * ldr + c];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

static int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

hipError_t launch_fill_zero(void* p, int64_t n, int elem_bytes, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (elem_bytes == 8)
    hipLaunchKernelGGL(fill_zero_kernel<double>, dim3(grid_for(n)), dim3(256), 0, s, (double*)p, n);
  else
    hipLaunchKernelGGL(fill_zero_kernel<float>, dim3(grid_for(n)), dim3(256), 0, s, (float*)p, n);
  return hipGetLastError();
}

hipError_t launch_convert(void* dst, int dst_bytes, const void* src, int src_bytes, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (dst_bytes == 4 && src_bytes == 8)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(n)), dim3(256), 0, s, (float*)dst,
                       (const double*)src, n);
  else if (dst_bytes == 8 && src_bytes == 4)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0, s, (double*)dst,
                       (const float*)src, n);
  else if (dst_bytes == 8 && src_bytes == 8)  // kernels, not memcpy nodes, so graphs capture them plainly
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3(grid_for(n)), dim3(256), 0, s, (double*)dst,
                       (const double*)src, n);
  else
    hipLaunchKernelGGL((convert_kernel<float, float>), dim3(grid_for(n)), dim3(256), 0, s, (float*)dst,
                       (const float*)src, n);
  return hipGetLastError();
}

hipError_t launch_copy2d(void* dst, int dst_bytes, int ldd, const void* src, int src_bytes, int lds_, int n, int nrhs,
                         hipStream_t s) {
  const int g = grid_for((int64_t)n * nrhs);
  if (dst_bytes == 8 && src_bytes == 8)
    hipLaunchKernelGGL((copy2d_kernel<double, double>), dim3(g), dim3(256), 0, s, (double*)dst, ldd,
                       (const double*)src, lds_, n, nrhs);
  else if (dst_bytes == 4 && src_bytes == 8)
    hipLaunchKernelGGL((copy2d_kernel<float, double>), dim3(g), dim3(256), 0, s, (float*)dst, ldd,
                       (const double*)src, lds_, n, nrhs);
  else if (dst_bytes == 8 && src_bytes == 4)
    hipLaunchKernelGGL((copy2d_kernel<double, float>), dim3(g), dim3(256), 0, s, (double*)dst, ldd,
                       (const float*)src, lds_, n, nrhs);
  else
    hipLaunchKernelGGL((copy2d_kernel<float, float>), dim3(g), dim3(256), 0, s, (float*)dst, ldd, (const float*)src,
                       lds_, n, nrhs);
  return hipGetLastError();
}

hipError_t launch_add_correction(double* x, int ldx, const void* d, int d_bytes, int ldd, int n, int nrhs,
                                 hipStream_t s) {
  const int g = grid_for((int64_t)n * nrhs);
  if (d_bytes == 8)
    hipLaunchKernelGGL(add_correction_kernel<double>, dim3(g), dim3(256), 0, s, x, ldx, (const double*)d, ldd, n,
                       nrhs);
  else
    hipLaunchKernelGGL(add_correction_kernel<float>, dim3(g), dim3(256), 0, s, x, ldx, (const float*)d, ldd, n, nrhs);
  return hipGetLastError();
}

template <typename T, int MODE>
static hipError_t launch_hmv_t(const MatvecPlanDev& p, const T* A, const T* x, int ldx, const T* y, int ldy, T* out,
                               int ldo, int nrhs, T* t, T sd, hipStream_t s) {
  if (p.n_lowrank > 0 && p.t_elems > 0) {
    T* tp = t + p.t_elems;
    hipLaunchKernelGGL((hmv_lowrank_part_kernel<T, MODE>), dim3(p.n_lowrank, kHmvSplit), dim3(256), 0, s, p.blocks,
                       p.lr_list, A, x, ldx, nrhs, tp, p.t_elems);
    const int g = (int)std::min<int64_t>((p.t_elems + 255) / 256, 1024);
    hipLaunchKernelGGL((hmv_sum_kernel<T>), dim3(g), dim3(256), 0, s, t, (const T*)tp, p.t_elems, kHmvSplit);
  } else if (p.n_lowrank > 0) {
    hipLaunchKernelGGL((hmv_lowrank_kernel<T, MODE>), dim3(p.n_lowrank), dim3(256), 0, s, p.blocks, p.lr_list, A, x,
                       ldx, nrhs, t);
  }
  hipLaunchKernelGGL((hmv_rows_kernel<T, MODE>), dim3(p.n_tiles), dim3(256), 0, s, p.blocks, p.tiles, p.contrib, A, x,
                     ldx, (const T*)t, y, ldy, out, ldo, nrhs, sd);
  return hipGetLastError();
}

hipError_t launch_hmatvec(const MatvecPlanDev& p, const double* A, const double* x, int ldx, const double* y, int ldy,
                          double* out, int ldo, int nrhs, double* t, hipStream_t s) {
  return launch_hmv_t<double, MODE_FP64>(p, A, x, ldx, y, ldy, out, ldo, nrhs, t, -1.0, s);
}

hipError_t launch_happly(const MatvecPlanDev& p, int elem_bytes, const void* A, const void* x, int ldx, void* out,
                         int ldo, int nrhs, void* t, hipStream_t s) {
  if (elem_bytes == 8)
    return launch_hmv_t<double, MODE_FP64>(p, (const double*)A, (const double*)x, ldx, nullptr, 0, (double*)out, ldo,
                                           nrhs, (double*)t, 1.0, s);
  if (elem_bytes == 4)
    return launch_hmv_t<float, MODE_FP32>(p, (const float*)A, (const float*)x, ldx, nullptr, 0, (float*)out, ldo, nrhs,
                                          (float*)t, 1.0f, s);
  return hipErrorInvalidValue;
}

hipError_t launch_segnorm2(const double* r, int ldr, const int* seg, int nseg, int nrhs, double* out, hipStream_t s) {
  hipLaunchKernelGGL(segnorm2_kernel, dim3(nseg * nrhs), dim3(256), 0, s, r, ldr, seg, nrhs, out);
  return hipGetLastError();
}

hipError_t launch_colnorm2(const double* r, int ldr, int n, int nrhs, double* out, hipStream_t s) {
  hipLaunchKernelGGL(colnorm2_kernel, dim3(nrhs), dim3(256), 0, s, r, ldr, n, out);
  return hipGetLastError();
}

// Sketch values of the library's generator, one thread per value:
// splitmix64's i-th output mixes seed + (i + 1) * golden.
__global__ void sketch_fill_kernel(void* out, int64_t count, int elem_bytes, uint64_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const double v = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
    if (elem_bytes == 8) static_cast<double*>(out)[i] = v;
    else static_cast<float*>(out)[i] = (float)v;
  }
}

__global__ void clear_i32_kernel(int32_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

hipError_t launch_sketch_fill(void* out, int64_t count, int elem_bytes, uint64_t seed, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int g = (int)std::min<int64_t>((count + 255) / 256, 1024);
  hipLaunchKernelGGL(sketch_fill_kernel, dim3(g), dim3(256), 0, s, out, count, elem_bytes, seed);
  return hipGetLastError();
}

hipError_t launch_clear_i32(int32_t* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(clear_i32_kernel, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

// MFMA lane-layout probe: D = A B for one 16x16xK tile with A[i][k] = a[i*K+k],
// B[k][j] = b[k*16+j]; lane l writes its 4 accumulator values to out[4l..].
__global__ void mfma_probe_kernel(int mode, const double* a, const double* b, double* out) {
  const int l = threadIdx.x;
  if (mode == 0) {
    d4_t c = {0., 0., 0., 0.};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[(l & 15) * 4 + (l >> 4)], b[(l >> 4) * 16 + (l & 15)], c, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[4 * l + r] = c[r];
  } else if (mode == 1) {
    f4_t c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a[(l & 15) * 4 + (l >> 4)], (float)b[(l >> 4) * 16 + (l & 15)], c,
                                            0, 0, 0);
    for (int r = 0; r < 4; ++r) out[4 * l + r] = c[r];
  } else {
    f4_t c = {0.f, 0.f, 0.f, 0.f};
    float av[8], bv[8];
    for (int e = 0; e < 8; ++e) {
      av[e] = (float)a[(l & 15) * 32 + (l >> 4) * 8 + e];
      bv[e] = (float)b[((l >> 4) * 8 + e) * 16 + (l & 15)];
    }
    c = Mma<float, MODE_BF16>::mma(av, bv, c);
    for (int r = 0; r < 4; ++r) out[4 * l + r] = c[r];
  }
}

hipError_t launch_mfma_probe(int mode, const double* a, const double* b, double* out, hipStream_t s) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, mode, a, b, out);
  return hipGetLastError();
}

hipError_t launch_single_op(int precision, int threads, int reps, const SingleOp& op, void* scratch, const void* rnd,
                            int32_t* status, hipStream_t s) {
  if (precision == 0) return vm_single_fp64(threads, reps, op, scratch, rnd, status, s);
  if (precision == 1) return vm_single_fp32(threads, reps, op, scratch, rnd, status, s);
  return vm_single_mixed(threads, reps, op, scratch, rnd, status, s);
}

}  // namespace pp
