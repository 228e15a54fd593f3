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

// Four elements per thread, one thread each (16-byte aligned arrays; the
// tail of n % 4 elements by the last threads): a large grid with every load
// of the conversion in flight at once.
template <typename D, typename S>
__global__ __launch_bounds__(256) void convert4_kernel(D* dst, const S* src, int64_t n) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * 4;
  if (i + 4 <= n) {
    S v[4];
    if constexpr (sizeof(S) == 8) {
      const d2_t a = gload<d2_t>(src + i), b = gload<d2_t>(src + i + 2);
      v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
    } else {
      const f4_t a = gload<f4_t>(src + i);
      v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
    }
    if constexpr (sizeof(D) == 8) {
      gstore<d2_t>(dst + i, d2_t{(double)v[0], (double)v[1]});
      gstore<d2_t>(dst + i + 2, d2_t{(double)v[2], (double)v[3]});
    } else {
      gstore<f4_t>(dst + i, f4_t{(float)v[0], (float)v[1], (float)v[2], (float)v[3]});
    }
  } else {
    for (int64_t j = i; j < n; ++j) dst[j] = (D)src[j];
  }
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

// ---------------------------------------------------------------- H-matrix-vector product
// out = y + sd * sum(D x) - sum(U (V^T x)) over a block list, nrhs columns:
// sd = -1 and y = b for the FP64 residual r = b - A_H x; sd = +1 and y = 0
// for applying a hierarchical inverse (whose off-diagonal blocks are stored
// as -U V^T).  Three launches, all HBM-bound skinny products:
//   1. hmv_vtx_kernel: partial products of V_b^T x over row parts of V_b (at
//      most kPartRows rows), one workgroup per (part, 16-row tile of the
//      rank, 64-column group);
//   2. hmv_sum_kernel: t_b = the partials of block b added in part order;
//   3. hmv_rows_kernel: one workgroup per (16-row group of a row tile,
//      64-column group).
// The waves of a workgroup split the K chunks (16 wide) of its output tile
// round-robin and issue the loads of several chunks before their MFMAs, so
// every wave keeps whole chunks in flight (a wave that waited for one chunk
// at a time would be latency-bound at a few percent of the HBM rate); the
// four partial tiles are added through LDS in wave order, so results are
// deterministic.  Small LDS footprint (no VM arena): several workgroups per
// CU.

// Operands of one 16-wide K chunk for lane (li, g) of a wave: A[li][k] (times
// sign) and B[k][16 cc + li] at k = kb + s (kb = 16 chunk + 4 g, s < 4) for
// k < kz, zero from there on (kz <= K; addresses clamped to k = K - 1, which
// exists).
template <typename T, int NC>
struct HmvChunk {
  T a[4];
  T b[NC][4];
};

template <typename T, int NC>
__device__ __forceinline__ void hmv_load(HmvChunk<T, NC>& u, const T* arow, int a_step, bool a_vec, const T* bbase,
                                         int ldb, int col, int ncol, int kb, int K, int kz, T sign) {
  if (a_vec) {  // whole chunk inside K, a_step 1, 16-byte aligned runs
    if constexpr (sizeof(T) == 8) {
      const d2_t v0 = gload<d2_t>(arow + kb), v1 = gload<d2_t>(arow + kb + 2);
      u.a[0] = v0.x, u.a[1] = v0.y, u.a[2] = v1.x, u.a[3] = v1.y;
    } else {
      const f4_t v = gload<f4_t>(arow + kb);
      u.a[0] = v.x, u.a[1] = v.y, u.a[2] = v.z, u.a[3] = v.w;
    }
  } else {
#pragma unroll
    for (int s = 0; s < 4; ++s) u.a[s] = gload<T>(arow + (int64_t)min(kb + s, K - 1) * a_step);
  }
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    const int c = min(col + 16 * cc, ncol - 1);
#pragma unroll
    for (int s = 0; s < 4; ++s) u.b[cc][s] = gload<T>(bbase + (int64_t)min(kb + s, K - 1) * ldb + c);
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const bool in = kb + s < kz;
    u.a[s] = in ? u.a[s] * sign : T(0);
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) u.b[cc][s] = in ? u.b[cc][s] : T(0);
  }
}

// Adds the four waves' partial 16 x (16 NC) tiles (LDS, wave order) and
// hands element (i, c) of the tile, i = threadIdx / 16, c = 16 cc +
// threadIdx % 16, to emit(i, cc, value).
template <typename T, int MODE, int NC, typename Emit>
__device__ __forceinline__ void hmv_reduce(T* red, const typename Mma<T, MODE>::acc_t (&acc)[NC], Emit emit) {
  const int l = lane_id(), w = wave_id();
#pragma unroll
  for (int cc = 0; cc < NC; ++cc)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[((w * NC + cc) * 4 + r) * 64 + l] = acc[cc][r];
  __syncthreads();
  const int i = threadIdx.x >> 4, j = threadIdx.x & 15;
  // lane and accumulator register holding row i, column j of an MFMA tile
  const int ls = MODE == MODE_FP64 ? 16 * (i & 3) + j : 16 * (i >> 2) + j;
  const int rs = MODE == MODE_FP64 ? i >> 2 : i & 3;
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    T v = red[((0 * NC + cc) * 4 + rs) * 64 + ls];
    v += red[((1 * NC + cc) * 4 + rs) * 64 + ls];
    v += red[((2 * NC + cc) * 4 + rs) * 64 + ls];
    v += red[((3 * NC + cc) * 4 + rs) * 64 + ls];
    emit(i, cc, v);
  }
}

// Stage 1: rows [k0, k1) of V_b (part s of block b) against x: the 16 x 16 NC
// tile (rows 16 rt.. of V^T) of the partial product, into the part's area
// (rank x nrhs, pitch nrhs).
template <typename T, int MODE, int NC>
__global__ __launch_bounds__(256) void hmv_vtx_kernel(const MatvecBlock* blocks, const int* parts, const T* A,
                                                       const T* x, int ldx, int nrhs, T* tp) {
  using M = Mma<T, MODE>;
  __shared__ T red[4 * NC * 4 * 64];
  const MatvecBlock b = blocks[parts[2 * blockIdx.x]];
  const int part = parts[2 * blockIdx.x + 1];
  const int i0 = 16 * blockIdx.y, c0 = 64 * blockIdx.z;
  if (i0 >= b.rank || c0 >= nrhs) return;  // whole workgroup
  const int k0 = part * kPartRows, K = min(b.nx - k0, kPartRows);
  const int l = lane_id(), w = wave_id(), li = l & 15, g = l >> 4;
  // op(A)[i][k] = V[k0 + k][i0 + i] (V: nx x rank, row pitch rank); B = x rows col0 + k0..
  const T* arow = A + b.off_v + (int64_t)k0 * b.rank + min(i0 + li, b.rank - 1);
  const T* bbase = x + (int64_t)(b.col0 + k0) * ldx;
  typename M::acc_t acc[NC];
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) acc[cc] = M::zero();
  const int nch = (K + 15) >> 4;
  for (int ch = w; ch < nch; ch += 8) {  // two chunks per pass, loads of both in flight
    HmvChunk<T, NC> u0, u1;
    const bool two = ch + 4 < nch;  // else the second chunk repeats the first, zeroed
    hmv_load<T, NC>(u0, arow, b.rank, false, bbase, ldx, c0 + li, nrhs, 16 * ch + 4 * g, K, K, T(1));
    hmv_load<T, NC>(u1, arow, b.rank, false, bbase, ldx, c0 + li, nrhs, 16 * (two ? ch + 4 : ch) + 4 * g, K,
                    two ? K : 0, T(1));
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) acc[cc] = M::mma_chunk(u0.a, u0.b[cc], acc[cc]);
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) acc[cc] = M::mma_chunk(u1.a, u1.b[cc], acc[cc]);
  }
  T* o = tp + b.p_off + (int64_t)part * b.p_stride;
  hmv_reduce<T, MODE, NC>(red, acc, [&](int i, int cc, T v) {
    const int c = c0 + 16 * cc + (threadIdx.x & 15);
    if (i0 + i < b.rank && c < nrhs) o[(int64_t)(i0 + i) * nrhs + c] = v;
  });
}

// Ordered sum over s < n of p[s * stride] by the four 64-thread groups of a
// workgroup: group g adds s = g, g + 4, ... with eight loads in flight, and
// thread l of group 0 returns the four group sums of column l added in group
// order (other threads return 0).  Every thread of the workgroup calls it.
template <typename T>
__device__ __forceinline__ T ordered_sum4(const T* p, int64_t stride, int n, bool valid, T* red) {
  const int grp = threadIdx.x >> 6, l = threadIdx.x & 63;
  T v = T(0);
  if (valid) {
    int s = grp;
    for (; s + 28 < n; s += 32) {
      T q[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) q[j] = p[(int64_t)(s + 4 * j) * stride];
#pragma unroll
      for (int j = 0; j < 8; ++j) v += q[j];
    }
    for (; s < n; s += 4) v += p[(int64_t)s * stride];
  }
  red[threadIdx.x] = v;
  __syncthreads();
  T r = T(0);
  if (grp == 0) r = ((red[l] + red[64 + l]) + red[128 + l]) + red[192 + l];
  __syncthreads();
  return r;
}

// Stage 2: t_b = the partials of block b added (ordered_sum4 order), 64
// elements per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void hmv_sum_kernel(const MatvecBlock* blocks, const int* lr_list, T* t,
                                                       const T* tp, int nrhs) {
  __shared__ T red[256];
  const MatvecBlock b = blocks[lr_list[blockIdx.x]];
  const int n = b.rank * nrhs, e = 64 * blockIdx.y + (threadIdx.x & 63);
  if (64 * (int)blockIdx.y >= n) return;  // whole workgroup
  const T v = ordered_sum4<T>(tp + b.p_off + min(e, n - 1), b.p_stride, b.nparts, e < n, red);
  if (threadIdx.x < 64 && e < n) t[b.t_off + e] = v;
}

// One leaf covering a row group, as the waves of hmv_rows_kernel address it:
// lane row i of op(A) at A + a_off + i * pitch (K values), B rows at
// (x or t) + b_off with pitch ldb.
struct HmvLeaf {
  int64_t a_off, b_off;
  int32_t pitch, K, ldb, kind;  // kind bit 0: low rank (B from t, sign -1); bit 1: 16-byte A runs
};
constexpr int kHmvMaxLeaves = 64;

// Stage 3: rows [r0, r0 + 16) of row tile blockIdx.x, columns c0 .. c0 + 16 NC.
// The leaves covering the tile are described in LDS first; their (leaf, K
// chunk) units are dealt to the waves round-robin, UB units per pass with
// all their loads in flight.
template <typename T, int MODE, int NC>
__global__ __launch_bounds__(256) void hmv_rows_kernel(const MatvecBlock* blocks, const MatvecTile* tiles,
                                                        const int* contrib, const T* A, const T* x, int ldx,
                                                        const T* t, const T* y, int ldy, T* out, int ldo, int nrhs,
                                                        T sd, HmvExtra ex) {
  using M = Mma<T, MODE>;
  // units per pass: about 64 VGPRs of operands per lane
  constexpr int UB = sizeof(T) == 8 ? (NC == 1 ? 4 : 2) : (NC == 1 ? 8 : NC == 2 ? 4 : 2);
  __shared__ T red[4 * NC * 4 * 64];
  __shared__ HmvLeaf leaf[kHmvMaxLeaves];
  const MatvecTile tl = tiles[blockIdx.x];
  const int r0 = 16 * blockIdx.y, c0 = 64 * blockIdx.z;
  if (c0 >= nrhs) return;  // whole workgroup
  if (r0 >= tl.rows) {      // whole workgroup: a row group past a short tile adds nothing to the norms
    if (ex.nrm_part != nullptr && threadIdx.x < 64 && c0 + (int)threadIdx.x < nrhs)
      ex.nrm_part[(int64_t)(c0 + threadIdx.x) * gridDim.x * (kTileRows / 16) + blockIdx.x * (kTileRows / 16) +
                  blockIdx.y] = 0.;
    return;
  }
  const int rows = min(16, tl.rows - r0);
  const int l = lane_id(), w = wave_id(), li = l & 15, g = l >> 4;
  const int ri = min(li, rows - 1);  // rows past the tile repeat its last row and are not stored
  const bool a_al = ((uintptr_t)A & 15) == 0;
  auto describe = [&](int q) {
    const MatvecBlock b = blocks[contrib[2 * (tl.first + q)]];
    const int64_t lrow = tl.row0 - b.row0 + r0;
    HmvLeaf d;
    if (b.type == 1) {
      d.pitch = b.nx;
      d.a_off = b.off + lrow * b.nx;
      d.b_off = (int64_t)b.col0 * ldx;
      d.ldb = ldx;
      d.kind = 0;
    } else {
      d.pitch = b.rank;
      d.a_off = b.off_u + lrow * b.rank;
      d.b_off = b.t_off;
      d.ldb = nrhs;
      d.kind = 1;
    }
    d.K = d.pitch;
    if (a_al && ((d.pitch * (int)sizeof(T)) & 15) == 0 && ((d.a_off * (int64_t)sizeof(T)) & 15) == 0) d.kind |= 2;
    return d;
  };
  const int nq = min(tl.count, kHmvMaxLeaves);
  for (int q = threadIdx.x; q < nq; q += blockDim.x) leaf[q] = describe(q);
  __syncthreads();
  typename M::acc_t acc[NC];
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) acc[cc] = M::zero();
  // unit iterator over (leaf q, chunk ch), this wave's units w, w + 4, ...
  int q = 0, ch = w;
  HmvLeaf d{};
  auto fetch = [&]() { d = q < kHmvMaxLeaves ? leaf[q] : describe(q); };
  auto seek = [&]() {  // move (q, ch) inside leaf q, or q = tl.count
    while (q < tl.count) {
      fetch();
      const int nch = (d.K + 15) >> 4;
      if (ch < nch) return;
      ch -= nch;
      ++q;
    }
  };
  seek();
  while (q < tl.count) {
    HmvChunk<T, NC> u[UB];
#pragma unroll
    for (int j = 0; j < UB; ++j) {
      if (q < tl.count) {
        const bool full = 16 * ch + 16 <= d.K;
        const bool lr = d.kind & 1;
        hmv_load<T, NC>(u[j], A + d.a_off + (int64_t)ri * d.pitch, 1, (d.kind & 2) && full, (lr ? t : x) + d.b_off,
                        d.ldb, c0 + li, nrhs, 16 * ch + 4 * g, d.K, d.K, lr ? T(-1) : sd);
        ch += 4;
        if (ch >= ((d.K + 15) >> 4)) seek();
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          u[j].a[s] = T(0);
#pragma unroll
          for (int cc = 0; cc < NC; ++cc) u[j].b[cc][s] = T(0);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < UB; ++j)
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) acc[cc] = M::mma_chunk(u[j].a, u[j].b[cc], acc[cc]);
  }
  const int64_t row_base = tl.row0 + r0;
  double sq[NC];
  hmv_reduce<T, MODE, NC>(red, acc, [&](int i, int cc, T v) {
    const int c = c0 + 16 * cc + (threadIdx.x & 15);
    sq[cc] = 0.;
    if (i < rows && c < nrhs) {
      const int64_t row = row_base + i;
      const T o = (y != nullptr ? y[row * ldy + c] : T(0)) + v;
      if (out != nullptr) out[row * ldo + c] = o;
      if (ex.out2 != nullptr) {
        if (ex.out2_bytes == 4)
          ((float*)ex.out2)[row * ex.ld2 + c] = (float)o;
        else
          ((double*)ex.out2)[row * ex.ld2 + c] = (double)o;
      }
      if (ex.xo != nullptr) {
        double* px = ex.xo + row * ex.ldxo + c;
        *px = ex.xmode == 2 ? *px + (double)o : (double)o;
      }
      sq[cc] = (double)o * (double)o;
    }
  });
  if (ex.nrm_part != nullptr) {  // column sums of squares of the 16 rows, in row order
    __shared__ double sqs[16 * 16 * NC];
    const int i = threadIdx.x >> 4, j = threadIdx.x & 15;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) sqs[(i * NC + cc) * 16 + j] = sq[cc];
    __syncthreads();
    if (threadIdx.x < 16 * NC) {
      const int cc = threadIdx.x >> 4, c = c0 + threadIdx.x;
      double a = sqs[cc * 16 + j];
      for (int r = 1; r < 16; ++r) a += sqs[(r * NC + cc) * 16 + j];
      if (c < nrhs)
        ex.nrm_part[(int64_t)c * gridDim.x * (kTileRows / 16) + blockIdx.x * (kTileRows / 16) + blockIdx.y] = a;
    }
  }
}

// Per (member, column) sums of squares from the per-workgroup column sums of
// a residual product (column-major: column c's n_wg = n_tiles kTileRows / 16
// sums at part + c n_wg): one workgroup per (member g, column c) adds the
// sums of member g's row tiles [first, first + count) (seg_tiles pairs), each
// thread a fixed stride of them, then a fixed-order tree.
__global__ __launch_bounds__(256) void rownorm_sum_kernel(const double* part, const int* seg_tiles, int n_wg, int nrhs,
                                                           double* out) {
  __shared__ double red[256];
  const int g = blockIdx.x / nrhs, c = blockIdx.x % nrhs;
  constexpr int G = kTileRows / 16;
  const double* p = part + (int64_t)c * n_wg + (int64_t)seg_tiles[2 * g] * G;
  const int n = seg_tiles[2 * g + 1] * G;
  double v = 0.;
  for (int i = threadIdx.x; i < n; i += 256) v += p[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Per (member, column) sums of squares, stage 1: workgroup (chunk j, member
// g) sums rows [kNormRows j, kNormRows (j + 1)) of the member, each thread
// over a fixed row stride of one column, the threads of a column added in
// order.
__global__ __launch_bounds__(256) void segnorm2_part_kernel(const double* r, int ldr, const int* seg, int nrhs,
                                                             int nchunks, double* part) {
  __shared__ double red[256];
  const int g = blockIdx.y, j = blockIdx.x;
  const int row0 = seg[2 * g], rows = seg[2 * g + 1];
  const int i0 = kNormRows * j, i1 = min(rows, i0 + kNormRows);
  double* o = part + ((int64_t)g * nchunks + j) * nrhs;
  for (int cb = 0; cb < nrhs; cb += 256) {
    const int nc = min(256, nrhs - cb), per = 256 / nc;
    const int c = (int)threadIdx.x % nc, p = (int)threadIdx.x / nc;
    double s = 0.;
    if (p < per)
      for (int i = i0 + p; i < i1; i += per) {
        const double v = r[(int64_t)(row0 + i) * ldr + cb + c];
        s += v * v;
      }
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < nc) {
      double a = red[threadIdx.x];
      for (int q = 1; q < per; ++q) a += red[q * nc + threadIdx.x];
      o[cb + threadIdx.x] = a;
    }
    __syncthreads();
  }
}

// stage 2: out[g nrhs + c] = the member's chunks added (ordered_sum4
// order), 64 (member, column) pairs per workgroup
__global__ __launch_bounds__(256) void segnorm2_sum_kernel(const double* part, int nseg, int nrhs, int nchunks,
                                                            double* out) {
  __shared__ double red[256];
  const int n = nseg * nrhs, e = 64 * blockIdx.x + (threadIdx.x & 63), ec = min(e, n - 1);
  const int g = ec / nrhs, c = ec % nrhs;
  const double v = ordered_sum4<double>(part + (int64_t)g * nchunks * nrhs + c, nrhs, nchunks, e < n, red);
  if (threadIdx.x < 64 && e < n) out[e] = v;
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
  const bool al = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0;
  const int64_t g4 = (n + 1023) / 1024;  // 256 threads x 4 elements per workgroup
  if (al && g4 <= 0x7fffffff) {
    // kernels, not memcpy nodes, so graphs capture them plainly
    if (dst_bytes == 4 && src_bytes == 8)
      hipLaunchKernelGGL((convert4_kernel<float, double>), dim3((unsigned)g4), dim3(256), 0, s, (float*)dst,
                         (const double*)src, n);
    else if (dst_bytes == 8 && src_bytes == 4)
      hipLaunchKernelGGL((convert4_kernel<double, float>), dim3((unsigned)g4), dim3(256), 0, s, (double*)dst,
                         (const float*)src, n);
    else if (dst_bytes == 8 && src_bytes == 8)
      hipLaunchKernelGGL((convert4_kernel<double, double>), dim3((unsigned)g4), dim3(256), 0, s, (double*)dst,
                         (const double*)src, n);
    else
      hipLaunchKernelGGL((convert4_kernel<float, float>), dim3((unsigned)g4), dim3(256), 0, s, (float*)dst,
                         (const float*)src, n);
    return hipGetLastError();
  }
  if (dst_bytes == 4 && src_bytes == 8)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(n)), dim3(256), 0, s, (float*)dst,
                       (const double*)src, n);
  else if (dst_bytes == 8 && src_bytes == 4)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0, s, (double*)dst,
                       (const float*)src, n);
  else if (dst_bytes == 8 && src_bytes == 8)
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

template <typename T, int MODE, int NC>
static void launch_hmv_nc(const MatvecPlanDev& p, const T* A, const T* x, int ldx, const T* y, int ldy, T* out,
                          int ldo, int nrhs, T* t, T sd, const HmvExtra& ex, hipStream_t s) {
  const int ncg = (nrhs + 63) / 64;
  if (p.n_lowrank > 0 && p.max_rank > 0) {
    T* tp = t + p.t_elems;
    if (p.n_parts > 0)
      hipLaunchKernelGGL((hmv_vtx_kernel<T, MODE, NC>), dim3(p.n_parts, (p.max_rank + 15) / 16, ncg), dim3(256), 0, s,
                         p.blocks, p.parts, A, x, ldx, nrhs, tp);
    hipLaunchKernelGGL((hmv_sum_kernel<T>), dim3(p.n_lowrank, (p.max_rank * nrhs + 63) / 64), dim3(256), 0, s,
                       p.blocks, p.lr_list, t, (const T*)tp, nrhs);
  }
  hipLaunchKernelGGL((hmv_rows_kernel<T, MODE, NC>), dim3(p.n_tiles, kTileRows / 16, ncg), dim3(256), 0, s, p.blocks,
                     p.tiles, p.contrib, A, x, ldx, (const T*)t, y, ldy, out, ldo, nrhs, sd, ex);
}

// Column groups of 64 (NC = 4) past 64 right-hand sides; below, NC covers
// the columns in 16-wide MFMA tiles.
template <typename T, int MODE>
static hipError_t launch_hmv_t(const MatvecPlanDev& p, const T* A, const T* x, int ldx, const T* y, int ldy, T* out,
                               int ldo, int nrhs, T* t, T sd, const HmvExtra& ex, hipStream_t s) {
  if (nrhs < 1 || p.n_tiles < 1) return nrhs < 1 ? hipErrorInvalidValue : hipSuccess;
  if (ex.out2 != nullptr && ex.out2_bytes != 4 && ex.out2_bytes != 8) return hipErrorInvalidValue;
  switch (nrhs > 64 ? 4 : (nrhs + 15) / 16) {
    case 1: launch_hmv_nc<T, MODE, 1>(p, A, x, ldx, y, ldy, out, ldo, nrhs, t, sd, ex, s); break;
    case 2: launch_hmv_nc<T, MODE, 2>(p, A, x, ldx, y, ldy, out, ldo, nrhs, t, sd, ex, s); break;
    case 3: launch_hmv_nc<T, MODE, 3>(p, A, x, ldx, y, ldy, out, ldo, nrhs, t, sd, ex, s); break;
    default: launch_hmv_nc<T, MODE, 4>(p, A, x, ldx, y, ldy, out, ldo, nrhs, t, sd, ex, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_hmatvec(const MatvecPlanDev& p, const double* A, const double* x, int ldx, const double* y, int ldy,
                          double* out, int ldo, int nrhs, double* t, const HmvExtra& ex, hipStream_t s) {
  return launch_hmv_t<double, MODE_FP64>(p, A, x, ldx, y, ldy, out, ldo, nrhs, t, -1.0, ex, s);
}

hipError_t launch_happly(const MatvecPlanDev& p, int elem_bytes, const void* A, const void* x, int ldx, void* out,
                         int ldo, int nrhs, void* t, const HmvExtra& ex, hipStream_t s) {
  if (elem_bytes == 8)
    return launch_hmv_t<double, MODE_FP64>(p, (const double*)A, (const double*)x, ldx, nullptr, 0, (double*)out, ldo,
                                           nrhs, (double*)t, 1.0, ex, s);
  if (elem_bytes == 4)
    return launch_hmv_t<float, MODE_FP32>(p, (const float*)A, (const float*)x, ldx, nullptr, 0, (float*)out, ldo, nrhs,
                                          (float*)t, 1.0f, ex, s);
  return hipErrorInvalidValue;
}

hipError_t launch_rownorm_sum(const double* part, const int* seg_tiles, int n_tiles, int nseg, int nrhs, double* out,
                              hipStream_t s) {
  if (nseg < 1 || nrhs < 1 || n_tiles < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rownorm_sum_kernel, dim3(nseg * nrhs), dim3(256), 0, s, part, seg_tiles,
                     n_tiles * (kTileRows / 16), nrhs, out);
  return hipGetLastError();
}

hipError_t launch_segnorm2(const double* r, int ldr, const int* seg, int nseg, int nrhs, int nchunks, double* part,
                           double* out, hipStream_t s) {
  if (nseg < 1 || nrhs < 1 || nchunks < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(segnorm2_part_kernel, dim3(nchunks, nseg), dim3(256), 0, s, r, ldr, seg, nrhs, nchunks, part);
  hipLaunchKernelGGL(segnorm2_sum_kernel, dim3((nseg * nrhs + 63) / 64), dim3(256), 0, s, (const double*)part, nseg,
                     nrhs, nchunks, out);
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
