// H-matrix construction on the device from a kernel function of points.
//
//   * dense leaves: every entry evaluated (one thread per entry, row-major,
//     coalesced stores);
//   * low-rank leaves: fixed-rank adaptive cross approximation with partial
//     pivoting, one workgroup per leaf.  Step k takes the residual row of the
//     current pivot row i, picks the column j* of largest magnitude, stores
//     V[:,k] = R(i,:) / R(i,j*), then U[:,k] = R(:,j*), and the next pivot
//     row is the unused row with the largest |U[t,k]|.  Ties go to the
//     smallest index; a residual row that is exactly zero is skipped for the
//     smallest unused row.  Residual entries are A(t,j) - sum_l U[t,l] V[j,l]
//     accumulated with one fused multiply-add per l in increasing l, so the
//     result is a fixed function of the kernel values (the CPU restatement
//     that checks it does the same operations).
//
// Kernel entries: kind 0: exp(-|x_i - x_j| / param) (1-D points), kind 1:
// 1 / (|p_i - p_j| + param) (2-D points), plus `shift` where the global row
// equals the global column; evaluated without contraction (the products and
// sums of the distance are rounded separately).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "kernels.h"

// No contraction anywhere in this file: every fused multiply-add is an
// explicit fma(), so the CPU restatement can repeat the arithmetic exactly.
#pragma clang fp contract(off)

namespace pp {

namespace {

template <int KIND>
__device__ inline double kernel_entry(const double* __restrict__ pts, int i, int j, double param, double shift) {
#pragma clang fp contract(off)
  double a;
  if (KIND == 0) {
    const double d = fabs(pts[i] - pts[j]);
    a = exp(-d / param);
  } else {
    const double dx = pts[2 * i] - pts[2 * j], dy = pts[2 * i + 1] - pts[2 * j + 1];
    const double s = dx * dx + dy * dy;
    a = 1.0 / (sqrt(s) + param);
  }
  if (i == j) a = a + shift;
  return a;
}

template <int KIND>
__global__ void __launch_bounds__(256) dense_fill_kernel(const BuildLeaf* __restrict__ leaves,
                                                         const double* __restrict__ pts, double param, double shift,
                                                         double* __restrict__ arena) {
  const BuildLeaf L = leaves[blockIdx.y];
  const int64_t n = (int64_t)L.ny * L.nx;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / L.nx), j = (int)(e - (int64_t)i * L.nx);
    arena[L.off + e] = kernel_entry<KIND>(pts, L.row0 + i, L.col0 + j, param, shift);
  }
}

struct amax {
  double m;  // |value|
  double v;  // value
  int idx;   // INT_MAX = none
};

__device__ inline bool better(const amax& a, const amax& b) { return a.m > b.m || (a.m == b.m && a.idx < b.idx); }

// Block-wide argmax (256 threads); every thread gets the result.
__device__ inline amax block_argmax(amax x, amax* red) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    amax y;
    y.m = __shfl_xor(x.m, o);
    y.v = __shfl_xor(x.v, o);
    y.idx = __shfl_xor(x.idx, o);
    if (better(y, x)) x = y;
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();  // red[] free from the previous use
  if ((threadIdx.x & 63) == 0) red[w] = x;
  __syncthreads();
  amax b = red[0];
#pragma unroll
  for (int q = 1; q < 4; ++q)
    if (better(red[q], b)) b = red[q];
  return b;
}

template <int KIND>
__global__ void __launch_bounds__(256) aca_kernel(const BuildLeaf* __restrict__ leaves, const double* __restrict__ pts,
                                                  double param, double shift, double* __restrict__ arena,
                                                  int32_t* __restrict__ pivots) {
  const BuildLeaf L = leaves[blockIdx.x];
  double* __restrict__ U = arena + L.off;
  double* __restrict__ V = arena + L.off_v;
  // rank r columns at row pitch ld (columns r..ld-1 are left zero)
  const int r = L.rank, ld = L.ld, ny = L.ny, nx = L.nx, tid = threadIdx.x;
  __shared__ double vec[kBuildMaxRank];
  __shared__ uint32_t used[kBuildMaxRows / 32];
  __shared__ amax red[4];
  __shared__ int s_next;
  for (int q = tid; q < (ny + 31) / 32; q += 256) used[q] = 0u;
  int32_t* piv = pivots ? pivots + L.aux : nullptr;
  int i = 0, k = 0;
  __syncthreads();
  while (k < r) {
    for (int l = tid; l < k; l += 256) vec[l] = U[(int64_t)i * ld + l];
    __syncthreads();
    // residual row i into V[:, k]
    amax best{-1.0, 0.0, INT_MAX};
    for (int j = tid; j < nx; j += 256) {
      double a = kernel_entry<KIND>(pts, L.row0 + i, L.col0 + j, param, shift);
      const double* vr = V + (int64_t)j * ld;
      for (int l = 0; l < k; ++l) a = fma(-vec[l], vr[l], a);
      V[(int64_t)j * ld + k] = a;
      const double m = fabs(a);
      if (m > best.m) best = amax{m, a, j};
    }
    const amax pc = block_argmax(best, red);
    if (tid == 0) used[i >> 5] |= 1u << (i & 31);
    __syncthreads();
    if (pc.m == 0.0) {
      // row i is represented exactly: retry with the smallest unused row
      if (tid == 0) {
        int nxt = -1;
        for (int q = 0; q < (ny + 31) / 32 && nxt < 0; ++q) {
          const uint32_t free_bits = ~used[q];
          if (free_bits != 0u) {
            const int t = q * 32 + __builtin_ctz(free_bits);
            if (t < ny) nxt = t;
          }
        }
        s_next = nxt;
      }
      __syncthreads();
      i = s_next;
      if (i < 0) break;
      continue;
    }
    const int jp = pc.idx;
    const double d = pc.v;
    for (int j = tid; j < nx; j += 256) V[(int64_t)j * ld + k] = V[(int64_t)j * ld + k] / d;
    for (int l = tid; l < k; l += 256) vec[l] = V[(int64_t)jp * ld + l];
    __syncthreads();
    // residual column j* into U[:, k]; next pivot row among the unused ones
    best = amax{-1.0, 0.0, INT_MAX};
    for (int t = tid; t < ny; t += 256) {
      double a = kernel_entry<KIND>(pts, L.row0 + t, L.col0 + jp, param, shift);
      const double* ur = U + (int64_t)t * ld;
      for (int l = 0; l < k; ++l) a = fma(-ur[l], vec[l], a);
      U[(int64_t)t * ld + k] = a;
      const double m = fabs(a);
      if (m > best.m && !((used[t >> 5] >> (t & 31)) & 1u)) best = amax{m, a, t};
    }
    const amax pr = block_argmax(best, red);
    if (piv && tid == 0) {
      piv[2 * k] = i;
      piv[2 * k + 1] = jp;
    }
    ++k;
    i = pr.idx == INT_MAX ? -1 : pr.idx;
    if (i < 0) break;
    __syncthreads();
  }
  // columns the rows ran out for stay zero
  for (int c = k; c < ld; ++c) {
    for (int t = tid; t < ny; t += 256) U[(int64_t)t * ld + c] = 0.0;
    for (int j = tid; j < nx; j += 256) V[(int64_t)j * ld + c] = 0.0;
    if (piv && tid == 0 && c < r) {
      piv[2 * c] = -1;
      piv[2 * c + 1] = -1;
    }
  }
}

// ---- truncation of a rank-k cross approximation to rank r (one workgroup
// per leaf).  X = U_k (then V_k), m x kp, pitch kp, columns >= k zero:
//   1. columns scaled to unit norm (norms d);
//   2. shifted CholeskyQR3: three times G = X^T X (the first with the shift
//      11 (m kp + kp (kp + 1)) eps tr(G) on the diagonal; zero columns get a
//      unit diagonal), G = R^T R, X <- X R^{-1}, Rt <- R Rt; so that
//      U_k = Q_u (Rt_u D_u), V_k = Q_v (Rt_v D_v);
//   3. core M = (Rt_u D_u)(Rt_v D_v)^T (kp x kp), one-sided Jacobi M J = N
//      (cyclic round-robin pairs, rotation when |g| > 1e-15 sqrt(a b), at
//      most 30 sweeps);
//   4. the r columns of N of largest norm (first on ties): U = Q_u N_sel,
//      V = Q_v J_sel.
// Every sum runs in increasing index order by one thread.

// G = X^T X, full kp x kp, 8 x 8 tiles per thread.
__device__ void gram(const double* __restrict__ X, int m, int kp, double* __restrict__ G) {
  const int nt = kp / 8;
  for (int tt = threadIdx.x; tt < nt * nt; tt += 256) {
    const int a0 = (tt / nt) * 8, b0 = (tt % nt) * 8;
    double acc[8][8];
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
      for (int f = 0; f < 8; ++f) acc[e][f] = 0.0;
    for (int t = 0; t < m; ++t) {
      const double* row = X + (int64_t)t * kp;
      double xa[8], xb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        xa[e] = row[a0 + e];
        xb[e] = row[b0 + e];
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
#pragma unroll
        for (int f = 0; f < 8; ++f) acc[e][f] = fma(xa[e], xb[f], acc[e][f]);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
      for (int f = 0; f < 8; ++f) G[(a0 + e) * kp + b0 + f] = acc[e][f];
  }
}

// X (m x kp) = Q R with the accumulated R in Rt (upper); d: column norms.
// Returns false when a Cholesky pivot is not positive.
__device__ bool orthonormalize(double* __restrict__ X, int m, int kp, double* __restrict__ d,
                               double* __restrict__ G, double* __restrict__ Rt, double* __restrict__ T,
                               int* s_fail) {
  const int tid = threadIdx.x;
  if (tid < kp) {
    double acc = 0.0;
    for (int t = 0; t < m; ++t) acc = fma(X[(int64_t)t * kp + tid], X[(int64_t)t * kp + tid], acc);
    d[tid] = sqrt(acc);
  }
  __syncthreads();
  for (int t = tid; t < m; t += 256)
    for (int c = 0; c < kp; ++c)
      if (d[c] > 0.0) X[(int64_t)t * kp + c] = X[(int64_t)t * kp + c] / d[c];
  for (int e = tid; e < kp * kp; e += 256) Rt[e] = (e / kp == e % kp) ? 1.0 : 0.0;
  if (tid == 0) *s_fail = 0;
  __syncthreads();
  for (int pass = 0; pass < 3; ++pass) {
    gram(X, m, kp, G);
    __syncthreads();
    if (tid == 0) {
      double tr = 0.0;
      for (int c = 0; c < kp; ++c) tr = tr + G[c * kp + c];
      const double shift = pass == 0 ? 11.0 * ((double)m * kp + (double)kp * (kp + 1)) * 0x1.0p-53 * tr : 0.0;
      for (int c = 0; c < kp; ++c) G[c * kp + c] = d[c] > 0.0 ? G[c * kp + c] + shift : 1.0;
    }
    __syncthreads();
    // right-looking Cholesky G = R^T R, R in the upper triangle
    for (int j = 0; j < kp; ++j) {
      const double piv = G[j * kp + j];
      if (!(piv > 0.0)) {
        if (tid == 0) *s_fail = 1;
        return false;  // uniform: every thread read the same pivot
      }
      const double rjj = sqrt(piv);
      __syncthreads();
      for (int c = j + tid; c < kp; c += 256) G[j * kp + c] = c == j ? rjj : G[j * kp + c] / rjj;
      __syncthreads();
      const int w = kp - j - 1;
      for (int e = tid; e < w * w; e += 256) {
        const int a = j + 1 + e / w, c = j + 1 + e % w;
        if (c >= a) G[a * kp + c] = fma(-G[j * kp + a], G[j * kp + c], G[a * kp + c]);
      }
      __syncthreads();
    }
    // X <- X R^{-1}, row by row
    for (int t = tid; t < m; t += 256) {
      double* x = X + (int64_t)t * kp;
      for (int c = 0; c < kp; ++c) {
        double v = x[c];
        for (int a = 0; a < c; ++a) v = fma(-x[a], G[a * kp + c], v);
        x[c] = v / G[c * kp + c];
      }
    }
    // Rt <- R Rt (both upper)
    for (int e = tid; e < kp * kp; e += 256) {
      const int a = e / kp, c = e % kp;
      double v = 0.0;
      for (int b = a; b <= c; ++b) v = fma(G[a * kp + b], Rt[b * kp + c], v);
      T[e] = c >= a ? v : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < kp * kp; e += 256) Rt[e] = T[e];
    __syncthreads();
  }
  return true;
}

__global__ void __launch_bounds__(256) recompress_kernel(RecompressLeaf* __restrict__ leaves,
                                                         double* __restrict__ scratch, double* __restrict__ arena) {
  RecompressLeaf* L = leaves + blockIdx.x;
  const int ny = L->ny, nx = L->nx, kp = L->kp, r = L->r, tid = threadIdx.x;
  double* Qu = scratch + L->u;
  double* Qv = scratch + L->v;
  double* W = scratch + L->work;
  double *Ru = W, *Rv = W + kp * kp, *G = W + 2 * kp * kp, *T = W + 3 * kp * kp, *N = W + 4 * kp * kp,
         *J = W + 5 * kp * kp, *du = W + 6 * kp * kp, *dv = du + kp;
  __shared__ int s_fail;
  __shared__ int s_any;
  __shared__ double sig[kBuildMaxRank];
  __shared__ int sel[kBuildMaxRank];
  if (!orthonormalize(Qu, ny, kp, du, G, Ru, T, &s_fail) || !orthonormalize(Qv, nx, kp, dv, G, Rv, T, &s_fail)) {
    if (tid == 0) L->status = 1;
    return;
  }
  // N = (Ru Du)(Rv Dv)^T, column-major (column b at N + b kp); J = I
  for (int e = tid; e < kp * kp; e += 256) {
    const int b = e / kp, a = e % kp;
    double v = 0.0;
    for (int c = a > b ? a : b; c < kp; ++c) v = fma(Ru[a * kp + c] * du[c], Rv[b * kp + c] * dv[c], v);
    N[e] = v;
    J[e] = a == b ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 30; ++sweep) {
    if (tid == 0) s_any = 0;
    __syncthreads();
    for (int rd = 0; rd < kp - 1; ++rd) {
      if (tid < kp / 2) {
        const int i0 = tid, i1 = kp - 1 - tid;
        const int p0 = i0 == 0 ? 0 : 1 + (i0 - 1 + rd) % (kp - 1);
        const int p1 = 1 + (i1 - 1 + rd) % (kp - 1);
        const int p = p0 < p1 ? p0 : p1, q = p0 < p1 ? p1 : p0;
        double* mp = N + p * kp;
        double* mq = N + q * kp;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = 0; i < kp; ++i) {
          al = fma(mp[i], mp[i], al);
          be = fma(mq[i], mq[i], be);
          ga = fma(mp[i], mq[i], ga);
        }
        if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
          double* jp = J + p * kp;
          double* jq = J + q * kp;
          for (int i = 0; i < kp; ++i) {
            const double x = mp[i], y = mq[i];
            mp[i] = c * x - sn * y;
            mq[i] = sn * x + c * y;
            const double u = jp[i], v = jq[i];
            jp[i] = c * u - sn * v;
            jq[i] = sn * u + c * v;
          }
          s_any = 1;
        }
      }
      __syncthreads();
    }
    if (!s_any) break;
    __syncthreads();
  }
  if (tid < kp) {
    double acc = 0.0;
    for (int i = 0; i < kp; ++i) acc = fma(N[tid * kp + i], N[tid * kp + i], acc);
    sig[tid] = sqrt(acc);
  }
  __syncthreads();
  if (tid == 0) {
    for (int o = 0; o < r; ++o) {
      int b = -1;
      for (int c = 0; c < kp; ++c)
        if (sig[c] >= 0.0 && (b < 0 || sig[c] > sig[b])) b = c;
      sel[o] = b;
      if (b >= 0) sig[b] = -1.0;
    }
  }
  __syncthreads();
  double* Uo = arena + L->out_u;
  double* Vo = arena + L->out_v;
  for (int t = tid; t < ny; t += 256)
    for (int o = 0; o < r; ++o) {
      const int b = sel[o];
      double v = 0.0;
      if (b >= 0)
        for (int a = 0; a < kp; ++a) v = fma(Qu[(int64_t)t * kp + a], N[b * kp + a], v);
      Uo[(int64_t)t * r + o] = v;
    }
  for (int t = tid; t < nx; t += 256)
    for (int o = 0; o < r; ++o) {
      const int b = sel[o];
      double v = 0.0;
      if (b >= 0)
        for (int a = 0; a < kp; ++a) v = fma(Qv[(int64_t)t * kp + a], J[b * kp + a], v);
      Vo[(int64_t)t * r + o] = v;
    }
  if (tid == 0) L->status = 0;
}

}  // namespace

hipError_t launch_recompress(RecompressLeaf* leaves, int n, double* scratch, double* arena, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(recompress_kernel, dim3(n), dim3(256), 0, s, leaves, scratch, arena);
  return hipGetLastError();
}

hipError_t launch_build(int kind, const BuildLeaf* dense, int n_dense, int max_dense_elems, const BuildLeaf* lowrank,
                        int n_lowrank, const double* pts, double param, double shift, double* arena,
                        int32_t* pivots, hipStream_t s) {
  if (n_dense > 0) {
    const int bx = std::max(1, std::min(256, (max_dense_elems + 255) / 256));
    dim3 grid(bx, n_dense);
    if (kind == 0)
      hipLaunchKernelGGL(dense_fill_kernel<0>, grid, dim3(256), 0, s, dense, pts, param, shift, arena);
    else
      hipLaunchKernelGGL(dense_fill_kernel<1>, grid, dim3(256), 0, s, dense, pts, param, shift, arena);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (n_lowrank > 0) {
    if (kind == 0)
      hipLaunchKernelGGL(aca_kernel<0>, dim3(n_lowrank), dim3(256), 0, s, lowrank, pts, param, shift, arena, pivots);
    else
      hipLaunchKernelGGL(aca_kernel<1>, dim3(n_lowrank), dim3(256), 0, s, lowrank, pts, param, shift, arena, pivots);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace pp
