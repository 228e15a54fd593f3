// H-matrix construction on the device from a kernel function of points.
//
//   * dense leaves: every entry evaluated (one thread per entry, row-major,
//     coalesced stores);
//   * low-rank leaves: partially pivoted adaptive cross approximation of rank
//     k, then (k > r) truncation to the leaf's rank r.
//
// Cross approximation, step s: the residual row of pivot row i is stored as
// V[:, s] and its column j* of largest magnitude picked; V[:, s] /= R(i, j*);
// then U[:, s] = R(:, j*) and the next pivot row is the unused row with the
// largest |U[t, s]|.  Ties go to the smallest index.  A residual row that is
// exactly zero leaves step s empty (zero columns) and the next step starts at
// the smallest unused row.  Residual entries are A(t, j) - sum_l U[t, l]
// V[j, l], one fused multiply-add per l in increasing l.  Every step is five
// launches over all leaves at once (row residual, row pick, normalisation,
// column residual, column pick); each leaf's rows and columns are cut into
// 256-element slices, one workgroup each, so a large leaf spreads over many
// compute units.
//
// Truncation (pitch kp, a multiple of 32): unit columns; shifted CholeskyQR3
// of U_k and V_k (column norms and Gram matrices summed per 256-row slice,
// the slices then added in order; X <- X R^{-1} with R^{-1} formed
// explicitly); one-sided Jacobi SVD of the kp x kp core (cyclic round-robin
// pairs, each pair's dot products as 64 lane sums combined by an xor
// butterfly, rotation when |g| > 1e-15 sqrt(a b), sweeps until none rotates,
// at most 30); the r columns of largest norm.
//
// Kernel entries: kind 0: exp(-|x_i - x_j| / param) (1-D points), kind 1:
// 1 / (|p_i - p_j| + param) (2-D points), plus `shift` where the global row
// equals the global column.  No contraction anywhere in this file: every
// fused multiply-add is an explicit fma(), so the CPU restatement repeats the
// arithmetic exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "kernels.h"

#pragma clang fp contract(off)

namespace pp {

namespace {

template <int KIND>
__device__ inline double kernel_entry(const double* __restrict__ pts, int i, int j, double param, double shift) {
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
__device__ inline double entry(const BuildArgs& A, int i, int j) {
  return kernel_entry<KIND>(A.pts, i, j, A.param, A.shift);
}

template <int KIND>
__global__ void __launch_bounds__(256) dense_fill_kernel(const BuildLeaf* __restrict__ leaves, BuildArgs A,
                                                         double* __restrict__ arena) {
  const BuildLeaf L = leaves[blockIdx.y];
  const int64_t n = (int64_t)L.ny * L.nx;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / L.nx), j = (int)(e - (int64_t)i * L.nx);
    arena[L.off + e] = entry<KIND>(A, L.row0 + i, L.col0 + j);
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
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = x;
  __syncthreads();
  amax b = red[0];
#pragma unroll
  for (int q = 1; q < 4; ++q)
    if (better(red[q], b)) b = red[q];
  return b;
}

__device__ inline bool active(const AcaLeaf& L, const AcaState& S, int step) { return step < L.k && S.i >= 0; }

// ---- cross approximation, one step per five launches

// residual row i into V[:, step] over a column slice; slice argmax
template <int KIND>
__global__ void __launch_bounds__(256) aca_row_kernel(const AcaLeaf* __restrict__ leaves, AcaState* __restrict__ st,
                                                      const BuildSlice* __restrict__ slices, BuildPartial* part,
                                                      BuildArgs A, int step) {
  const BuildSlice sl = slices[blockIdx.x];
  const AcaLeaf L = leaves[sl.leaf];
  const AcaState S = st[sl.leaf];
  __shared__ double vec[kBuildMaxRank];
  __shared__ amax red[4];
  if (!active(L, S, step)) return;
  const int i = S.i, tid = threadIdx.x;
  for (int l = tid; l < step; l += 256) vec[l] = L.U[(int64_t)i * L.ld + l];
  __syncthreads();
  amax best{-1.0, 0.0, INT_MAX};
  if (tid < sl.count) {
    const int j = sl.first + tid;
    double a = entry<KIND>(A, L.row0 + i, L.col0 + j);
    const double* vr = L.V + (int64_t)j * L.ld;
    for (int l = 0; l < step; ++l) a = fma(-vec[l], vr[l], a);
    L.V[(int64_t)j * L.ld + step] = a;
    best = amax{fabs(a), a, j};
  }
  const amax b = block_argmax(best, red);
  if (tid == 0) part[sl.part] = BuildPartial{b.m, b.v, b.idx, 0};
}

// column pivot of each leaf; an exactly zero row leaves the step empty
__global__ void __launch_bounds__(64) aca_pick_row_kernel(const AcaLeaf* __restrict__ leaves, AcaState* st,
                                                          const BuildPartial* __restrict__ part, int step) {
  const AcaLeaf L = leaves[blockIdx.x];
  AcaState S = st[blockIdx.x];
  if (!active(L, S, step) || threadIdx.x != 0) return;
  amax b{-1.0, 0.0, INT_MAX};
  for (int q = 0; q < L.ncol; ++q) {
    const BuildPartial p = part[L.col_part + q];
    const amax c{p.m, p.v, p.idx};
    if (better(c, b)) b = c;
  }
  L.used[S.i >> 5] |= 1u << (S.i & 31);
  if (b.m == 0.0) {
    S.zero = 1;
    int nxt = -1;
    for (int q = 0; q < (L.ny + 31) / 32 && nxt < 0; ++q) {
      const uint32_t free_bits = ~L.used[q];
      if (free_bits != 0u) {
        const int t = q * 32 + __builtin_ctz(free_bits);
        if (t < L.ny) nxt = t;
      }
    }
    S.next = nxt;
  } else {
    S.zero = 0;
    S.jp = b.idx;
    S.d = b.v;
  }
  st[blockIdx.x] = S;
}

// V[:, step] /= R(i, j*) over a column slice
__global__ void __launch_bounds__(256) aca_scale_kernel(const AcaLeaf* __restrict__ leaves,
                                                        const AcaState* __restrict__ st,
                                                        const BuildSlice* __restrict__ slices, int step) {
  const BuildSlice sl = slices[blockIdx.x];
  const AcaLeaf L = leaves[sl.leaf];
  const AcaState S = st[sl.leaf];
  if (!active(L, S, step) || S.zero || (int)threadIdx.x >= sl.count) return;
  double* v = L.V + (int64_t)(sl.first + threadIdx.x) * L.ld + step;
  *v = *v / S.d;
}

// residual column j* into U[:, step] over a row slice; argmax over the
// slice's unused rows
template <int KIND>
__global__ void __launch_bounds__(256) aca_col_kernel(const AcaLeaf* __restrict__ leaves,
                                                      const AcaState* __restrict__ st,
                                                      const BuildSlice* __restrict__ slices, BuildPartial* part,
                                                      BuildArgs A, int step) {
  const BuildSlice sl = slices[blockIdx.x];
  const AcaLeaf L = leaves[sl.leaf];
  const AcaState S = st[sl.leaf];
  __shared__ double vec[kBuildMaxRank];
  __shared__ amax red[4];
  if (!active(L, S, step)) return;
  const int tid = threadIdx.x;
  if (S.zero) {
    if (tid < sl.count) L.U[(int64_t)(sl.first + tid) * L.ld + step] = 0.0;
    return;
  }
  const int jp = S.jp;
  for (int l = tid; l < step; l += 256) vec[l] = L.V[(int64_t)jp * L.ld + l];
  __syncthreads();
  amax best{-1.0, 0.0, INT_MAX};
  if (tid < sl.count) {
    const int t = sl.first + tid;
    double a = entry<KIND>(A, L.row0 + t, L.col0 + jp);
    const double* ur = L.U + (int64_t)t * L.ld;
    for (int l = 0; l < step; ++l) a = fma(-ur[l], vec[l], a);
    L.U[(int64_t)t * L.ld + step] = a;
    if (!((L.used[t >> 5] >> (t & 31)) & 1u)) best = amax{fabs(a), a, t};
  }
  const amax b = block_argmax(best, red);
  if (tid == 0) part[sl.part] = BuildPartial{b.m, b.v, b.idx, 0};
}

// next pivot row of each leaf; the pivot pair of the step
__global__ void __launch_bounds__(64) aca_pick_col_kernel(const AcaLeaf* __restrict__ leaves, AcaState* st,
                                                          const BuildPartial* __restrict__ part, int step) {
  const AcaLeaf L = leaves[blockIdx.x];
  AcaState S = st[blockIdx.x];
  if (!active(L, S, step) || threadIdx.x != 0) return;
  if (S.zero) {
    S.i = S.next;
  } else {
    amax b{-1.0, 0.0, INT_MAX};
    for (int q = 0; q < L.nrow; ++q) {
      const BuildPartial p = part[L.row_part + q];
      const amax c{p.m, p.v, p.idx};
      if (better(c, b)) b = c;
    }
    if (L.piv) {
      L.piv[2 * step] = S.i;
      L.piv[2 * step + 1] = S.jp;
    }
    S.i = b.idx == INT_MAX ? -1 : b.idx;
  }
  st[blockIdx.x] = S;
}

// zero U, V (pitch ld, all ld columns) and the pivots before the steps
__global__ void __launch_bounds__(256) aca_clear_kernel(const AcaLeaf* __restrict__ leaves,
                                                        const BuildSlice* __restrict__ slices, int rows) {
  const BuildSlice sl = slices[blockIdx.x];
  const AcaLeaf L = leaves[sl.leaf];
  double* X = rows ? L.U : L.V;
  for (int e = threadIdx.x; e < sl.count * L.ld; e += 256) X[(int64_t)sl.first * L.ld + e] = 0.0;
  if (rows && sl.first == 0 && L.piv)
    for (int e = threadIdx.x; e < 2 * L.k; e += 256) L.piv[e] = -1;
}

// ---- truncation

// partial column sums of squares of a factor slice (thread c: column c)
__global__ void __launch_bounds__(256) tr_colsq_kernel(const TruncSlice* __restrict__ slices) {
  const TruncSlice sl = slices[blockIdx.x];
  const int c = threadIdx.x;
  if (c >= sl.kp) return;
  double acc = 0.0;
  for (int t = 0; t < sl.count; ++t) {
    const double x = sl.X[(int64_t)(sl.first + t) * sl.kp + c];
    acc = fma(x, x, acc);
  }
  sl.part[c] = acc;
}

// column norms of each factor (sum of the slices in order)
__global__ void __launch_bounds__(256) tr_norm_kernel(const TruncFactor* __restrict__ fs) {
  const TruncFactor F = fs[blockIdx.x];
  const int c = threadIdx.x;
  if (c >= F.kp) return;
  double tot = 0.0;
  for (int q = 0; q < F.nslices; ++q) tot = tot + F.parts[(int64_t)q * F.part_stride + c];
  F.d[c] = sqrt(tot);
}

__global__ void __launch_bounds__(256) tr_scale_kernel(const TruncSlice* __restrict__ slices) {
  const TruncSlice sl = slices[blockIdx.x];
  if ((int)threadIdx.x >= sl.count) return;
  double* x = sl.X + (int64_t)(sl.first + threadIdx.x) * sl.kp;
  for (int c = 0; c < sl.kp; ++c)
    if (sl.d[c] > 0.0) x[c] = x[c] / sl.d[c];
}

// partial Gram matrix of a slice (8 x 8 tiles per thread)
__global__ void __launch_bounds__(256) tr_gram_kernel(const TruncSlice* __restrict__ slices) {
  const TruncSlice sl = slices[blockIdx.x];
  const int kp = sl.kp, nt = kp / 8;
  const double* X = sl.X + (int64_t)sl.first * kp;
  for (int tt = threadIdx.x; tt < nt * nt; tt += 256) {
    const int a0 = (tt / nt) * 8, b0 = (tt % nt) * 8;
    double acc[8][8];
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
      for (int f = 0; f < 8; ++f) acc[e][f] = 0.0;
    for (int t = 0; t < sl.count; ++t) {
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
      for (int f = 0; f < 8; ++f) sl.gpart[(a0 + e) * kp + b0 + f] = acc[e][f];
  }
}

// G = sum of the slice Gram matrices in slice order (grid: factor x 256 entries)
__global__ void __launch_bounds__(256) tr_gsum_kernel(TruncFactor* fs) {
  TruncFactor* F = fs + blockIdx.x;
  const int kp = F->kp, e = blockIdx.y * 256 + threadIdx.x;
  if (F->status || e >= kp * kp) return;
  double tot = 0.0;
  for (int q = 0; q < F->nslices; ++q) tot = tot + F->gparts[(int64_t)q * F->gpart_stride + e];
  F->G[e] = tot;
}

// per factor (G summed by tr_gsum_kernel): shift (first pass),
// Cholesky G = R^T R (R in the upper triangle of G), R^{-1} (Rt <- R Rt
// follows in tr_rprod_kernel)
__global__ void __launch_bounds__(256) tr_chol_kernel(TruncFactor* fs, int pass) {
  TruncFactor* F = fs + blockIdx.x;
  const int kp = F->kp, tid = threadIdx.x;
  double *G = F->G, *Rt = F->Rt;
  const double* d = F->d;
  if (F->status) return;
  if (pass == 0)
    for (int e = tid; e < kp * kp; e += 256) Rt[e] = (e / kp == e % kp) ? 1.0 : 0.0;
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0;
    for (int c = 0; c < kp; ++c) tr = tr + G[c * kp + c];
    const double shift = pass == 0 ? 11.0 * ((double)F->m * kp + (double)kp * (kp + 1)) * 0x1.0p-53 * tr : 0.0;
    for (int c = 0; c < kp; ++c) G[c * kp + c] = d[c] > 0.0 ? G[c * kp + c] + shift : 1.0;
  }
  __syncthreads();
  for (int j = 0; j < kp; ++j) {
    const double piv = G[j * kp + j];
    if (!(piv > 0.0)) {
      if (tid == 0) F->status = 1;
      return;  // uniform: every thread read the same pivot
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
  // R^{-1} column by column (thread c): y[c] = 1 / R[c][c], then upward
  double* Ri = F->Ri;
  if (tid < kp) {
    const int c = tid;
    for (int a = c + 1; a < kp; ++a) Ri[a * kp + c] = 0.0;
    Ri[c * kp + c] = 1.0 / G[c * kp + c];
    for (int a = c - 1; a >= 0; --a) {
      double v = 0.0;
      for (int b = a + 1; b <= c; ++b) v = fma(G[a * kp + b], Ri[b * kp + c], v);
      Ri[a * kp + c] = -v / G[a * kp + a];
    }
  }
}

// T = R Rt (both upper; grid: factor x 256 entries), then Rt = T
__global__ void __launch_bounds__(256) tr_rprod_kernel(TruncFactor* fs) {
  const TruncFactor& F = fs[blockIdx.x];
  const int kp = F.kp, e = blockIdx.y * 256 + threadIdx.x;
  if (F.status || e >= kp * kp) return;
  const int a = e / kp, c = e % kp;
  double v = 0.0;
  for (int b = a; b <= c; ++b) v = fma(F.G[a * kp + b], F.Rt[b * kp + c], v);
  F.T[e] = c >= a ? v : 0.0;
}

__global__ void __launch_bounds__(256) tr_rcopy_kernel(TruncFactor* fs) {
  const TruncFactor& F = fs[blockIdx.x];
  const int kp = F.kp, e = blockIdx.y * 256 + threadIdx.x;
  if (F.status || e >= kp * kp) return;
  F.Rt[e] = F.T[e];
}

// rows of a slice <- row R^{-1}
__global__ void __launch_bounds__(256) tr_trsm_kernel(const TruncSlice* __restrict__ slices,
                                                      const TruncFactor* __restrict__ fs) {
  const TruncSlice sl = slices[blockIdx.x];
  const TruncFactor& F = fs[sl.factor];
  if (F.status || (int)threadIdx.x >= sl.count) return;
  const int kp = sl.kp;
  const double* Ri = F.Ri;
  double* x = sl.X + (int64_t)(sl.first + threadIdx.x) * kp;
  // x <- x R^{-1}: column c sums x[a] Ri[a][c] over a <= c; blocks of 8
  // columns from the last one down, so the block written is never read again
  for (int c0 = kp - 8; c0 >= 0; c0 -= 8) {
    double acc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = 0.0;
    for (int a = 0; a < c0 + 8; ++a) {
      const double xa = x[a];
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (a <= c0 + e) acc[e] = fma(xa, Ri[a * kp + c0 + e], acc[e]);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) x[c0 + e] = acc[e];
  }
}

// per leaf: core N = (Ru Du)(Rv Dv)^T (column-major) and J = I
__global__ void __launch_bounds__(256) tr_core_kernel(TruncLeaf* tl, const TruncFactor* __restrict__ fs) {
  TruncLeaf* L = tl + blockIdx.x;
  const TruncFactor& Fu = fs[2 * blockIdx.x];
  const TruncFactor& Fv = fs[2 * blockIdx.x + 1];
  const int kp = L->kp, tid = threadIdx.x;
  if (Fu.status || Fv.status) {
    if (tid == 0) L->status = 1;
    return;
  }
  const double *Ru = Fu.Rt, *Rv = Fv.Rt, *du = Fu.d, *dv = Fv.d;
  double *N = L->N, *J = L->J;
  for (int e = tid; e < kp * kp; e += 256) {
    const int b = e / kp, a = e % kp;
    double v = 0.0;
    for (int c = a > b ? a : b; c < kp; ++c) v = fma(Ru[a * kp + c] * du[c], Rv[b * kp + c] * dv[c], v);
    N[e] = v;
    J[e] = a == b ? 1.0 : 0.0;
  }
}

// sum over the 64 lanes of a wave: xor butterfly (every lane ends with the
// same value: each step adds two lanes' values in either order)
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o);
  return v;
}

// one round of the cyclic one-sided Jacobi sweep for every leaf: pair i0 of
// round rd on one wave; lane l sums elements l, l + 64, ... in order
__global__ void __launch_bounds__(256) tr_jacobi_kernel(TruncLeaf* tl, int rd) {
  TruncLeaf* L = tl + blockIdx.x;
  const int kp = L->kp, lane = threadIdx.x & 63;
  const int i0 = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (L->status || rd >= kp - 1 || i0 >= kp / 2) return;
  const int i1 = kp - 1 - i0;
  const int p0 = i0 == 0 ? 0 : 1 + (i0 - 1 + rd) % (kp - 1);
  const int p1 = 1 + (i1 - 1 + rd) % (kp - 1);
  const int p = p0 < p1 ? p0 : p1, q = p0 < p1 ? p1 : p0;
  double* mp = L->N + p * kp;
  double* mq = L->N + q * kp;
  double al = 0.0, be = 0.0, ga = 0.0;
  for (int i = lane; i < kp; i += 64) {
    al = fma(mp[i], mp[i], al);
    be = fma(mq[i], mq[i], be);
    ga = fma(mp[i], mq[i], ga);
  }
  al = wave_sum(al);
  be = wave_sum(be);
  ga = wave_sum(ga);
  if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
    const double zeta = (be - al) / (2.0 * ga);
    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
    double* jp = L->J + p * kp;
    double* jq = L->J + q * kp;
    for (int i = lane; i < kp; i += 64) {
      const double x = mp[i], y = mq[i];
      mp[i] = c * x - sn * y;
      mq[i] = sn * x + c * y;
      const double u = jp[i], v = jq[i];
      jp[i] = c * u - sn * v;
      jq[i] = sn * u + c * v;
    }
    if (lane == 0) L->rotated = 1;
  }
}

// per leaf: the r columns of N of largest norm
__global__ void __launch_bounds__(256) tr_select_kernel(TruncLeaf* tl) {
  TruncLeaf* L = tl + blockIdx.x;
  const int kp = L->kp, r = L->r, tid = threadIdx.x;
  if (L->status) return;
  __shared__ double sig[kBuildMaxRank];
  if (tid < kp) {
    double acc = 0.0;
    for (int i = 0; i < kp; ++i) acc = fma(L->N[tid * kp + i], L->N[tid * kp + i], acc);
    sig[tid] = sqrt(acc);
  }
  __syncthreads();
  if (tid == 0) {
    for (int o = 0; o < r; ++o) {
      int b = -1;
      for (int c = 0; c < kp; ++c)
        if (sig[c] >= 0.0 && (b < 0 || sig[c] > sig[b])) b = c;
      L->sel[o] = b;
      if (b >= 0) sig[b] = -1.0;
    }
  }
}

__global__ void tr_clear_rotated_kernel(TruncLeaf* tl, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tl[i].rotated = 0;
}

__global__ void tr_any_rotated_kernel(const TruncLeaf* tl, int n, int* any) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int a = 0;
  for (int i = 0; i < n; ++i) a |= tl[i].rotated;
  *any = a;
}

// output rows of a slice: U = Q_u N_sel (factor 2l), V = Q_v J_sel (2l + 1)
__global__ void __launch_bounds__(256) tr_out_kernel(const TruncSlice* __restrict__ slices,
                                                     const TruncLeaf* __restrict__ tl) {
  const TruncSlice sl = slices[blockIdx.x];
  const TruncLeaf& L = tl[sl.factor >> 1];
  if (L.status || (int)threadIdx.x >= sl.count) return;
  const int kp = sl.kp, r = L.r, t = sl.first + threadIdx.x;
  const double* x = sl.X + (int64_t)t * kp;
  const double* M = (sl.factor & 1) ? L.J : L.N;
  double* out = ((sl.factor & 1) ? L.out_v : L.out_u) + (int64_t)t * r;
  for (int o = 0; o < r; ++o) {
    const int b = L.sel[o];
    double v = 0.0;
    if (b >= 0)
      for (int a = 0; a < kp; ++a) v = fma(x[a], M[b * kp + a], v);
    out[o] = v;
  }
}

}  // namespace

hipError_t launch_dense_fill(int kind, const BuildLeaf* dense, int n_dense, int max_elems, const BuildArgs& a,
                             double* arena, hipStream_t s) {
  if (n_dense <= 0) return hipSuccess;
  dim3 grid(std::max(1, std::min(256, (max_elems + 255) / 256)), n_dense);
  if (kind == 0)
    hipLaunchKernelGGL(dense_fill_kernel<0>, grid, dim3(256), 0, s, dense, a, arena);
  else
    hipLaunchKernelGGL(dense_fill_kernel<1>, grid, dim3(256), 0, s, dense, a, arena);
  return hipGetLastError();
}

hipError_t launch_aca(int kind, const AcaLeaf* leaves, AcaState* st, int n_leaves, int max_k,
                      const BuildSlice* col_slices, int n_col, const BuildSlice* row_slices, int n_row,
                      BuildPartial* part, const BuildArgs& a, hipStream_t s) {
  if (n_leaves <= 0) return hipSuccess;
  hipLaunchKernelGGL(aca_clear_kernel, dim3(n_row), dim3(256), 0, s, leaves, row_slices, 1);
  hipLaunchKernelGGL(aca_clear_kernel, dim3(n_col), dim3(256), 0, s, leaves, col_slices, 0);
  for (int step = 0; step < max_k; ++step) {
    if (kind == 0)
      hipLaunchKernelGGL(aca_row_kernel<0>, dim3(n_col), dim3(256), 0, s, leaves, st, col_slices, part, a, step);
    else
      hipLaunchKernelGGL(aca_row_kernel<1>, dim3(n_col), dim3(256), 0, s, leaves, st, col_slices, part, a, step);
    hipLaunchKernelGGL(aca_pick_row_kernel, dim3(n_leaves), dim3(64), 0, s, leaves, st, part, step);
    hipLaunchKernelGGL(aca_scale_kernel, dim3(n_col), dim3(256), 0, s, leaves, st, col_slices, step);
    if (kind == 0)
      hipLaunchKernelGGL(aca_col_kernel<0>, dim3(n_row), dim3(256), 0, s, leaves, st, row_slices, part, a, step);
    else
      hipLaunchKernelGGL(aca_col_kernel<1>, dim3(n_row), dim3(256), 0, s, leaves, st, row_slices, part, a, step);
    hipLaunchKernelGGL(aca_pick_col_kernel, dim3(n_leaves), dim3(64), 0, s, leaves, st, part, step);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_truncate(TruncLeaf* leaves, int n_leaves, int max_kp, TruncFactor* factors,
                           const TruncSlice* slices, int n_slices, int* flag, hipStream_t s) {
  if (n_leaves <= 0) return hipSuccess;
  hipLaunchKernelGGL(tr_colsq_kernel, dim3(n_slices), dim3(256), 0, s, slices);
  hipLaunchKernelGGL(tr_norm_kernel, dim3(2 * n_leaves), dim3(256), 0, s, factors);
  hipLaunchKernelGGL(tr_scale_kernel, dim3(n_slices), dim3(256), 0, s, slices);
  for (int pass = 0; pass < 3; ++pass) {
    hipLaunchKernelGGL(tr_gram_kernel, dim3(n_slices), dim3(256), 0, s, slices);
    hipLaunchKernelGGL(tr_gsum_kernel, dim3(2 * n_leaves, (max_kp * max_kp + 255) / 256), dim3(256), 0, s, factors);
    hipLaunchKernelGGL(tr_chol_kernel, dim3(2 * n_leaves), dim3(256), 0, s, factors, pass);
    hipLaunchKernelGGL(tr_rprod_kernel, dim3(2 * n_leaves, (max_kp * max_kp + 255) / 256), dim3(256), 0, s,
                       factors);
    hipLaunchKernelGGL(tr_rcopy_kernel, dim3(2 * n_leaves, (max_kp * max_kp + 255) / 256), dim3(256), 0, s,
                       factors);
    hipLaunchKernelGGL(tr_trsm_kernel, dim3(n_slices), dim3(256), 0, s, slices, factors);
  }
  hipLaunchKernelGGL(tr_core_kernel, dim3(n_leaves), dim3(256), 0, s, leaves, factors);
  // sweeps until no leaf rotates (at most 30): the flag is read back per sweep
  for (int sweep = 0; sweep < 30; ++sweep) {
    hipLaunchKernelGGL(tr_clear_rotated_kernel, dim3((n_leaves + 255) / 256), dim3(256), 0, s, leaves, n_leaves);
    for (int rd = 0; rd < max_kp - 1; ++rd)
      hipLaunchKernelGGL(tr_jacobi_kernel, dim3(n_leaves, (max_kp / 2 + 3) / 4), dim3(256), 0, s, leaves, rd);
    hipLaunchKernelGGL(tr_any_rotated_kernel, dim3(1), dim3(64), 0, s, leaves, n_leaves, flag);
    int any = 0;
    hipError_t e = hipMemcpyAsync(&any, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (!any) break;
  }
  hipLaunchKernelGGL(tr_select_kernel, dim3(n_leaves), dim3(256), 0, s, leaves);
  hipLaunchKernelGGL(tr_out_kernel, dim3(n_slices), dim3(256), 0, s, slices, leaves);
  return hipGetLastError();
}

}  // namespace pp
