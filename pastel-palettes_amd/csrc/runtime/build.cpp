// H-matrix construction on the device (C ABI): leaf tables for the dense
// fill and cross-approximation kernels.
#include <algorithm>
#include <vector>

#include "../matrix/Hierarchical.h"
#include "runtime_internal.h"

using namespace ppr;
using ppc::from_exception;
using ppc::set_error;

namespace {

// Rank of the cross approximation of a low-rank leaf: max(4r, 64) (at most
// 256 and the block's smaller side), truncated to r afterwards; r itself when
// that is no larger (the approximation is then exact or as good as it gets).
int aca_rank(int ny, int nx, int r) {
  const int k = std::min(std::min(std::max(4 * r, 64), pp::kBuildMaxRank), std::min(ny, nx));
  return k > r ? k : r;
}

}  // namespace

extern "C" {

int pp_hmatrix_build_kernel(pp_ctx* ctx, const pp_hmatrix* h, int kind, const double* points, double param,
                            double shift, double* arena, int32_t* pivots, void* hip_stream) {
  try {
    if (ctx == nullptr || h == nullptr || points == nullptr || arena == nullptr ||
        (kind != PP_KERNEL_EXP1D && kind != PP_KERNEL_INV2D) || !(param > 0.0))
      return set_error(PP_ERR_INVALID, "bad build args");
    std::vector<pp::BuildLeaf> dense, direct, approx;
    std::vector<pp::RecompressLeaf> trunc;
    int max_dense = 0;
    int64_t piv = 0, scratch = 0;
    for (const h_block* b : h->leaves) {
      pp::BuildLeaf L{};
      L.row0 = b->row0;
      L.col0 = b->col0;
      L.ny = b->ny;
      L.nx = b->nx;
      if (b->type == element_t::dense) {
        L.off = b->off;
        dense.push_back(L);
        max_dense = (int)std::max<int64_t>(max_dense, std::min<int64_t>((int64_t)b->ny * b->nx, 1 << 30));
        continue;
      }
      if (b->type != element_t::low_rank) continue;
      if (b->rank > pp::kBuildMaxRank || b->ny > pp::kBuildMaxRows)
        return set_error(PP_ERR_INVALID, "low-rank leaf beyond the builder's limits (rank 256, 65536 rows)");
      const int k = aca_rank(b->ny, b->nx, b->rank);
      L.rank = k;
      L.aux = piv;
      piv += 2 * (int64_t)k;
      if (k == b->rank) {
        L.ld = k;
        L.off = b->off_u;
        L.off_v = b->off_v;
        direct.push_back(L);
        continue;
      }
      pp::RecompressLeaf R{};
      R.ny = b->ny;
      R.nx = b->nx;
      R.k = k;
      R.kp = (k + 31) / 32 * 32;
      R.r = b->rank;
      R.u = scratch;
      R.v = R.u + (int64_t)R.ny * R.kp;
      R.work = R.v + (int64_t)R.nx * R.kp;
      scratch = (R.work + 6 * (int64_t)R.kp * R.kp + 2 * R.kp + 31) / 32 * 32;
      R.out_u = b->off_u;
      R.out_v = b->off_v;
      L.ld = R.kp;
      L.off = R.u;
      L.off_v = R.v;
      approx.push_back(L);
      trunc.push_back(R);
    }
    // the largest approximations first (they take longest)
    auto big_first = [](const pp::BuildLeaf& a, const pp::BuildLeaf& b) {
      return (int64_t)(a.ny + a.nx) * a.rank > (int64_t)(b.ny + b.nx) * b.rank;
    };
    std::stable_sort(direct.begin(), direct.end(), big_first);
    std::vector<size_t> order(approx.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return big_first(approx[a], approx[b]); });
    std::vector<pp::BuildLeaf> approx_s;
    std::vector<pp::RecompressLeaf> trunc_s;
    for (size_t i : order) {
      approx_s.push_back(approx[i]);
      trunc_s.push_back(trunc[i]);
    }
    pp::BuildLeaf *d_dense = nullptr, *d_direct = nullptr, *d_approx = nullptr;
    pp::RecompressLeaf* d_trunc = nullptr;
    double* d_scratch = nullptr;
    int rc = ctx_alloc_t(ctx, (int64_t)dense.size(), &d_dense);
    if (rc == PP_OK) rc = ctx_alloc_t(ctx, (int64_t)direct.size(), &d_direct);
    if (rc == PP_OK) rc = ctx_alloc_t(ctx, (int64_t)approx_s.size(), &d_approx);
    if (rc == PP_OK) rc = ctx_alloc_t(ctx, (int64_t)trunc_s.size(), &d_trunc);
    if (rc == PP_OK) rc = ctx_alloc_t(ctx, scratch, &d_scratch);
    hipStream_t st = pick(ctx, hip_stream);
    hipError_t e = hipSuccess;
    auto up = [&](void* dst, const void* src, size_t bytes) {
      if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    if (rc == PP_OK) {
      up(d_dense, dense.data(), sizeof(pp::BuildLeaf) * dense.size());
      up(d_direct, direct.data(), sizeof(pp::BuildLeaf) * direct.size());
      up(d_approx, approx_s.data(), sizeof(pp::BuildLeaf) * approx_s.size());
      up(d_trunc, trunc_s.data(), sizeof(pp::RecompressLeaf) * trunc_s.size());
      if (e == hipSuccess)
        e = pp::launch_build(kind, d_dense, (int)dense.size(), max_dense, d_direct, (int)direct.size(), points, param,
                             shift, arena, pivots, st);
      if (e == hipSuccess)
        e = pp::launch_build(kind, nullptr, 0, 0, d_approx, (int)approx_s.size(), points, param, shift, d_scratch,
                             pivots, st);
      if (e == hipSuccess) e = pp::launch_recompress(d_trunc, (int)trunc_s.size(), d_scratch, arena, st);
      if (e == hipSuccess && !trunc_s.empty())
        e = hipMemcpyAsync(trunc_s.data(), d_trunc, sizeof(pp::RecompressLeaf) * trunc_s.size(),
                           hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) rc = hip_fail(e, "H-matrix build");
    }
    for (size_t i = 0; rc == PP_OK && i < trunc_s.size(); ++i)
      if (trunc_s[i].status != 0)
        rc = set_error(PP_ERR_INTERNAL, "truncation of a cross approximation failed (Cholesky pivot)");
    if (d_dense) ctx_free(ctx, d_dense);
    if (d_direct) ctx_free(ctx, d_direct);
    if (d_approx) ctx_free(ctx, d_approx);
    if (d_trunc) ctx_free(ctx, d_trunc);
    if (d_scratch) ctx_free(ctx, d_scratch);
    return rc;
  } catch (...) {
    return from_exception();
  }
}

int64_t pp_hmatrix_build_pivots(const pp_hmatrix* h) {
  if (h == nullptr) return set_error(PP_ERR_INVALID, "null matrix");
  int64_t n = 0;
  for (const h_block* b : h->leaves)
    if (b->type == element_t::low_rank) n += 2 * (int64_t)aca_rank(b->ny, b->nx, b->rank);
  return n;
}

}  // extern "C"
