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

// Rank of the cross approximation of a low-rank leaf of rank r:
// max(4r, 64, min(rows, cols) / 16), at most 256 and the block's smaller side;
// truncated to r afterwards.  r itself when that is no larger (the
// approximation is then exact or as good as it gets).
int aca_rank(int ny, int nx, int r) {
  const int small = std::min(ny, nx);
  const int k = std::min(std::min(std::max(std::max(4 * r, 64), small / 16), pp::kBuildMaxRank), small);
  return k > r ? k : r;
}

// Device scratch carved from one allocation (8-byte elements, 256-byte
// aligned pieces).
struct carver {
  int64_t used = 0;
  int64_t take(int64_t elems) {
    const int64_t at = used;
    used += (elems + 31) / 32 * 32;
    return at;
  }
};

}  // namespace

extern "C" {

int pp_hmatrix_build_kernel(pp_ctx* ctx, const pp_hmatrix* h, int kind, const double* points, double param,
                            double shift, double* arena, int32_t* pivots, void* hip_stream) {
  try {
    if (ctx == nullptr || h == nullptr || points == nullptr || arena == nullptr ||
        (kind != PP_KERNEL_EXP1D && kind != PP_KERNEL_INV2D) || !(param > 0.0))
      return set_error(PP_ERR_INVALID, "bad build args");
    const int S = pp::kBuildSlice;
    std::vector<pp::BuildLeaf> dense;
    struct lr {
      const h_block* b;
      int k, kp;
      int64_t piv;
    };
    std::vector<lr> lrs;
    int max_dense = 0, max_k = 0;
    int64_t piv = 0;
    for (const h_block* b : h->leaves) {
      if (b->type == element_t::dense) {
        pp::BuildLeaf L{};
        L.row0 = b->row0;
        L.col0 = b->col0;
        L.ny = b->ny;
        L.nx = b->nx;
        L.off = b->off;
        dense.push_back(L);
        max_dense = (int)std::max<int64_t>(max_dense, std::min<int64_t>((int64_t)b->ny * b->nx, 1 << 30));
      } else if (b->type == element_t::low_rank) {
        if (b->rank > pp::kBuildMaxRank || b->ny > pp::kBuildMaxRows)
          return set_error(PP_ERR_INVALID, "low-rank leaf beyond the builder's limits (rank 256, 65536 rows)");
        const int k = aca_rank(b->ny, b->nx, b->rank);
        lrs.push_back(lr{b, k, k > b->rank ? (k + 31) / 32 * 32 : k, piv});
        piv += 2 * (int64_t)k;
        max_k = std::max(max_k, k);
      }
    }
    // scratch plan (element offsets; pointers are fixed up after allocation)
    carver cv;
    std::vector<int64_t> o_u(lrs.size(), -1), o_v(lrs.size(), -1), o_used(lrs.size());
    const int64_t o_state = cv.take(((int64_t)lrs.size() * sizeof(pp::AcaState) + 7) / 8);
    int n_col = 0, n_row = 0;
    for (size_t i = 0; i < lrs.size(); ++i) {
      n_col += (lrs[i].b->nx + S - 1) / S;
      n_row += (lrs[i].b->ny + S - 1) / S;
      o_used[i] = cv.take((lrs[i].b->ny + 63) / 64);
      if (lrs[i].k > lrs[i].b->rank) {
        o_u[i] = cv.take((int64_t)lrs[i].b->ny * lrs[i].kp);
        o_v[i] = cv.take((int64_t)lrs[i].b->nx * lrs[i].kp);
      }
    }
    const int64_t o_part = cv.take(((int64_t)(n_col + n_row) * sizeof(pp::BuildPartial) + 7) / 8);
    const int64_t o_flag = cv.take(1);
    int max_kp = 0;
    // truncation: per factor G, Rt, T, d and the slice partials; per leaf N, J, sel
    std::vector<size_t> trunc;
    for (size_t i = 0; i < lrs.size(); ++i)
      if (o_u[i] >= 0) trunc.push_back(i);
    struct fplan {
      int64_t G, Rt, T, d, Ri, parts, gparts;
      int nsl;
    };
    std::vector<fplan> fp(2 * trunc.size());
    std::vector<int64_t> o_N(trunc.size()), o_J(trunc.size()), o_sel(trunc.size());
    for (size_t q = 0; q < trunc.size(); ++q) {
      const lr& L = lrs[trunc[q]];
      const int64_t kp = L.kp;
      max_kp = std::max(max_kp, L.kp);
      for (int f = 0; f < 2; ++f) {
        const int m = f == 0 ? L.b->ny : L.b->nx;
        fplan& F = fp[2 * q + f];
        F.nsl = (m + S - 1) / S;
        F.G = cv.take(kp * kp);
        F.Rt = cv.take(kp * kp);
        F.T = cv.take(kp * kp);
        F.d = cv.take(kp);
        F.Ri = cv.take(kp * kp);
        F.parts = cv.take(F.nsl * kp);
        F.gparts = cv.take(F.nsl * kp * kp);
      }
      o_N[q] = cv.take(kp * kp);
      o_J[q] = cv.take(kp * kp);
      o_sel[q] = cv.take((L.b->rank + 1) / 2);
    }
    // device allocations of this call, freed on every way out
    struct freer {
      pp_ctx* c;
      std::vector<void*> p;
      ~freer() {
        for (void* x : p) ctx_free(c, x);
      }
    } owned{ctx, {}};
    double* d_s = nullptr;
    int rc = ctx_alloc_t(ctx, std::max<int64_t>(cv.used, 1), &d_s);
    if (rc) return rc;
    owned.p.push_back(d_s);
    // tables
    std::vector<pp::AcaLeaf> al(lrs.size());
    std::vector<pp::AcaState> as(lrs.size());
    std::vector<pp::BuildSlice> cs, rs;
    for (size_t i = 0; i < lrs.size(); ++i) {
      const lr& L = lrs[i];
      pp::AcaLeaf& A = al[i];
      A.row0 = L.b->row0;
      A.col0 = L.b->col0;
      A.ny = L.b->ny;
      A.nx = L.b->nx;
      A.k = L.k;
      A.ld = L.kp;
      A.U = o_u[i] >= 0 ? d_s + o_u[i] : arena + L.b->off_u;
      A.V = o_v[i] >= 0 ? d_s + o_v[i] : arena + L.b->off_v;
      A.piv = pivots ? pivots + L.piv : nullptr;
      A.used = reinterpret_cast<uint32_t*>(d_s + o_used[i]);
      A.col_part = (int)cs.size();
      for (int j = 0; j < A.nx; j += S) cs.push_back(pp::BuildSlice{(int)i, j, std::min(S, A.nx - j), 0});
      A.ncol = (int)cs.size() - A.col_part;
      as[i] = pp::AcaState{0, 0, 0, 0, 0.0};
    }
    for (size_t i = 0; i < lrs.size(); ++i) {
      pp::AcaLeaf& A = al[i];
      A.row_part = n_col + (int)rs.size();
      for (int t = 0; t < A.ny; t += S) rs.push_back(pp::BuildSlice{(int)i, t, std::min(S, A.ny - t), 0});
      A.nrow = n_col + (int)rs.size() - A.row_part;
    }
    for (size_t q = 0; q < cs.size(); ++q) cs[q].part = (int)q;
    for (size_t q = 0; q < rs.size(); ++q) rs[q].part = n_col + (int)q;
    std::vector<pp::TruncFactor> tf(2 * trunc.size());
    std::vector<pp::TruncLeaf> tl(trunc.size());
    std::vector<pp::TruncSlice> ts;
    for (size_t q = 0; q < trunc.size(); ++q) {
      const size_t i = trunc[q];
      const lr& L = lrs[i];
      for (int f = 0; f < 2; ++f) {
        const fplan& P = fp[2 * q + f];
        pp::TruncFactor& F = tf[2 * q + f];
        F.G = d_s + P.G;
        F.Rt = d_s + P.Rt;
        F.T = d_s + P.T;
        F.d = d_s + P.d;
        F.Ri = d_s + P.Ri;
        F.parts = d_s + P.parts;
        F.gparts = d_s + P.gparts;
        F.part_stride = L.kp;
        F.gpart_stride = (int64_t)L.kp * L.kp;
        F.kp = L.kp;
        F.nslices = P.nsl;
        F.m = f == 0 ? L.b->ny : L.b->nx;
        F.status = 0;
        double* X = f == 0 ? d_s + o_u[i] : d_s + o_v[i];
        for (int sq = 0; sq < P.nsl; ++sq) {
          pp::TruncSlice T{};
          T.X = X;
          T.part = F.parts + (int64_t)sq * F.part_stride;
          T.gpart = F.gparts + (int64_t)sq * F.gpart_stride;
          T.d = F.d;
          T.first = sq * S;
          T.count = std::min(S, F.m - sq * S);
          T.kp = L.kp;
          T.factor = (int)(2 * q + f);
          ts.push_back(T);
        }
      }
      pp::TruncLeaf& T = tl[q];
      T.N = d_s + o_N[q];
      T.J = d_s + o_J[q];
      T.out_u = arena + L.b->off_u;
      T.out_v = arena + L.b->off_v;
      T.sel = reinterpret_cast<int32_t*>(d_s + o_sel[q]);
      T.kp = L.kp;
      T.r = L.b->rank;
      T.status = 0;
      T.rotated = 0;
    }
    // device copies of the tables
    pp::BuildLeaf* d_dense = nullptr;
    pp::AcaLeaf* d_al = nullptr;
    pp::BuildSlice *d_cs = nullptr, *d_rs = nullptr;
    pp::TruncFactor* d_tf = nullptr;
    pp::TruncLeaf* d_tl = nullptr;
    pp::TruncSlice* d_ts = nullptr;
    // one device array per table
    auto table = [&](auto** dp, const auto& v) {
      if (rc == PP_OK) rc = ctx_alloc_t(ctx, (int64_t)v.size(), dp);
      if (rc == PP_OK) owned.p.push_back(*dp);
    };
    table(&d_dense, dense);
    table(&d_al, al);
    table(&d_cs, cs);
    table(&d_rs, rs);
    table(&d_tf, tf);
    table(&d_tl, tl);
    table(&d_ts, ts);
    hipStream_t st = pick(ctx, hip_stream);
    hipError_t e = hipSuccess;
    auto up = [&](void* dst, const void* src, size_t bytes) {
      if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    if (rc == PP_OK) {
      up(d_dense, dense.data(), sizeof(pp::BuildLeaf) * dense.size());
      up(d_al, al.data(), sizeof(pp::AcaLeaf) * al.size());
      up(d_cs, cs.data(), sizeof(pp::BuildSlice) * cs.size());
      up(d_rs, rs.data(), sizeof(pp::BuildSlice) * rs.size());
      up(d_tf, tf.data(), sizeof(pp::TruncFactor) * tf.size());
      up(d_tl, tl.data(), sizeof(pp::TruncLeaf) * tl.size());
      up(d_ts, ts.data(), sizeof(pp::TruncSlice) * ts.size());
      up(d_s + o_state, as.data(), sizeof(pp::AcaState) * as.size());
      for (size_t i = 0; e == hipSuccess && i < lrs.size(); ++i)
        e = hipMemsetAsync(d_s + o_used[i], 0, sizeof(uint32_t) * ((lrs[i].b->ny + 31) / 32), st);
      const pp::BuildArgs args{points, param, shift};
      if (e == hipSuccess) e = pp::launch_dense_fill(kind, d_dense, (int)dense.size(), max_dense, args, arena, st);
      if (e == hipSuccess)
        e = pp::launch_aca(kind, d_al, reinterpret_cast<pp::AcaState*>(d_s + o_state), (int)lrs.size(), max_k, d_cs,
                           (int)cs.size(), d_rs, (int)rs.size(), reinterpret_cast<pp::BuildPartial*>(d_s + o_part),
                           args, st);
      if (e == hipSuccess)
        e = pp::launch_truncate(d_tl, (int)tl.size(), max_kp, d_tf, d_ts, (int)ts.size(),
                                reinterpret_cast<int*>(d_s + o_flag), st);
      if (e == hipSuccess && !tl.empty())
        e = hipMemcpyAsync(tl.data(), d_tl, sizeof(pp::TruncLeaf) * tl.size(), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) rc = hip_fail(e, "H-matrix build");
    }
    for (size_t q = 0; rc == PP_OK && q < tl.size(); ++q)
      if (tl[q].status != 0)
        rc = set_error(PP_ERR_INTERNAL, "truncation of a cross approximation failed (Cholesky pivot)");
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
