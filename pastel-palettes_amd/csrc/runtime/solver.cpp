// Factor + solve + FP64 iterative refinement over one or many H-matrices.
//
// A batch of H-matrices is handled as one block-diagonal system: their FP64
// arenas are concatenated (member i at arena_off[i]), their rows are stacked
// (member i at row_off[i]), one program factors all of them in a single
// persistent launch and one program solves all of them.  Refinement runs on
// the stacked system with per-member, per-column convergence checks.
//
// Precision: the factor arena holds the storage type of the precision mode
// (FP64, or FP32 for FP32 / mixed); the FP64 arena stays resident for the
// residual r = b - A_H x, which is always computed in FP64.
//
// Each member's arena is followed by its leaf-inverse area.  The factor
// program also forms L^{-1} and U^{-1} of every dense diagonal leaf there
// (those ops depend only on the getrf of their own leaf and fill idle
// workers), and the solve applies each leaf with one product instead of a
// substitution.  The FP64 copy holds identity blocks in the inverse areas, so
// the arena restore before every factorization re-seeds them.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../io/wire.h"
#include "../matrix/Hierarchical.h"
#include "runtime_internal.h"

using namespace ppr;
using ppc::from_exception;
using ppc::set_error;

struct pp_solver {
  pp_ctx* ctx = nullptr;
  std::vector<const pp_hmatrix*> hms;
  std::vector<int64_t> arena_off;  // elements
  std::vector<int64_t> row_off;
  int64_t n_total = 0, storage_total = 0;
  int precision = 0, workers = 0, wpi = 4, nrhs_max = 0, rhs_split = 1;
  double* d_A = nullptr;  // FP64 arenas (residual)
  void* d_F = nullptr;    // factor arenas (storage precision)
  pp_program* fprog = nullptr;
  pp_stream* fstream = nullptr;
  struct SolveProg {
    pp_program* prog = nullptr;
    pp_stream* stream = nullptr;
  };
  std::map<int, SolveProg> sprogs;
  void* d_rhs = nullptr;    // n_total x nrhs_max, storage precision
  double* d_t = nullptr;    // matvec low-rank products
  double* d_nrm = nullptr;  // [members * nrhs_max] r norms, then the same for b
  double* d_npart = nullptr;  // [members * norm_chunks * nrhs_max] first-stage norm sums
  int norm_chunks = 1;        // kNormRows chunks of the largest member
  double* h_nrm = nullptr;    // pinned host copy of d_nrm
  int32_t* h_status = nullptr;  // pinned host copy of the factor launch's status words
  int* d_seg = nullptr;     // member row ranges (row0, rows)
  int* d_seg_tiles = nullptr;  // member row-tile ranges of the residual plan (first, count)
  double* d_rpart = nullptr;   // residual workgroups' column sums of squares
  hipEvent_t ev_done = nullptr;  // end of the work a host wait covers
  pp::MatvecPlanDev plan{};  // A_H (FP64 arena): residual
  // Hierarchical-inverse solve (every member HODLR-shaped): x = U^{-1} L^{-1} b
  // as two H-matrix products over the inverse areas of the factor arena.
  bool hinv = false;
  int solve_mode = PP_SOLVE_HINV;
  pp::MatvecPlanDev lplan{}, uplan{};
  void* d_rhs2 = nullptr;  // n_total x nrhs_max, storage precision
  void* d_tinv = nullptr;  // V^T x products of lplan / uplan, storage precision
  double factor_flops = 0;
  double last_factor_ms = 0;
  // the factor arena holds factors of the current matrices: set by a
  // factorization whose launch reported no VM error, or by loading factors;
  // cleared when a factorization starts
  bool factored = false;
  std::vector<double> member_res;
  bool use_graph = false;
  hipGraphExec_t factor_graph = nullptr;
  hipStream_t graph_stream = nullptr;
  std::vector<void*> owned;
};

namespace {

const Hierarchical* hier(const pp_hmatrix* h) { return h->h.get(); }

// Uploads a block list as a matvec plan: one row tile per dense diagonal
// block (they must tile the stacked rows), each with the blocks covering it.
int upload_plan(pp_solver* s, std::vector<pp::MatvecBlock> blocks, pp::MatvecPlanDev* plan,
                std::vector<pp::MatvecTile>* tiles_out = nullptr) {
  std::vector<int> lr;
  std::vector<std::pair<int64_t, int>> ranges;  // global row0, rows
  for (size_t i = 0; i < blocks.size(); ++i) {
    const auto& b = blocks[i];
    if (b.type == (int)element_t::low_rank) lr.push_back((int)i);
    if (b.type == (int)element_t::dense && b.row0 == b.col0) ranges.emplace_back(b.row0, b.ny);
  }
  if (s->n_total > 0x7fffffffLL) return set_error(PP_ERR_INVALID, "stacked rows exceed int32");
  std::sort(ranges.begin(), ranges.end());
  // V^T x partial products: parts of at most kPartRows rows of V, each
  // block's partials contiguous (room for nrhs_max columns)
  std::vector<int> parts;
  int64_t p_elems = 0;
  int max_rank = 0;
  for (int i : lr) {
    auto& b = blocks[i];
    b.part0 = (int32_t)(parts.size() / 2);
    b.nparts = b.rank > 0 && b.nx > 0 ? (b.nx + pp::kPartRows - 1) / pp::kPartRows : 0;
    b.p_off = p_elems;
    b.p_stride = (int64_t)b.rank * s->nrhs_max;
    p_elems += b.nparts * b.p_stride;
    for (int q = 0; q < b.nparts; ++q) {
      parts.push_back(i);
      parts.push_back(q);
    }
    max_rank = std::max(max_rank, b.nparts > 0 ? b.rank : 0);
  }
  // Row tiles of at most kTileRows rows, each product workgroup taking 16 of
  // them: an N = 16384 product at leaf 256 has 1024 workgroups.
  constexpr int kTileRows = pp::kTileRows;
  std::vector<pp::MatvecTile> tiles;
  std::vector<int> contrib;
  std::vector<int64_t> rows_of(blocks.size(), 0);  // rows of the tiles each block contributes to
  for (const auto& r : ranges) {
    for (int64_t r0 = r.first; r0 < r.first + r.second; r0 += kTileRows) {
      pp::MatvecTile t{};
      t.row0 = (int32_t)r0;
      t.rows = (int)std::min<int64_t>(kTileRows, r.first + r.second - r0);
      t.first = (int)(contrib.size() / 2);
      for (size_t i = 0; i < blocks.size(); ++i) {
        const auto& b = blocks[i];
        if (b.row0 <= r0 && r0 + t.rows <= (int64_t)b.row0 + b.ny) {
          contrib.push_back((int)i);
          contrib.push_back(0);
          rows_of[i] += t.rows;
        }
      }
      t.count = (int)(contrib.size() / 2) - t.first;
      tiles.push_back(t);
    }
  }
  int64_t covered = 0;
  for (const auto& t : tiles) covered += t.rows;
  if (covered != s->n_total) return set_error(PP_ERR_INVALID, "matvec plan: diagonal blocks do not tile the rows");
  // a block whose rows start or end inside a row tile would be left out of
  // that tile: every block must be covered by whole tiles
  for (size_t i = 0; i < blocks.size(); ++i)
    if (rows_of[i] != blocks[i].ny)
      return set_error(PP_ERR_INVALID, "matvec plan: block " + std::to_string(i) + " is not aligned to the row tiles");
  pp::MatvecBlock* d_blocks = nullptr;
  pp::MatvecTile* d_tiles = nullptr;
  int *d_contrib = nullptr, *d_lr = nullptr, *d_parts = nullptr;
  int rc;
  if ((rc = ctx_alloc_t(s->ctx, blocks.size(), &d_blocks))) return rc;
  s->owned.push_back(d_blocks);
  if ((rc = ctx_alloc_t(s->ctx, tiles.size(), &d_tiles))) return rc;
  s->owned.push_back(d_tiles);
  if ((rc = ctx_alloc_t(s->ctx, std::max<size_t>(contrib.size(), 2), &d_contrib))) return rc;
  s->owned.push_back(d_contrib);
  if ((rc = ctx_alloc_t(s->ctx, std::max<size_t>(lr.size(), 1), &d_lr))) return rc;
  s->owned.push_back(d_lr);
  if ((rc = ctx_alloc_t(s->ctx, std::max<size_t>(parts.size(), 2), &d_parts))) return rc;
  s->owned.push_back(d_parts);
  HIP_TRY(hipMemcpy(d_blocks, blocks.data(), sizeof(pp::MatvecBlock) * blocks.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d_tiles, tiles.data(), sizeof(pp::MatvecTile) * tiles.size(), hipMemcpyHostToDevice));
  if (!contrib.empty())
    HIP_TRY(hipMemcpy(d_contrib, contrib.data(), sizeof(int) * contrib.size(), hipMemcpyHostToDevice));
  if (!lr.empty()) HIP_TRY(hipMemcpy(d_lr, lr.data(), sizeof(int) * lr.size(), hipMemcpyHostToDevice));
  if (!parts.empty()) HIP_TRY(hipMemcpy(d_parts, parts.data(), sizeof(int) * parts.size(), hipMemcpyHostToDevice));
  plan->blocks = d_blocks;
  plan->tiles = d_tiles;
  plan->contrib = d_contrib;
  plan->lr_list = d_lr;
  plan->n_tiles = (int)tiles.size();
  plan->n_lowrank = (int)lr.size();
  plan->parts = d_parts;
  plan->n_parts = (int)(parts.size() / 2);
  plan->max_rank = max_rank;
  plan->p_elems = p_elems;
  if (tiles_out) *tiles_out = std::move(tiles);
  return PP_OK;
}

// Residual plan over the FP64 arenas: every leaf of every member.
int build_plan(pp_solver* s) {
  std::vector<pp::MatvecBlock> blocks;
  int64_t t_off = 0;
  for (size_t m = 0; m < s->hms.size(); ++m) {
    std::vector<const h_block*> lv;
    hier(s->hms[m])->leaves(lv);
    const int64_t ro = s->row_off[m], ao = s->arena_off[m];
    for (const h_block* b : lv) {
      pp::MatvecBlock mb{};
      mb.type = (int)b->type;
      mb.row0 = (int32_t)(b->row0 + ro);
      mb.col0 = (int32_t)(b->col0 + ro);
      mb.ny = b->ny;
      mb.nx = b->nx;
      mb.rank = b->rank;
      mb.off = b->off + ao;
      mb.off_u = b->off_u + ao;
      mb.off_v = b->off_v + ao;
      if (b->type == element_t::low_rank) {
        mb.t_off = t_off;
        t_off += (int64_t)b->rank * s->nrhs_max;
      }
      blocks.push_back(mb);
    }
  }
  int rc;
  std::vector<pp::MatvecTile> tiles;
  if ((rc = upload_plan(s, blocks, &s->plan, &tiles))) return rc;
  s->plan.t_elems = t_off;
  if ((rc = ctx_alloc_t(s->ctx, std::max<int64_t>(t_off + s->plan.p_elems, 1), &s->d_t))) return rc;
  s->owned.push_back(s->d_t);
  std::vector<int> seg;
  int max_rows = 1;
  for (size_t m = 0; m < s->hms.size(); ++m) {
    seg.push_back((int)s->row_off[m]);
    seg.push_back(hier(s->hms[m])->getNx_abs());
    max_rows = std::max(max_rows, seg.back());
  }
  s->norm_chunks = (max_rows + pp::kNormRows - 1) / pp::kNormRows;
  // each member's row tiles are consecutive (tiles follow the stacked rows)
  std::vector<int> seg_tiles;
  for (size_t m = 0; m < s->hms.size(); ++m) {
    const int64_t r0 = s->row_off[m], r1 = r0 + hier(s->hms[m])->getNx_abs();
    int first = -1, count = 0;
    for (size_t i = 0; i < tiles.size(); ++i)
      if (tiles[i].row0 >= r0 && tiles[i].row0 < r1) {
        if (first < 0) first = (int)i;
        if ((int)i != first + count) return set_error(PP_ERR_INTERNAL, "matvec plan: member tiles not consecutive");
        ++count;
      }
    seg_tiles.push_back(std::max(first, 0));
    seg_tiles.push_back(count);
  }
  if ((rc = ctx_alloc_t(s->ctx, seg_tiles.size(), &s->d_seg_tiles))) return rc;
  s->owned.push_back(s->d_seg_tiles);
  HIP_TRY(hipMemcpy(s->d_seg_tiles, seg_tiles.data(), sizeof(int) * seg_tiles.size(), hipMemcpyHostToDevice));
  if ((rc = ctx_alloc_t(s->ctx, (int64_t)tiles.size() * (pp::kTileRows / 16) * s->nrhs_max, &s->d_rpart))) return rc;
  s->owned.push_back(s->d_rpart);
  if ((rc = ctx_alloc_t(s->ctx, (int64_t)s->hms.size() * s->norm_chunks * s->nrhs_max, &s->d_npart))) return rc;
  s->owned.push_back(s->d_npart);
  if ((rc = ctx_alloc_t(s->ctx, seg.size(), &s->d_seg))) return rc;
  s->owned.push_back(s->d_seg);
  HIP_TRY(hipMemcpy(s->d_seg, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice));
  return PP_OK;
}

// Plans applying L^{-1} and U^{-1} from the inverse areas of the factor
// arena, when every member is HODLR-shaped; otherwise the solve runs the VM
// solve program.
int build_inverse_plans(pp_solver* s) {
  for (const pp_hmatrix* h : s->hms)
    if (!hier(h)->hodlrShaped()) return PP_OK;
  std::vector<pp::MatvecBlock> lb, ub;
  int64_t tl = 0, tu = 0;
  for (size_t m = 0; m < s->hms.size(); ++m) {
    const Hierarchical* h = hier(s->hms[m]);
    const int64_t base = s->arena_off[m] + inverse_base(h), ro = s->row_off[m];
    std::vector<Hierarchical::inv_block> leaves;
    std::vector<Hierarchical::inv_node> nodes;
    h->inverseLayout(leaves);
    h->inverseNodes(nodes);
    for (const auto& e : leaves) {
      pp::MatvecBlock d{};
      d.type = (int)element_t::dense;
      d.row0 = d.col0 = (int32_t)(ro + e.row0);
      d.ny = d.nx = e.nb;
      d.off = base + e.inv_off;
      lb.push_back(d);
      d.off = base + e.inv_off + ((int64_t)e.nb * e.nb + 31) / 32 * 32;
      ub.push_back(d);
    }
    for (const auto& n : nodes) {
      pp::MatvecBlock l{};
      l.type = (int)element_t::low_rank;
      l.row0 = (int32_t)(ro + n.row0 + n.na);
      l.col0 = (int32_t)(ro + n.row0);
      l.ny = n.nb;
      l.nx = n.na;
      l.rank = n.rl;
      l.off_u = base + n.l_u;
      l.off_v = base + n.l_v;
      l.t_off = tl;
      tl += (int64_t)n.rl * s->nrhs_max;
      lb.push_back(l);
      pp::MatvecBlock u{};
      u.type = (int)element_t::low_rank;
      u.row0 = (int32_t)(ro + n.row0);
      u.col0 = (int32_t)(ro + n.row0 + n.na);
      u.ny = n.na;
      u.nx = n.nb;
      u.rank = n.ru;
      u.off_u = base + n.u_u;
      u.off_v = base + n.u_v;
      u.t_off = tu;
      tu += (int64_t)n.ru * s->nrhs_max;
      ub.push_back(u);
    }
  }
  int rc;
  if ((rc = upload_plan(s, lb, &s->lplan))) return rc;
  if ((rc = upload_plan(s, ub, &s->uplan))) return rc;
  s->lplan.t_elems = tl;
  s->uplan.t_elems = tu;
  const int eb = elem_bytes(s->precision);
  if ((rc = ctx_alloc(s->ctx,
                      (uint64_t)std::max<int64_t>(std::max(tl + s->lplan.p_elems, tu + s->uplan.p_elems), 1) * eb,
                      &s->d_tinv)))
    return rc;
  s->owned.push_back(s->d_tinv);
  if ((rc = ctx_alloc(s->ctx, (uint64_t)s->n_total * s->nrhs_max * eb, &s->d_rhs2))) return rc;
  s->owned.push_back(s->d_rhs2);
  s->hinv = true;
  return PP_OK;
}

void destroy_solver(pp_solver* s) {
  if (s == nullptr) return;
  if (s->factor_graph) (void)hipGraphExecDestroy(s->factor_graph);
  if (s->fstream) pp_stream_destroy(s->ctx, s->fstream);
  if (s->fprog) pp_program_destroy(s->fprog);
  for (auto& kv : s->sprogs) {
    if (kv.second.stream) pp_stream_destroy(s->ctx, kv.second.stream);
    if (kv.second.prog) pp_program_destroy(kv.second.prog);
  }
  for (void* p : s->owned) (void)ctx_free(s->ctx, p);
  if (s->h_nrm) (void)hipHostFree(s->h_nrm);
  if (s->h_status) (void)hipHostFree(s->h_status);
  if (s->ev_done) (void)hipEventDestroy(s->ev_done);
  delete s;
}

// pointer table of a program over the members: slot 3i = factor arena of
// member i, 3i+1 = temp (owned by the stream), 3i+2 = its rhs rows.
std::vector<void*> slot_table(pp_solver* s, int nrhs) {
  const int eb = elem_bytes(s->precision);
  std::vector<void*> ptrs;
  for (size_t m = 0; m < s->hms.size(); ++m) {
    ptrs.push_back(static_cast<char*>(s->d_F) + s->arena_off[m] * eb);
    ptrs.push_back(nullptr);
    ptrs.push_back(static_cast<char*>(s->d_rhs) + s->row_off[m] * nrhs * eb);
  }
  return ptrs;
}

int solver_create(pp_ctx* ctx, const pp_hmatrix* const* hs, int count, const double* const* storages, bool on_device,
                  int precision, int workers, int wpi, int nrhs_max, int rhs_split, int mode, pp_solver** out) {
  if (ctx == nullptr || hs == nullptr || count < 1 || storages == nullptr || out == nullptr || precision < 0 ||
      precision > 2 || workers < 1 || nrhs_max < 1 || rhs_split < 1 || (wpi != 4 && wpi != 8) ||
      mode < PP_SOLVE_HINV || mode > PP_SOLVE_SUBST)
    return set_error(PP_ERR_INVALID, "bad solver args");
  auto s = new pp_solver();
  s->solve_mode = mode;
  s->ctx = ctx;
  s->precision = precision;
  s->workers = workers;
  s->wpi = wpi;
  s->nrhs_max = nrhs_max;
  s->rhs_split = rhs_split;
  for (int i = 0; i < count; ++i) {
    if (hs[i] == nullptr || storages[i] == nullptr) {
      destroy_solver(s);
      return set_error(PP_ERR_INVALID, "null member");
    }
    s->hms.push_back(hs[i]);
    s->arena_off.push_back(s->storage_total);
    s->row_off.push_back(s->n_total);
    const Hierarchical* h = hier(hs[i]);
    s->storage_total += mode == PP_SOLVE_SUBST ? (h->getStorage() + 63) / 64 * 64
                                               : (inverse_base(h) + h->inverseStorage() + 63) / 64 * 64;
    s->n_total += hier(hs[i])->getNx_abs();
  }
  int rc = PP_OK;
  const int eb = elem_bytes(precision);
  do {
    if ((rc = ctx_alloc_t(ctx, s->storage_total, &s->d_A))) break;
    s->owned.push_back(s->d_A);
    hipError_t ez = hipMemset(s->d_A, 0, 8 * s->storage_total);
    if (ez != hipSuccess) {
      rc = hip_fail(ez, "clear matrix arena");
      break;
    }
    for (int i = 0; i < count && rc == PP_OK; ++i) {
      const Hierarchical* h = hier(hs[i]);
      hipError_t e = hipMemcpy(s->d_A + s->arena_off[i], storages[i], 8 * h->getStorage(),
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        rc = hip_fail(e, "upload matrix");
        break;
      }
      if (mode == PP_SOLVE_SUBST) continue;
      std::vector<double> ident(h->inverseStorage());
      if (pp_hmatrix_inverse_identity(hs[i], ident.data()) != PP_OK) {
        rc = PP_ERR_INTERNAL;
        break;
      }
      e = hipMemcpy(s->d_A + s->arena_off[i] + inverse_base(h), ident.data(), 8 * ident.size(),
                    hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = hip_fail(e, "upload inverse identities");
    }
    if (rc) break;
    if ((rc = ctx_alloc(ctx, (uint64_t)s->storage_total * eb, &s->d_F))) break;
    s->owned.push_back(s->d_F);
    if ((rc = ctx_alloc(ctx, (uint64_t)s->n_total * nrhs_max * eb, &s->d_rhs))) break;
    s->owned.push_back(s->d_rhs);
    if ((rc = ctx_alloc_t(ctx, 2 * (int64_t)count * nrhs_max, &s->d_nrm))) break;
    s->owned.push_back(s->d_nrm);
    hipError_t he = hipHostMalloc((void**)&s->h_nrm, sizeof(double) * 2 * (size_t)count * nrhs_max, 0);
    if (he == hipSuccess) he = hipHostMalloc((void**)&s->h_status, 8 * sizeof(int32_t), 0);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&s->ev_done, hipEventDisableTiming);
    if (he != hipSuccess) {
      rc = hip_fail(he, "hipHostMalloc");
      break;
    }
    if ((rc = build_plan(s))) break;
    if (mode == PP_SOLVE_HINV && (rc = build_inverse_plans(s))) break;
    if (mode == PP_SOLVE_SUBST) rc = pp_program_factor_batch(hs, count, workers, &s->fprog);
    else rc = ppc::factor_inverse_program(hs, count, workers, mode == PP_SOLVE_HINV, &s->fprog);
    if (rc) break;
    std::vector<void*> ptrs = slot_table(s, nrhs_max);
    if ((rc = pp_stream_upload_program(ctx, s->fprog, const_cast<const void* const*>(ptrs.data()), (int)ptrs.size(),
                                       precision, &s->fstream)))
      break;
    // H-LU FLOPs only: the inversions are part of the solve's set-up
    s->factor_flops = s->fprog->primary_flops >= 0 ? (double)s->fprog->primary_flops : (double)s->fprog->tree.getFlops();
    s->fstream->role = 1;
    s->member_res.assign(count, 0.);
  } while (false);
  if (rc != PP_OK) {
    destroy_solver(s);
    return rc;
  }
  *out = s;
  return PP_OK;
}

// Solve programs are built for a few column widths only (powers of two up to
// nrhs_max): a request with fewer columns runs the next wider program over
// zero-padded right-hand sides (columns never mix), so a stream of requests
// with varying column counts reuses a handful of programs.
int program_width(const pp_solver* s, int nrhs) {
  int w = 1;
  while (w < nrhs) w *= 2;
  return std::min(w, s->nrhs_max);
}

int solve_prog(pp_solver* s, int nrhs, pp_solver::SolveProg** out) {
  auto it = s->sprogs.find(nrhs);
  if (it == s->sprogs.end()) {
    pp_solver::SolveProg sp;
    std::vector<int32_t> nr(s->hms.size(), nrhs), ld(s->hms.size(), nrhs);
    int rc = s->solve_mode == PP_SOLVE_SUBST
                 ? pp_program_solve_batch(s->hms.data(), (int)s->hms.size(), nr.data(), ld.data(), s->rhs_split,
                                          s->workers, &sp.prog)
                 : pp_program_solve_inv_batch(s->hms.data(), (int)s->hms.size(), nr.data(), ld.data(),
                                              s->rhs_split, s->workers, &sp.prog);
    if (rc) return rc;
    std::vector<void*> ptrs = slot_table(s, nrhs);
    rc = pp_stream_upload_program(s->ctx, sp.prog, const_cast<const void* const*>(ptrs.data()), (int)ptrs.size(),
                                  s->precision, &sp.stream);
    if (rc) {
      pp_program_destroy(sp.prog);
      return rc;
    }
    it = s->sprogs.emplace(nrhs, sp).first;
  }
  *out = &it->second;
  return PP_OK;
}

// Check the device status words of a stream after its work completed.
int check_status(pp_stream* st) {
  int32_t v[8];
  HIP_TRY(hipMemcpy(v, st->d_status, sizeof(v), hipMemcpyDeviceToHost));
  return stream_status(st, v, nullptr);
}

// Waits for the work enqueued on st so far by polling an event (the refinement
// loop waits once per sweep for a few hundred bytes of norms; a blocking
// synchronisation would add its wake-up latency to every sweep).
int wait_host(pp_solver* s, hipStream_t st) {
  HIP_TRY(hipEventRecord(s->ev_done, st));
  for (;;) {
    const hipError_t q = hipEventQuery(s->ev_done);
    if (q == hipSuccess) return PP_OK;
    if (q != hipErrorNotReady) return hip_fail(q, "hipEventQuery");
    std::this_thread::yield();
  }
}

// Waits for the factorization launch on st, with its status words copied
// into pinned host memory by the same stream (one wait).
int finish_factor(pp_solver* s, hipStream_t st) {
  HIP_TRY(hipMemcpyAsync(s->h_status, s->fstream->d_status, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (const int rc = wait_host(s, st)) return rc;
  s->last_factor_ms = pp_stream_last_ms(s->fstream);
  const int rc = stream_status(s->fstream, s->h_status, nullptr);
  s->factored = rc == PP_OK;
  return rc;
}

int enqueue_factor(pp_solver* s, hipStream_t st, bool events) {
  hipError_t e = pp::launch_convert(s->d_F, elem_bytes(s->precision), s->d_A, 8, s->storage_total, st);
  if (e != hipSuccess) return hip_fail(e, "restore factor arena");
  return enqueue_launch(s->ctx, s->fstream, s->wpi, st, events);
}

}  // namespace

extern "C" {

int pp_solver_create(pp_ctx* ctx, const pp_hmatrix* h, const double* storage, int precision, int workers,
                     int wavefronts_per_inst, int nrhs_max, int rhs_split, pp_solver** out) {
  try {
    return solver_create(ctx, &h, 1, &storage, false, precision, workers, wavefronts_per_inst, nrhs_max, rhs_split,
                         PP_SOLVE_HINV, out);
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_create_device(pp_ctx* ctx, const pp_hmatrix* h, const double* storage_dev, int precision, int workers,
                            int wavefronts_per_inst, int nrhs_max, int rhs_split, pp_solver** out) {
  try {
    return solver_create(ctx, &h, 1, &storage_dev, true, precision, workers, wavefronts_per_inst, nrhs_max, rhs_split,
                         PP_SOLVE_HINV, out);
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_create_batch(pp_ctx* ctx, const pp_hmatrix* const* hs, int count, const double* const* storages,
                           int on_device, int precision, int workers, int wavefronts_per_inst, int nrhs_max,
                           int rhs_split, pp_solver** out) {
  try {
    return solver_create(ctx, hs, count, storages, on_device != 0, precision, workers, wavefronts_per_inst, nrhs_max,
                         rhs_split, PP_SOLVE_HINV, out);
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_create_batch_mode(pp_ctx* ctx, const pp_hmatrix* const* hs, int count, const double* const* storages,
                                int on_device, int precision, int workers, int wavefronts_per_inst, int nrhs_max,
                                int rhs_split, int solve_mode, pp_solver** out) {
  try {
    return solver_create(ctx, hs, count, storages, on_device != 0, precision, workers, wavefronts_per_inst, nrhs_max,
                         rhs_split, solve_mode, out);
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_num_members(const pp_solver* s) { return s ? (int)s->hms.size() : PP_ERR_INVALID; }

int64_t pp_solver_rows(const pp_solver* s) { return s ? s->n_total : PP_ERR_INVALID; }

int pp_solver_member(const pp_solver* s, int i, int64_t* row_off, int* n, double* last_residual) {
  if (s == nullptr || i < 0 || i >= (int)s->hms.size()) return set_error(PP_ERR_INVALID, "bad member");
  if (row_off) *row_off = s->row_off[i];
  if (n) *n = hier(s->hms[i])->getNx_abs();
  if (last_residual) *last_residual = s->member_res[i];
  return PP_OK;
}

int pp_solver_member_arena(const pp_solver* s, int i, int64_t* arena_off) {
  if (s == nullptr || arena_off == nullptr || i < 0 || i >= (int)s->hms.size())
    return set_error(PP_ERR_INVALID, "bad member");
  *arena_off = s->arena_off[i];
  return PP_OK;
}

int64_t pp_solver_storage_elems(const pp_solver* s) { return s ? s->storage_total : PP_ERR_INVALID; }

int pp_solver_use_graph(pp_solver* s, int enable) {
  if (s == nullptr) return set_error(PP_ERR_INVALID, "null solver");
  s->use_graph = enable != 0;
  return PP_OK;
}

int pp_solver_factor(pp_solver* s, void* hip_stream) {
  try {
    if (s == nullptr) return set_error(PP_ERR_INVALID, "null solver");
    hipStream_t st = pick(s->ctx, hip_stream);
    s->factored = false;
    if (!s->use_graph) {
      const int rc = enqueue_factor(s, st, true);
      return rc ? rc : finish_factor(s, st);
    }
    // capture once (restore + temp fills + epoch bump + VM), replay after
    if (s->factor_graph == nullptr || s->graph_stream != st) {
      if (s->factor_graph) {
        HIP_TRY(hipGraphExecDestroy(s->factor_graph));
        s->factor_graph = nullptr;
      }
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_factor(s, st, false);
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&s->factor_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
      s->graph_stream = st;
    }
    HIP_TRY(hipEventRecord(s->fstream->ev0, st));
    HIP_TRY(hipGraphLaunch(s->factor_graph, st));
    HIP_TRY(hipEventRecord(s->fstream->ev1, st));
    s->fstream->launched = true;
    return finish_factor(s, st);
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_residual(pp_solver* s, const double* b, const double* x, double* r, int nrhs, int ld, void* hip_stream) {
  if (s == nullptr || b == nullptr || x == nullptr || r == nullptr || nrhs < 1 || nrhs > s->nrhs_max || ld < nrhs)
    return set_error(PP_ERR_INVALID, "bad residual args");
  hipStream_t st = pick(s->ctx, hip_stream);
  hipError_t e = pp::launch_hmatvec(s->plan, s->d_A, x, ld, b, ld, r, ld, nrhs, s->d_t, pp::HmvExtra(), st);
  if (e != hipSuccess) return hip_fail(e, "hmatvec");
  HIP_TRY(hipStreamSynchronize(st));
  return PP_OK;
}

int pp_solver_solve(pp_solver* s, const double* b, double* x, int nrhs, int ld, double tol, int max_iter,
                    void* hip_stream, pp_solve_report* rep) {
  try {
    if (s == nullptr || b == nullptr || x == nullptr || nrhs < 1 || nrhs > s->nrhs_max || ld < nrhs || max_iter < 0)
      return set_error(PP_ERR_INVALID, "bad solve args");
    if (!s->factored) return set_error(PP_ERR_INVALID, "solve needs a successful factorization (or loaded factors)");
    hipStream_t st = pick(s->ctx, hip_stream);
    const auto t0 = std::chrono::steady_clock::now();
    const int eb = elem_bytes(s->precision);
    const int M = (int)s->hms.size();
    pp_solver::SolveProg* sp = nullptr;
    // hierarchical-inverse solves work at the request's width; VM solve
    // programs at a power-of-two width over zero-padded columns
    const int w = s->hinv ? nrhs : program_width(s, nrhs);
    int rc = PP_OK;
    if (!s->hinv && (rc = solve_prog(s, w, &sp))) return rc;
    // Hierarchical inverse: x = F^{-1} d_rhs (d_rhs, pitch w), the U^{-1}
    // product going straight into x (xmode 1: set, 2: add).  VM solve
    // programs solve in place in d_rhs, whose result is then moved into x.
    auto apply_inverse = [&](int xmode) -> int {
      hipError_t e2;
      if (!s->hinv) {
        if (const int r2 = enqueue_launch(s->ctx, sp->stream, s->wpi, st, true)) return r2;
        e2 = xmode == 1 ? pp::launch_copy2d(x, 8, ld, s->d_rhs, eb, w, (int)s->n_total, nrhs, st)
                        : pp::launch_add_correction(x, ld, s->d_rhs, eb, w, (int)s->n_total, nrhs, st);
        return e2 == hipSuccess ? PP_OK : hip_fail(e2, "update x");
      }
      pp::HmvExtra to_x;
      to_x.xo = x;
      to_x.ldxo = ld;
      to_x.xmode = xmode;
      e2 = pp::launch_happly(s->lplan, eb, s->d_F, s->d_rhs, w, s->d_rhs2, w, nrhs, s->d_tinv, pp::HmvExtra(), st);
      if (e2 == hipSuccess)
        e2 = pp::launch_happly(s->uplan, eb, s->d_F, s->d_rhs2, w, nullptr, w, nrhs, s->d_tinv, to_x, st);
      return e2 == hipSuccess ? PP_OK : hip_fail(e2, "apply hierarchical inverse");
    };
    // norms: r of member m, column c at h_nrm[m nrhs + c]; b at the same
    // place + M nrhs_max (copied with the first residual's norms)
    const double* rn = s->h_nrm;
    const double* bn = s->h_nrm + (int64_t)M * s->nrhs_max;
    double* d_bn = s->d_nrm + (int64_t)M * s->nrhs_max;
    hipError_t e = pp::launch_segnorm2(b, ld, s->d_seg, M, nrhs, s->norm_chunks, s->d_npart, d_bn, st);
    if (e != hipSuccess) return hip_fail(e, "norm b");
    // x = F^{-1} b
    if (w > nrhs) {
      e = pp::launch_fill_zero(s->d_rhs, s->n_total * w, eb, st);
      if (e != hipSuccess) return hip_fail(e, "clear rhs");
    }
    e = pp::launch_copy2d(s->d_rhs, eb, w, b, 8, ld, (int)s->n_total, nrhs, st);
    if (e != hipSuccess) return hip_fail(e, "copy rhs");
    if ((rc = apply_inverse(1))) return rc;
    // the residual product leaves r in d_rhs (the next sweep's right-hand
    // side) and its workgroups' column sums of squares in d_rpart
    pp::HmvExtra to_rhs;
    to_rhs.out2 = s->d_rhs;
    to_rhs.out2_bytes = eb;
    to_rhs.ld2 = w;
    to_rhs.nrm_part = s->d_rpart;
    double res = 0.;
    int it = 0;
    bool conv = false;
    for (;;) {
      // r = b - A x (FP64)
      e = pp::launch_hmatvec(s->plan, s->d_A, x, ld, b, ld, nullptr, nrhs, nrhs, s->d_t, to_rhs, st);
      if (e != hipSuccess) return hip_fail(e, "hmatvec");
      e = pp::launch_rownorm_sum(s->d_rpart, s->d_seg_tiles, s->plan.n_tiles, M, nrhs, s->d_nrm, st);
      if (e != hipSuccess) return hip_fail(e, "norm r");
      const int64_t ncopy = it == 0 ? 2LL * M * s->nrhs_max : (int64_t)M * nrhs;
      HIP_TRY(hipMemcpyAsync(s->h_nrm, s->d_nrm, 8 * ncopy, hipMemcpyDeviceToHost, st));
      if ((rc = wait_host(s, st))) return rc;
      if (it == 0 && sp != nullptr && (rc = check_status(sp->stream))) return rc;
      res = 0.;
      for (int m = 0; m < M; ++m) {
        double worst = 0.;
        for (int c = 0; c < nrhs; ++c) {
          const double bb = bn[m * nrhs + c], rr = rn[m * nrhs + c];
          worst = std::max(worst, bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr));
        }
        s->member_res[m] = worst;
        res = std::max(res, worst);
      }
      if (res <= tol) {
        conv = true;
        break;
      }
      if (it >= max_iter) break;
      ++it;
      if ((rc = apply_inverse(2))) return rc;
    }
    if (sp != nullptr && (rc = check_status(sp->stream))) return rc;
    if (rep) {
      rep->iterations = it;
      rep->converged = conv ? 1 : 0;
      rep->residual = res;
      rep->factor_ms = s->last_factor_ms;
      rep->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      rep->solve_ms = rep->total_ms;
    }
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_factor_arena(pp_solver* s, void** dptr, int* eb) {
  if (s == nullptr || dptr == nullptr || eb == nullptr) return set_error(PP_ERR_INVALID, "null");
  *dptr = s->d_F;
  *eb = elem_bytes(s->precision);
  return PP_OK;
}

// FACT payload: i32 precision | i32 solve_mode | i32 members | i32
// elem_bytes | i64 storage_total | i64 n_total | per member: u64 structure
// fingerprint | i64 arena_off | i64 row_off | u64 FNV-1a 64 of the FP64
// arenas | i64 elements | factor arenas (elements x elem_bytes).  The FP64
// arenas stay with the solver that loads the string: their checksum ties the
// factors to the matrices they were computed from.
namespace {

int fp64_arena_sum(pp_solver* s, uint64_t* sum) {
  std::vector<double> a((size_t)s->storage_total);
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(a.data(), s->d_A, 8 * a.size(), hipMemcpyDeviceToHost));
  *sum = ppw::fnv1a64(a.data(), 8 * a.size());
  return PP_OK;
}

void put_factor_header(ppw::writer& w, pp_solver* s, uint64_t arena_sum) {
  w.put<int32_t>(s->precision);
  w.put<int32_t>(s->solve_mode);
  w.put<int32_t>((int32_t)s->hms.size());
  w.put<int32_t>(elem_bytes(s->precision));
  w.put<int64_t>(s->storage_total);
  w.put<int64_t>(s->n_total);
  for (size_t m = 0; m < s->hms.size(); ++m) {
    w.put<uint64_t>(ppc::tree_fingerprint(s->hms[m]));
    w.put<int64_t>(s->arena_off[m]);
    w.put<int64_t>(s->row_off[m]);
  }
  w.put<uint64_t>(arena_sum);
}

}  // namespace

int64_t pp_solver_encode_factors(pp_solver* s, uint8_t* buf, uint64_t cap) {
  try {
    if (s == nullptr) return set_error(PP_ERR_INVALID, "null solver");
    const uint64_t eb = (uint64_t)elem_bytes(s->precision);
    ppw::writer head;
    put_factor_header(head, s, 0);
    const uint64_t payload = head.bytes.size() + 8 + eb * (uint64_t)s->storage_total;
    const uint64_t total = 16 + 16 + (payload + 7) / 8 * 8 + 8;
    if (buf == nullptr || cap < total) return (int64_t)total;
    uint64_t sum = 0;
    int rc = fp64_arena_sum(s, &sum);
    if (rc) return rc;
    ppw::writer w;
    put_factor_header(w, s, sum);
    w.put<int64_t>(s->storage_total);
    const size_t at = w.bytes.size();
    w.bytes.resize(at + eb * (size_t)s->storage_total);
    HIP_TRY(hipMemcpy(w.bytes.data() + at, s->d_F, eb * (size_t)s->storage_total, hipMemcpyDeviceToHost));
    ppw::container c;
    c.add(ppw::kFactor, std::move(w.bytes));
    if (c.size() != total) return set_error(PP_ERR_INTERNAL, "factor string size");
    c.write(buf);
    return (int64_t)total;
  } catch (const ppw::wire_error& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_decode_factors(pp_solver* s, const uint8_t* buf, uint64_t len) {
  try {
    if (s == nullptr) return set_error(PP_ERR_INVALID, "null solver");
    const auto secs = ppw::parse(buf, len);
    const ppw::section* f = ppw::find(secs, ppw::kFactor);
    if (f == nullptr) throw ppw::wire_error("no FACT section");
    ppw::writer mine;
    put_factor_header(mine, s, 0);
    const size_t hb = mine.bytes.size();
    ppw::reader r(f->data, f->length);
    r.need(hb);
    // everything up to the arena checksum must match this solver exactly
    if (std::memcmp(r.here(), mine.bytes.data(), hb - 8) != 0)
      throw ppw::wire_error("factors belong to another solver layout (precision, solve mode, members or structure)");
    r.skip(hb - 8);
    const uint64_t sum = r.get<uint64_t>();
    const int64_t n = r.count(elem_bytes(s->precision));
    if (n != s->storage_total || r.left() != (uint64_t)n * elem_bytes(s->precision))
      throw ppw::wire_error("factor arena size");
    uint64_t here = 0;
    const int rc = fp64_arena_sum(s, &here);
    if (rc) return rc;
    if (sum != here) throw ppw::wire_error("factors were computed from other matrix entries");
    s->factored = false;
    HIP_TRY(hipMemcpy(s->d_F, r.here(), (size_t)n * elem_bytes(s->precision), hipMemcpyHostToDevice));
    s->factored = true;
    return PP_OK;
  } catch (const ppw::wire_error& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (...) {
    return from_exception();
  }
}

int pp_solver_destroy(pp_solver* s) {
  destroy_solver(s);
  return PP_OK;
}

pp_stream* pp_solver_factor_stream(pp_solver* s) { return s ? s->fstream : nullptr; }
double pp_solver_factor_flops(const pp_solver* s) { return s ? s->factor_flops : 0.; }
double pp_solver_last_factor_ms(const pp_solver* s) { return s ? s->last_factor_ms : -1.; }

double pp_solver_solve_flops(const pp_solver* s, int nrhs) {
  if (s == nullptr) return 0.;
  double e = 0.;
  for (const pp_hmatrix* h : s->hms) e += (double)hier(h)->storedEntries();
  return 2.0 * e * nrhs;
}

}  // extern "C"
