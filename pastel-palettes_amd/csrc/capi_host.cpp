// C ABI, host half: rts queries, H-matrix structures and programs.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pastel.h"
#include "capi_common.h"
#include "h_ops/h_ops.h"
#include "instructions/instructions_manager.h"
#include "matrix/Hierarchical.h"
#include "rts/h_ops_dag.h"
#include "rts/scheduler.h"

namespace {
thread_local std::string g_last_error;
}

namespace ppc {

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int from_exception() {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return set_error(PP_ERR_NOMEM, "host allocation failed");
  } catch (const std::invalid_argument& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (const std::out_of_range& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (const std::exception& e) {
    return set_error(PP_ERR_INTERNAL, e.what());
  } catch (...) {
    return set_error(PP_ERR_INTERNAL, "unknown exception");
  }
}

}  // namespace ppc

using ppc::from_exception;
using ppc::set_error;

struct pp_schedule {
  std::unique_ptr<h_ops_dag> dag;
  std::unique_ptr<instructions_scheduler> sched;
};

extern "C" {

const char* pp_last_error(void) { return g_last_error.c_str(); }
const char* pp_version(void) { return "pastel-mi355x 0.1"; }
int64_t pp_min_inst_flops(void) { return _MIN_INST_FLOPS; }

int pp_rts_check_overlap(uint64_t ptr_a, uint64_t pitch_a, uint64_t width_a, uint64_t height_a, uint64_t ptr_b,
                         uint64_t pitch_b, uint64_t width_b, uint64_t height_b) {
  try {
    const DataMap2D a(reinterpret_cast<void*>(ptr_a), pitch_a, width_a, height_a);
    const DataMap2D b(reinterpret_cast<void*>(ptr_b), pitch_b, width_b, height_b);
    return a.checkOverlap(b) ? 1 : 0;
  } catch (...) {
    return from_exception();
  }
}

int pp_rts_check_dependency(int n_from, const uint64_t* maps_from, const int32_t* ro_from, int n_to,
                            const uint64_t* maps_to, const int32_t* ro_to) {
  try {
    if (n_from < 0 || n_to < 0) return set_error(PP_ERR_INVALID, "negative operand count");
    Operation from, to;
    for (int i = 0; i < n_from; ++i)
      from.addData(ro_from[i] != 0, DataMap2D(reinterpret_cast<void*>(maps_from[4 * i]), maps_from[4 * i + 1],
                                              maps_from[4 * i + 2], maps_from[4 * i + 3]));
    for (int i = 0; i < n_to; ++i)
      to.addData(ro_to[i] != 0, DataMap2D(reinterpret_cast<void*>(maps_to[4 * i]), maps_to[4 * i + 1],
                                          maps_to[4 * i + 2], maps_to[4 * i + 3]));
    return (int)to.checkDependencyFrom(from);
  } catch (...) {
    return from_exception();
  }
}

struct pp_dep_table {
  dependency_table t;
};

int pp_dep_table_create(pp_dep_table** out) {
  try {
    if (out == nullptr) return set_error(PP_ERR_INVALID, "null out");
    *out = new pp_dep_table();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

void pp_dep_table_destroy(pp_dep_table* t) { delete t; }

int pp_dep_table_update(pp_dep_table* t, int64_t inst_from, int64_t inst_to, int dep) {
  try {
    if (t == nullptr || dep < 0 || dep > 7) return set_error(PP_ERR_INVALID, "bad table update");
    t->t.update(inst_from, inst_to, static_cast<dependency_t>(dep));
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_dep_table_lookup(const pp_dep_table* t, int64_t inst_from, int64_t inst_to) {
  if (t == nullptr) return set_error(PP_ERR_INVALID, "null table");
  return static_cast<int>(t->t.lookup(inst_from, inst_to));
}

int pp_dep_table_list(const pp_dep_table* t, int64_t inst, int to, int64_t* out, int cap) {
  try {
    if (t == nullptr || cap < 0) return set_error(PP_ERR_INVALID, "bad table query");
    std::vector<int64_t> v;
    if (to) t->t.getDepTo(inst, v);
    else t->t.getDepFrom(inst, v);
    for (int i = 0; i < (int)v.size() && i < cap && out != nullptr; ++i) out[i] = v[i];
    return (int)v.size();
  } catch (...) {
    return from_exception();
  }
}

int64_t pp_dep_table_count(const pp_dep_table* t, int64_t inst, int to) {
  if (t == nullptr) return set_error(PP_ERR_INVALID, "null table");
  return (int64_t)(to ? t->t.countDepTo(inst) : t->t.countDepFrom(inst));
}

int pp_dep_table_clear(pp_dep_table* t) {
  if (t == nullptr) return set_error(PP_ERR_INVALID, "null table");
  t->t.clear();
  return PP_OK;
}

int pp_schedule_create(int n, const int64_t* flops, int n_edges, const int32_t* edges, const int32_t* masks,
                       int workers, pp_schedule** out) {
  try {
    if (out == nullptr || n < 0 || n_edges < 0 || workers < 1) return set_error(PP_ERR_INVALID, "bad schedule args");
    std::vector<long long> f(flops, flops + n);
    auto s = std::make_unique<pp_schedule>();
    s->dag = std::make_unique<h_ops_dag>(n, f.data(), n_edges, edges, masks);
    s->sched = std::make_unique<instructions_scheduler>(s->dag.get(), workers);
    *out = s.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_schedule_worker_length(const pp_schedule* s, int worker) {
  if (s == nullptr || worker < 0 || worker >= s->sched->getWorkers()) return set_error(PP_ERR_INVALID, "bad worker");
  return s->sched->getLength(worker);
}

static int fill_entries(const instructions_scheduler* sc, int worker, int32_t* insts, int32_t* ex_w,
                        int64_t* start_flops) {
  if (worker < 0 || worker >= sc->getWorkers()) return set_error(PP_ERR_INVALID, "bad worker");
  int i = 0;
  for (const instructions_queue* q = sc->getSchedule(worker); q != nullptr; q = q->getNext(), ++i) {
    if (insts) insts[i] = q->getInst();
    if (ex_w) ex_w[i] = q->getExW() ? 1 : 0;
    if (start_flops) start_flops[i] = q->getStartingFlops();
  }
  return i;
}

int pp_schedule_worker_entries(const pp_schedule* s, int worker, int32_t* insts, int32_t* ex_w, int64_t* start_flops) {
  if (s == nullptr) return set_error(PP_ERR_INVALID, "null schedule");
  return fill_entries(s->sched.get(), worker, insts, ex_w, start_flops);
}

int pp_schedule_comm_waits(const pp_schedule* s) { return s ? s->sched->getCommWaitCounts() : PP_ERR_INVALID; }

void pp_schedule_destroy(pp_schedule* s) { delete s; }

// ------------------------------------------------------------------ H-matrix

static int make_hmatrix(Hierarchical* h, pp_hmatrix** out) {
  auto m = std::make_unique<pp_hmatrix>();
  m->h.reset(h);
  m->h->leaves(m->leaves);
  *out = m.release();
  return PP_OK;
}

int pp_hmatrix_create_hodlr(int n, int levels, int rank, pp_hmatrix** out) {
  try {
    if (out == nullptr) return set_error(PP_ERR_INVALID, "null out");
    return make_hmatrix(Hierarchical::hodlr(n, levels, rank, 0), out);
  } catch (...) {
    return from_exception();
  }
}

int pp_hmatrix_create_dense(int n, int leaf, pp_hmatrix** out) {
  try {
    if (out == nullptr) return set_error(PP_ERR_INVALID, "null out");
    return make_hmatrix(Hierarchical::denseGrid(n, leaf, 0), out);
  } catch (...) {
    return from_exception();
  }
}

void pp_hmatrix_destroy(pp_hmatrix* h) { delete h; }
int pp_hmatrix_size(const pp_hmatrix* h) { return h ? h->h->getNx_abs() : PP_ERR_INVALID; }
int64_t pp_hmatrix_storage(const pp_hmatrix* h) { return h ? h->h->getStorage() : PP_ERR_INVALID; }
int64_t pp_hmatrix_inverse_base(const pp_hmatrix* h) { return h ? inverse_base(h->h.get()) : PP_ERR_INVALID; }
int64_t pp_hmatrix_inverse_elems(const pp_hmatrix* h) { return h ? h->h->inverseStorage() : PP_ERR_INVALID; }

int pp_hmatrix_inverse_count(const pp_hmatrix* h, int* leaves, int* nodes) {
  try {
    if (h == nullptr) return set_error(PP_ERR_INVALID, "null matrix");
    std::vector<Hierarchical::inv_block> lay;
    std::vector<Hierarchical::inv_node> nd;
    h->h->inverseLayout(lay);
    h->h->inverseNodes(nd);
    if (leaves) *leaves = (int)lay.size();
    if (nodes) *nodes = (int)nd.size();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_hmatrix_inverse_leaf(const pp_hmatrix* h, int i, int64_t info[4]) {
  try {
    if (h == nullptr || info == nullptr) return set_error(PP_ERR_INVALID, "null argument");
    std::vector<Hierarchical::inv_block> lay;
    h->h->inverseLayout(lay);
    if (i < 0 || i >= (int)lay.size()) return set_error(PP_ERR_INVALID, "leaf index out of range");
    info[0] = lay[i].row0;
    info[1] = lay[i].nb;
    info[2] = lay[i].inv_off;
    info[3] = lay[i].leaf_off;
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_hmatrix_inverse_node(const pp_hmatrix* h, int i, int64_t info[9]) {
  try {
    if (h == nullptr || info == nullptr) return set_error(PP_ERR_INVALID, "null argument");
    std::vector<Hierarchical::inv_node> nd;
    h->h->inverseNodes(nd);
    if (i < 0 || i >= (int)nd.size()) return set_error(PP_ERR_INVALID, "node index out of range");
    const auto& n = nd[i];
    const int64_t v[9] = {n.row0, n.na, n.nb, n.rl, n.ru, n.l_u, n.l_v, n.u_u, n.u_v};
    for (int k = 0; k < 9; ++k) info[k] = v[k];
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_hmatrix_inverse_identity(const pp_hmatrix* h, double* area) {
  try {
    if (h == nullptr || area == nullptr) return set_error(PP_ERR_INVALID, "null argument");
    std::vector<Hierarchical::inv_block> lay;
    h->h->inverseLayout(lay);
    std::fill(area, area + h->h->inverseStorage(), 0.0);
    for (const auto& e : lay) {
      const int64_t second = (((int64_t)e.nb * e.nb + 31) / 32) * 32;
      for (int i = 0; i < e.nb; ++i) {
        area[e.inv_off + (int64_t)i * e.nb + i] = 1.0;
        area[e.inv_off + second + (int64_t)i * e.nb + i] = 1.0;
      }
    }
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}
int pp_hmatrix_num_leaves(const pp_hmatrix* h) { return h ? (int)h->leaves.size() : PP_ERR_INVALID; }

int pp_hmatrix_leaf(const pp_hmatrix* h, int i, int64_t info[8]) {
  if (h == nullptr || i < 0 || i >= (int)h->leaves.size()) return set_error(PP_ERR_INVALID, "bad leaf index");
  const h_block* b = h->leaves[i];
  info[0] = (int64_t)b->type;
  info[1] = b->row0;
  info[2] = b->col0;
  info[3] = b->ny;
  info[4] = b->nx;
  info[5] = b->rank;
  info[6] = b->type == element_t::dense ? b->off : b->off_u;
  info[7] = b->off_v;
  return PP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ programs

pp_program::~pp_program() = default;

int ppc::program_insts(const pp_program* p) { return p->dag ? p->dag->getLength() : p->image->n_insts; }

const std::vector<int32_t>& ppc::program_stream(const pp_program* p, int worker) {
  return p->dag ? p->mgr->getStream(worker) : p->image->streams[worker];
}

namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// tree -> DAG -> schedule -> words
void finish_program(pp_program* p, int workers) {
  auto t0 = std::chrono::steady_clock::now();
  p->dag = std::make_unique<h_ops_dag>(&p->tree, 0, 0);
  p->dag_ms = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  p->sched = std::make_unique<instructions_scheduler>(p->dag.get(), workers);
  p->schedule_ms = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  p->mgr = std::make_unique<instructions_manager>(workers, p->dag.get(), p->sched.get(), &p->tree, 0);
  p->encode_ms = ms_since(t0);
  p->workers = workers;
  p->scratch = p->tree.maxScratch();
}

// Two-phase program: ops [0, F) (a factorization) scheduled exactly as the
// factorization alone, ops [F, n) (the inverse construction) by their own
// schedule, appended to every worker's stream after its factorization
// entries.  Each op of the second phase is preceded by waits for the
// first-phase ops it depends on; no first-phase op depends on the second, so
// every wait ends.  The second phase only fills the time a worker has left
// after its factorization work and never delays it.
void finish_two_phase(pp_program* p, int workers, int F) {
  auto t0 = std::chrono::steady_clock::now();
  const int n = p->tree.length(), I = n - F;
  h_ops_dag full(&p->tree, 0, 0);
  h_ops_dag dag_f(&p->tree, 0, F), dag_i(&p->tree, F, I);
  p->dag_ms = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  instructions_scheduler s_f(&dag_f, workers), s_i(&dag_i, workers);
  p->schedule_ms = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  instructions_manager m_f(workers, &dag_f, &s_f, &p->tree, 0), m_i(workers, &dag_i, &s_i, &p->tree, F);
  auto im = std::make_unique<pp_program_image>();
  im->n_insts = n;
  im->n_edges = full.getEdgeCount();
  im->flops = full.getFlops();
  im->dag_density = full.dag_density();
  im->preds.resize(n);
  for (int i = 0; i < n; ++i)
    for (const dependency_linked_list* d = full.getDepList_To(i); d != nullptr; d = d->getNext()) {
      im->preds[i].push_back(d->getInst());
      im->preds[i].push_back((int)d->getDep());
    }
  im->streams.resize(workers);
  im->sched.resize(workers);
  int64_t waits = 0;
  for (int w = 0; w < workers; ++w) {
    std::vector<int32_t>& out = im->streams[w];
    const std::vector<int32_t>& a = m_f.getStream(w);
    out.assign(a.begin(), a.end() - 1);  // without its finish word
    int64_t end_f = 0;
    for (const instructions_queue* q = s_f.getSchedule(w); q != nullptr; q = q->getNext()) {
      im->sched[w].push_back({q->getInst(), q->getExW() ? 1 : 0, q->getStartingFlops()});
      end_f = std::max<int64_t>(end_f, q->getStartingFlops());
      if (!q->getExW()) ++waits;
    }
    const std::vector<int32_t>& b = m_i.getStream(w);
    size_t at = 0;
    const instructions_queue* q = s_i.getSchedule(w);
    while (at < b.size() && b[at] != finish) {
      if (b[at] == signal_wait) {
        out.push_back(signal_wait);
        out.push_back(b[at + 1] + F);
        out.push_back(b[at + 2]);
        at += wait_l;
      } else {
        const int id = b[at + 1] + F, len = operationLength(b[at + 2]);
        for (const dependency_linked_list* d = full.getDepList_To(id); d != nullptr; d = d->getNext())
          if (d->getInst() < F) {
            out.push_back(signal_wait);
            out.push_back(d->getInst());
            out.push_back(0);
            im->sched[w].push_back({d->getInst(), 0, end_f});
            ++waits;
          }
        out.insert(out.end(), b.begin() + (ptrdiff_t)at, b.begin() + (ptrdiff_t)at + len);
        out[out.size() - len + 1] = id;
        at += len;
      }
      if (q != nullptr) {
        im->sched[w].push_back({q->getInst() + F, q->getExW() ? 1 : 0, end_f + q->getStartingFlops()});
        if (!q->getExW()) ++waits;
        q = q->getNext();
      }
    }
    out.push_back(finish);
  }
  im->comm_waits = waits;
  p->encode_ms = ms_since(t0);
  p->workers = workers;
  p->scratch = p->tree.maxScratch();
  p->image = std::move(im);
}

}  // namespace

extern "C" {

int pp_program_factor_batch(const pp_hmatrix* const* hs, int count, int workers, pp_program** out) {
  try {
    if (hs == nullptr || count < 1 || workers < 1 || out == nullptr) return set_error(PP_ERR_INVALID, "bad args");
    auto p = std::make_unique<pp_program>();
    p->n_slots = 3 * count;
    for (int i = 0; i < count; ++i) {
      if (hs[i] == nullptr) return set_error(PP_ERR_INVALID, "null matrix in batch");
      dev_temp tmp(3 * i + 1);
      std::unique_ptr<const h_index> root(hs[i]->h->getRootIndex());
      std::unique_ptr<h_ops_tree> t(hs[i]->h->generateOps_GETRF(root.get(), &tmp, 3 * i));
      p->tree.append(*t);
      p->temp_elems.push_back(tmp.size());
    }
    finish_program(p.get(), workers);
    *out = p.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_program_factor(const pp_hmatrix* h, int workers, pp_program** out) {
  return pp_program_factor_batch(&h, 1, workers, out);
}

int pp_program_factor_num_windows(const pp_hmatrix* h, int window_ops) {
  try {
    if (h == nullptr) return set_error(PP_ERR_INVALID, "null matrix");
    dev_temp tmp(1);
    std::unique_ptr<const h_index> root(h->h->getRootIndex());
    std::unique_ptr<h_ops_tree> t(h->h->generateOps_GETRF(root.get(), &tmp, 0));
    return window_ops <= 0 ? 1 : (t->length() + window_ops - 1) / window_ops;
  } catch (...) {
    return from_exception();
  }
}

int pp_program_factor_window(const pp_hmatrix* h, int workers, int window_ops, int index, pp_program** out) {
  try {
    if (h == nullptr || workers < 1 || out == nullptr) return set_error(PP_ERR_INVALID, "bad args");
    dev_temp tmp(1);
    std::unique_ptr<const h_index> root(h->h->getRootIndex());
    std::unique_ptr<h_ops_tree> t(h->h->generateOps_GETRF(root.get(), &tmp, 0));
    const int len = t->length();
    const int size = window_ops <= 0 ? len : window_ops;
    const int first = index * size;
    if (index < 0 || (first >= len && !(len == 0 && index == 0)))
      return set_error(PP_ERR_INVALID, "window index out of range");
    auto p = std::make_unique<pp_program>();
    p->n_slots = 3;
    for (int i = first; i < std::min(len, first + size); ++i) p->tree.push(h_op(t->getOp(i)));
    // every window addresses the whole factorization's temp area; only the
    // first clears it, later windows read what earlier ones left there
    p->temp_elems.push_back(tmp.size());
    p->zero_temps = index == 0;
    finish_program(p.get(), workers);
    *out = p.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_program_factor_inv_batch(const pp_hmatrix* const* hs, int count, int workers, pp_program** out) {
  return ppc::factor_inverse_program(hs, count, workers, true, out);
}

}  // extern "C"

int ppc::factor_inverse_program(const pp_hmatrix* const* hs, int count, int workers, bool hier, pp_program** out) {
  try {
    if (hs == nullptr || count < 1 || workers < 1 || out == nullptr) return set_error(PP_ERR_INVALID, "bad args");
    auto p = std::make_unique<pp_program>();
    p->n_slots = 3 * count;
    std::vector<dev_temp> tmps;
    for (int i = 0; i < count; ++i) {
      if (hs[i] == nullptr) return set_error(PP_ERR_INVALID, "null matrix in batch");
      tmps.emplace_back(3 * i + 1);
      const Hierarchical* h = hs[i]->h.get();
      std::unique_ptr<const h_index> root(h->getRootIndex());
      std::unique_ptr<h_ops_tree> t(h->generateOps_GETRF(root.get(), &tmps[i], 3 * i));
      p->tree.append(*t);
    }
    p->primary_flops = p->tree.getFlops();
    const int n_factor = p->tree.length();
    // the inversions follow every factorization in program order, so each
    // depends only on the factor blocks it reads; HODLR-shaped matrices get
    // the full hierarchical inverses, others the leaf inverses only
    for (int i = 0; i < count; ++i) {
      const Hierarchical* h = hs[i]->h.get();
      std::unique_ptr<const h_index> root(h->getRootIndex());
      std::unique_ptr<h_ops_tree> v(hier && h->hodlrShaped()
                                        ? h->generateOps_InvertH(root.get(), inverse_base(h), &tmps[i], 3 * i)
                                        : h->generateOps_Invert(root.get(), inverse_base(h), 3 * i));
      p->tree.append(*v);
      p->temp_elems.push_back(tmps[i].size());
    }
    finish_two_phase(p.get(), workers, n_factor);
    *out = p.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

extern "C" {

int pp_program_solve_batch(const pp_hmatrix* const* hs, int count, const int32_t* nrhs, const int32_t* ld,
                           int rhs_split, int workers, pp_program** out) {
  try {
    if (hs == nullptr || count < 1 || workers < 1 || out == nullptr || nrhs == nullptr || ld == nullptr)
      return set_error(PP_ERR_INVALID, "bad args");
    auto p = std::make_unique<pp_program>();
    p->n_slots = 3 * count;
    for (int i = 0; i < count; ++i) {
      if (hs[i] == nullptr) return set_error(PP_ERR_INVALID, "null matrix in batch");
      std::unique_ptr<const h_index> root(hs[i]->h->getRootIndex());
      std::unique_ptr<h_ops_tree> t(
          hs[i]->h->generateOps_Solve(root.get(), 3 * i + 2, 0, nrhs[i], ld[i], rhs_split, 3 * i));
      p->tree.append(*t);
      p->temp_elems.push_back(0);
    }
    finish_program(p.get(), workers);
    *out = p.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_program_solve_inv_batch(const pp_hmatrix* const* hs, int count, const int32_t* nrhs, const int32_t* ld,
                               int rhs_split, int workers, pp_program** out) {
  try {
    if (hs == nullptr || count < 1 || workers < 1 || out == nullptr || nrhs == nullptr || ld == nullptr)
      return set_error(PP_ERR_INVALID, "bad args");
    auto p = std::make_unique<pp_program>();
    p->n_slots = 3 * count;
    for (int i = 0; i < count; ++i) {
      if (hs[i] == nullptr) return set_error(PP_ERR_INVALID, "null matrix in batch");
      const Hierarchical* h = hs[i]->h.get();
      std::unique_ptr<const h_index> root(h->getRootIndex());
      std::unique_ptr<h_ops_tree> t(
          h->generateOps_Solve(root.get(), 3 * i + 2, 0, nrhs[i], ld[i], rhs_split, 3 * i, inverse_base(h)));
      p->tree.append(*t);
      p->temp_elems.push_back(0);
    }
    finish_program(p.get(), workers);
    *out = p.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_program_solve(const pp_hmatrix* h, int nrhs, int ld, int rhs_split, int workers, pp_program** out) {
  const int32_t n = nrhs, l = ld;
  return pp_program_solve_batch(&h, 1, &n, &l, rhs_split, workers, out);
}

void pp_program_destroy(pp_program* p) { delete p; }

int pp_program_get_info(const pp_program* p, pp_program_info* info) {
  if (p == nullptr || info == nullptr) return set_error(PP_ERR_INVALID, "null");
  std::memset(info, 0, sizeof(*info));
  const pp_program_image* im = p->image.get();
  info->n_insts = ppc::program_insts(p);
  info->workers = p->workers;
  info->n_slots = p->n_slots;
  info->n_edges = p->dag ? p->dag->getEdgeCount() : im->n_edges;
  info->flops = p->dag ? p->dag->getFlops() : im->flops;
  int64_t t = 0;
  for (int64_t e : p->temp_elems) t += e;
  info->temp_elems = t;
  info->scratch_elems = p->scratch;
  int64_t words = 0;
  for (int w = 0; w < p->workers; ++w) words += (int64_t)ppc::program_stream(p, w).size();
  info->total_words = words;
  info->comm_waits = p->dag ? p->sched->getCommWaitCounts() : im->comm_waits;
  info->dag_density = p->dag ? p->dag->dag_density() : im->dag_density;
  info->schedule_ms = p->schedule_ms;
  info->dag_ms = p->dag_ms;
  info->encode_ms = p->encode_ms;
  return PP_OK;
}

int64_t pp_program_temp_elems(const pp_program* p, int member) {
  if (p == nullptr || member < 0 || member >= (int)p->temp_elems.size()) return set_error(PP_ERR_INVALID, "bad member");
  return p->temp_elems[member];
}

int pp_program_worker_length(const pp_program* p, int worker) {
  if (p == nullptr || worker < 0 || worker >= p->workers) return set_error(PP_ERR_INVALID, "bad worker");
  return (int)ppc::program_stream(p, worker).size();
}

int pp_program_worker_words(const pp_program* p, int worker, int32_t* out, int cap) {
  if (p == nullptr || worker < 0 || worker >= p->workers) return set_error(PP_ERR_INVALID, "bad worker");
  const auto& s = ppc::program_stream(p, worker);
  const int n = (int)s.size();
  if (out != nullptr) std::memcpy(out, s.data(), sizeof(int32_t) * (size_t)std::min(n, cap));
  return n;
}

int pp_program_inst_words(const pp_program* p, int inst, int32_t* out, int cap) {
  if (p == nullptr || inst < 0 || inst >= p->tree.length()) return set_error(PP_ERR_INVALID, "bad instruction");
  const auto& w = p->tree.getOp(inst).words;
  const int n = (int)w.size();
  if (out != nullptr) {
    std::memcpy(out, w.data(), sizeof(int32_t) * (size_t)std::min(n, cap));
    if (cap > 1) out[1] = inst;
  }
  return n;
}

int64_t pp_program_inst_bytes(const pp_program* p, int inst) {
  if (p == nullptr || inst < 0 || inst >= p->tree.length()) return set_error(PP_ERR_INVALID, "bad instruction");
  int64_t b = 0;
  for (const auto& d : p->tree.getOp(inst).data.getData())
    b += (d.first ? 1 : 2) * (int64_t)(d.second.width * d.second.height);
  return b;
}

int64_t pp_program_inst_flops(const pp_program* p, int inst) {
  if (p == nullptr || inst < 0 || inst >= p->tree.length()) return set_error(PP_ERR_INVALID, "bad instruction");
  return p->tree.getOp(inst).flops;
}

int pp_program_inst_preds(const pp_program* p, int inst, int32_t* preds, int32_t* masks, int cap) {
  if (p == nullptr || inst < 0 || inst >= ppc::program_insts(p)) return set_error(PP_ERR_INVALID, "bad instruction");
  if (!p->dag) {
    const auto& v = p->image->preds[inst];
    const int n = (int)v.size() / 2;
    for (int i = 0; i < n && i < cap; ++i) {
      if (preds) preds[i] = v[2 * i];
      if (masks) masks[i] = v[2 * i + 1];
    }
    return n;
  }
  int n = 0;
  for (const dependency_linked_list* d = p->dag->getDepList_To(inst); d != nullptr; d = d->getNext(), ++n) {
    if (n < cap) {
      if (preds) preds[n] = d->getInst();
      if (masks) masks[n] = (int)d->getDep();
    }
  }
  return n;
}

int pp_program_schedule_length(const pp_program* p, int worker) {
  if (p == nullptr || worker < 0 || worker >= p->workers) return set_error(PP_ERR_INVALID, "bad worker");
  return p->dag ? p->sched->getLength(worker) : (int)p->image->sched[worker].size();
}

int pp_program_schedule_entries(const pp_program* p, int worker, int32_t* insts, int32_t* ex_w, int64_t* start_flops) {
  if (p == nullptr) return set_error(PP_ERR_INVALID, "null program");
  if (p->dag) return fill_entries(p->sched.get(), worker, insts, ex_w, start_flops);
  if (worker < 0 || worker >= p->workers) return set_error(PP_ERR_INVALID, "bad worker");
  const auto& v = p->image->sched[worker];
  for (size_t i = 0; i < v.size(); ++i) {
    if (insts) insts[i] = v[i].inst;
    if (ex_w) ex_w[i] = v[i].ex_w;
    if (start_flops) start_flops[i] = v[i].start_flops;
  }
  return (int)v.size();
}

}  // extern "C"
