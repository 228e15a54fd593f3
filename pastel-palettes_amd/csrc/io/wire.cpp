// Encoders and decoders of the byte-string format (wire.h) for H-matrices
// (structure + FP64 arena) and instruction programs.
//
// TREE payload: i32 ptr_id | i32 0 | i64 storage | i64 nodes | nodes in
//   pre-order, each i32 type | i32 ny | i32 nx | i32 rank | i32 py | i32 px
//   followed, for a hierarchical node, by i32 row_sizes[py] | i32
//   col_sizes[px] and then its py*px children row-major.  Positions, levels
//   and arena offsets are not stored: the decoder recomputes them with the
//   same storage plan and checks the arena size against `storage`.
// ARNA payload: i64 count | f64 elements[count] (count = storage).
// PROG payload: i32 workers | i32 n_slots | i32 n_insts | i32 n_members |
//   i32 flags (bit 0: a window after the first, temps carried over) | i32 0 |
//   i64 n_edges | i64 flops | i64 primary_flops | i64 scratch |
//   i64 comm_waits | f64 dag_density | i64 temp_elems[n_members] |
//   per instruction: i32 operation | i32 n_words | i64 flops | i64 scratch |
//     i32 words[n_words] (id slot 0) |
//   per instruction: i32 n_preds | (i32 inst, i32 mask)[n_preds] |
//   per worker: i32 n_entries | (i32 inst, i32 exec, i64 start_flops)[..] |
//   per worker: i32 n_words | i32 words[n_words].
// A decoded program's streams are checked against its instructions: every
// execute entry must repeat its instruction's words (with the id filled in),
// every wait must name an instruction, every stream must end with finish and
// every instruction must execute exactly once.
#include "wire.h"

#include <algorithm>
#include <memory>

#include "../../../include/pastel.h"
#include "../capi_common.h"

using ppc::set_error;
using ppw::wire_error;

namespace {

constexpr int kMaxDepth = 48;
constexpr int kMaxDim = 1 << 28;
constexpr int kMaxSplit = 64;

void put_tree(ppw::writer& w, const h_block* b, int64_t& nodes) {
  ++nodes;
  w.put<int32_t>((int32_t)b->type);
  w.put<int32_t>(b->ny);
  w.put<int32_t>(b->nx);
  w.put<int32_t>(b->type == element_t::low_rank ? b->rank : 0);
  w.put<int32_t>(b->type == element_t::hierarchical ? b->py : 0);
  w.put<int32_t>(b->type == element_t::hierarchical ? b->px : 0);
  if (b->type != element_t::hierarchical) return;
  w.put_vec(b->row_sizes);
  w.put_vec(b->col_sizes);
  for (const h_block* c : b->children) put_tree(w, c, nodes);
}

std::vector<uint8_t> tree_payload(const Hierarchical* h) {
  ppw::writer body;
  int64_t nodes = 0;
  put_tree(body, h->getRoot(), nodes);
  ppw::writer w;
  w.put<int32_t>(h->getPtrId());
  w.put<int32_t>(0);
  w.put<int64_t>(h->getStorage());
  w.put<int64_t>(nodes);
  w.put_vec(body.bytes);
  return std::move(w.bytes);
}

h_block* get_tree(ppw::reader& r, int row0, int col0, int level, int64_t& budget) {
  if (level > kMaxDepth) throw wire_error("tree deeper than the format allows");
  if (--budget < 0) throw wire_error("more nodes than declared");
  const int type = r.get<int32_t>(), ny = r.get<int32_t>(), nx = r.get<int32_t>(), rank = r.get<int32_t>(),
            py = r.get<int32_t>(), px = r.get<int32_t>();
  if (ny < 1 || nx < 1 || ny > kMaxDim || nx > kMaxDim) throw wire_error("bad block size");
  auto b = std::make_unique<h_block>();
  b->ny = ny;
  b->nx = nx;
  b->row0 = row0;
  b->col0 = col0;
  b->level = level;
  if (type == (int)element_t::dense) {
    if (rank != 0 || py != 0 || px != 0) throw wire_error("bad dense block");
    b->type = element_t::dense;
  } else if (type == (int)element_t::low_rank) {
    if (rank < 1 || rank > kMaxDim || py != 0 || px != 0) throw wire_error("bad low-rank block");
    b->type = element_t::low_rank;
    b->rank = rank;
  } else if (type == (int)element_t::hierarchical) {
    if (rank != 0 || py < 1 || px < 1 || py > kMaxSplit || px > kMaxSplit) throw wire_error("bad hierarchical block");
    b->type = element_t::hierarchical;
    b->py = py;
    b->px = px;
    b->row_sizes = r.get_vec<int32_t>(py);
    b->col_sizes = r.get_vec<int32_t>(px);
    int64_t sy = 0, sx = 0;
    for (int v : b->row_sizes) {
      if (v < 1) throw wire_error("bad partition");
      sy += v;
    }
    for (int v : b->col_sizes) {
      if (v < 1) throw wire_error("bad partition");
      sx += v;
    }
    if (sy != ny || sx != nx) throw wire_error("partition does not cover the block");
    b->children.assign((size_t)py * px, nullptr);
    int ro = row0;
    for (int i = 0; i < py; ++i) {
      int co = col0;
      for (int j = 0; j < px; ++j) {
        b->children[(size_t)i * px + j] = get_tree(r, ro, co, level + 1, budget);
        co += b->col_sizes[j];
      }
      ro += b->row_sizes[i];
    }
  } else {
    throw wire_error("unknown block type");
  }
  return b.release();
}

std::unique_ptr<Hierarchical> decode_tree(const ppw::section& s) {
  ppw::reader r(s.data, s.length);
  const int ptr_id = r.get<int32_t>();
  (void)r.get<int32_t>();
  const int64_t storage = r.get<int64_t>();
  int64_t nodes = r.count(24);
  std::unique_ptr<h_block> root(get_tree(r, 0, 0, 0, nodes));
  if (nodes != 0 || r.left() != 0) throw wire_error("node count does not match the tree");
  if (root->ny != root->nx) throw wire_error("root block is not square");
  auto h = std::make_unique<Hierarchical>(root.release(), ptr_id);
  if (h->getStorage() != storage) throw wire_error("arena size does not match the structure");
  return h;
}

int64_t emit(const ppw::container& c, uint8_t* buf, uint64_t cap) {
  const uint64_t n = c.size();
  if (buf != nullptr && cap >= n) c.write(buf);
  return (int64_t)n;
}

std::vector<uint8_t> program_payload(const pp_program* p) {
  pp_program_info info;
  if (pp_program_get_info(p, &info) != PP_OK) throw wire_error("program info");
  const int n = info.n_insts;
  if (p->tree.length() != n) throw wire_error("program has no instruction list");
  ppw::writer w;
  w.put<int32_t>(p->workers);
  w.put<int32_t>(p->n_slots);
  w.put<int32_t>(n);
  w.put<int32_t>((int32_t)p->temp_elems.size());
  w.put<int32_t>(p->zero_temps ? 0 : 1);
  w.put<int32_t>(0);
  w.put<int64_t>(info.n_edges);
  w.put<int64_t>(info.flops);
  w.put<int64_t>(p->primary_flops);
  w.put<int64_t>(p->scratch);
  w.put<int64_t>(info.comm_waits);
  w.put<double>(info.dag_density);
  w.put_vec(p->temp_elems);
  for (int i = 0; i < n; ++i) {
    const h_op& o = p->tree.getOp(i);
    w.put<int32_t>((int32_t)o.op);
    w.put<int32_t>((int32_t)o.words.size());
    w.put<int64_t>(o.flops);
    w.put<int64_t>(o.scratch);
    w.put_vec(o.words);
  }
  std::vector<int32_t> preds, masks;
  for (int i = 0; i < n; ++i) {
    const int k = pp_program_inst_preds(p, i, nullptr, nullptr, 0);
    preds.resize(k);
    masks.resize(k);
    pp_program_inst_preds(p, i, preds.data(), masks.data(), k);
    w.put<int32_t>(k);
    for (int j = 0; j < k; ++j) {
      w.put<int32_t>(preds[j]);
      w.put<int32_t>(masks[j]);
    }
  }
  for (int wk = 0; wk < p->workers; ++wk) {
    const int k = pp_program_schedule_length(p, wk);
    std::vector<int32_t> insts(k), ex(k);
    std::vector<int64_t> start(k);
    pp_program_schedule_entries(p, wk, insts.data(), ex.data(), start.data());
    w.put<int32_t>(k);
    for (int j = 0; j < k; ++j) {
      w.put<int32_t>(insts[j]);
      w.put<int32_t>(ex[j]);
      w.put<int64_t>(start[j]);
    }
  }
  for (int wk = 0; wk < p->workers; ++wk) {
    const auto& s = ppc::program_stream(p, wk);
    w.put<int32_t>((int32_t)s.size());
    w.put_vec(s);
  }
  return std::move(w.bytes);
}

// a count read as i32 that must fit in what remains at `unit` bytes per item
int get_count32(ppw::reader& r, int64_t unit) {
  const int32_t c = r.get<int32_t>();
  if (c < 0 || (unit > 0 && (uint64_t)c > r.left() / (uint64_t)unit)) throw wire_error("bad count");
  return c;
}

std::unique_ptr<pp_program> decode_program(const ppw::section& s) {
  ppw::reader r(s.data, s.length);
  auto p = std::make_unique<pp_program>();
  auto im = std::make_unique<pp_program_image>();
  p->workers = r.get<int32_t>();
  p->n_slots = r.get<int32_t>();
  const int n = r.get<int32_t>();
  const int members = r.get<int32_t>();
  const int flags = r.get<int32_t>();
  (void)r.get<int32_t>();
  // member i owns slots 3i (matrix), 3i + 1 (temps) and 3i + 2 (right-hand
  // sides); the upload indexes the temp slot of every member
  if (p->workers < 1 || p->workers > (1 << 16) || p->n_slots < 0 || p->n_slots > (1 << 24) || n < 0 ||
      members < 0 || (members > 0 && 3 * (int64_t)members - 2 >= p->n_slots) || (flags & ~1) != 0)
    throw wire_error("bad program header");
  p->zero_temps = (flags & 1) == 0;
  im->n_insts = n;
  im->n_edges = r.get<int64_t>();
  im->flops = r.get<int64_t>();
  p->primary_flops = r.get<int64_t>();
  p->scratch = r.get<int64_t>();
  im->comm_waits = r.get<int64_t>();
  im->dag_density = r.get<double>();
  if (p->scratch < 0) throw wire_error("bad scratch size");
  if ((uint64_t)members > r.left() / 8) throw wire_error("bad member count");
  p->temp_elems = r.get_vec<int64_t>(members);
  for (int64_t t : p->temp_elems)
    if (t < 0) throw wire_error("bad temp size");
  if ((uint64_t)n > r.left() / 24) throw wire_error("bad instruction count");
  for (int i = 0; i < n; ++i) {
    h_op o;
    const int op = r.get<int32_t>();
    const int nw = get_count32(r, 4);
    o.flops = r.get<int64_t>();
    o.scratch = r.get<int64_t>();
    o.words = r.get_vec<int>(nw);
    if (op < (int)nop || op > (int)trmm || nw != operationLength(op) || o.words[0] != execute || o.words[2] != op ||
        o.scratch < 0)
      throw wire_error("bad instruction " + std::to_string(i));
    // operand words as the device decoder reads them; the worker scratch the
    // routine touches must fit in the program's per-worker scratch
    std::string why;
    const int64_t need = h_op_check_words(o.words.data(), nw, p->n_slots, &why);
    if (need < 0) throw wire_error("instruction " + std::to_string(i) + ": " + why);
    if (need > p->scratch) throw wire_error("instruction " + std::to_string(i) + " needs more scratch than the program");
    o.op = (operation_t)op;
    p->tree.push(std::move(o));
  }
  im->preds.resize(n);
  for (int i = 0; i < n; ++i) {
    const int k = get_count32(r, 8);
    im->preds[i] = r.get_vec<int32_t>(2 * (int64_t)k);
    for (int j = 0; j < k; ++j)
      if (im->preds[i][2 * j] < 0 || im->preds[i][2 * j] >= n) throw wire_error("bad predecessor");
  }
  im->sched.resize(p->workers);
  for (int w = 0; w < p->workers; ++w) {
    const int k = get_count32(r, 16);
    auto& v = im->sched[w];
    v.resize(k);
    for (int j = 0; j < k; ++j) {
      v[j].inst = r.get<int32_t>();
      v[j].ex_w = r.get<int32_t>();
      v[j].start_flops = r.get<int64_t>();
      if (v[j].inst < 0 || v[j].inst >= n || (v[j].ex_w != 0 && v[j].ex_w != 1)) throw wire_error("bad schedule entry");
    }
  }
  im->streams.resize(p->workers);
  std::vector<int> executed(n, 0);
  for (int w = 0; w < p->workers; ++w) {
    const int k = get_count32(r, 4);
    auto& st = im->streams[w];
    st = r.get_vec<int32_t>(k);
    size_t at = 0;
    bool ended = false;
    while (at < st.size() && !ended) {
      const int oc = st[at];
      if (oc == finish) {
        ended = true;
        at += 1;
      } else if (oc == signal_wait) {
        if (at + 3 > st.size() || st[at + 1] < 0 || st[at + 1] >= n) throw wire_error("bad wait entry");
        at += 3;
      } else if (oc == execute) {
        if (at + 2 > st.size() || st[at + 1] < 0 || st[at + 1] >= n) throw wire_error("bad execute entry");
        const int id = st[at + 1];
        const auto& ow = p->tree.getOp(id).words;
        if (at + ow.size() > st.size() || st[at] != ow[0] ||
            !std::equal(ow.begin() + 2, ow.end(), st.begin() + (ptrdiff_t)at + 2))
          throw wire_error("stream entry differs from instruction " + std::to_string(id));
        ++executed[id];
        at += ow.size();
      } else {
        throw wire_error("bad opcode in stream");
      }
    }
    if (!ended || at != st.size()) throw wire_error("stream does not end with finish");
  }
  for (int i = 0; i < n; ++i)
    if (executed[i] != 1) throw wire_error("instruction " + std::to_string(i) + " not executed exactly once");
  if (r.left() != 0) throw wire_error("trailing bytes in program section");
  p->image = std::move(im);
  return p;
}

}  // namespace

extern "C" {

int64_t pp_hmatrix_encode(const pp_hmatrix* h, const double* arena, uint8_t* buf, uint64_t cap) {
  try {
    if (h == nullptr) return set_error(PP_ERR_INVALID, "null matrix");
    ppw::container c;
    c.add(ppw::kTree, tree_payload(h->h.get()));
    if (arena != nullptr) {
      ppw::writer w;
      const int64_t n = h->h->getStorage();
      w.put<int64_t>(n);
      w.put_bytes(arena, 8 * (size_t)n);
      c.add(ppw::kArena, std::move(w.bytes));
    }
    return emit(c, buf, cap);
  } catch (const wire_error& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (...) {
    return ppc::from_exception();
  }
}

int pp_hmatrix_decode(const uint8_t* buf, uint64_t len, double* arena, uint64_t arena_cap, int* has_arena,
                      pp_hmatrix** out) {
  try {
    if (out == nullptr) return set_error(PP_ERR_INVALID, "null output");
    const auto secs = ppw::parse(buf, len);
    const ppw::section* t = ppw::find(secs, ppw::kTree);
    if (t == nullptr) throw wire_error("no TREE section");
    std::unique_ptr<Hierarchical> h = decode_tree(*t);
    const ppw::section* a = ppw::find(secs, ppw::kArena);
    if (a != nullptr) {
      ppw::reader r(a->data, a->length);
      const int64_t n = r.count(8);
      if (n != h->getStorage() || r.left() != 8 * (uint64_t)n) throw wire_error("arena size does not match the structure");
      if (arena != nullptr) {
        if (arena_cap < (uint64_t)n) return set_error(PP_ERR_INVALID, "arena buffer too small");
        r.get_bytes(arena, 8 * (size_t)n);
      }
    } else if (arena != nullptr) {
      return set_error(PP_ERR_INVALID, "string holds no arena");
    }
    if (has_arena) *has_arena = a != nullptr;
    auto m = std::make_unique<pp_hmatrix>();
    m->h = std::move(h);
    m->h->leaves(m->leaves);
    *out = m.release();
    return PP_OK;
  } catch (const wire_error& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (...) {
    return ppc::from_exception();
  }
}

int64_t pp_program_encode(const pp_program* p, uint8_t* buf, uint64_t cap) {
  try {
    if (p == nullptr) return set_error(PP_ERR_INVALID, "null program");
    ppw::container c;
    c.add(ppw::kProgram, program_payload(p));
    return emit(c, buf, cap);
  } catch (const wire_error& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (...) {
    return ppc::from_exception();
  }
}

int pp_program_decode(const uint8_t* buf, uint64_t len, pp_program** out) {
  try {
    if (out == nullptr) return set_error(PP_ERR_INVALID, "null output");
    const auto secs = ppw::parse(buf, len);
    const ppw::section* s = ppw::find(secs, ppw::kProgram);
    if (s == nullptr) throw wire_error("no PROG section");
    *out = decode_program(*s).release();
    return PP_OK;
  } catch (const wire_error& e) {
    return set_error(PP_ERR_INVALID, e.what());
  } catch (...) {
    return ppc::from_exception();
  }
}

}  // extern "C"

uint64_t ppc::tree_fingerprint(const pp_hmatrix* h) {
  const std::vector<uint8_t> t = tree_payload(h->h.get());
  return ppw::fnv1a64(t.data(), t.size());
}
