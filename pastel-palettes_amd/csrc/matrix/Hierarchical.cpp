// H-matrix structure builders and storage planning.
#include "Hierarchical.h"

#include <stdexcept>

h_block::~h_block() {
  for (h_block* c : children) delete c;
}

namespace {

constexpr int64_t kAlign = 32;  // elements

int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

h_block* make_leaf(element_t t, int ny, int nx, int r0, int c0, int level, int rank = 0) {
  h_block* b = new h_block();
  b->type = t;
  b->ny = ny;
  b->nx = nx;
  b->row0 = r0;
  b->col0 = c0;
  b->level = level;
  b->rank = rank;
  return b;
}

h_block* make_hier(int ny, int nx, int r0, int c0, int level, const std::vector<int>& rows, const std::vector<int>& cols) {
  h_block* b = make_leaf(element_t::hierarchical, ny, nx, r0, c0, level);
  b->py = (int)rows.size();
  b->px = (int)cols.size();
  b->row_sizes = rows;
  b->col_sizes = cols;
  b->children.assign((size_t)b->py * b->px, nullptr);
  return b;
}

// Full quad-tree of dense blocks.
h_block* dense_tree(int ny, int nx, int r0, int c0, int level, int leaf) {
  if (ny <= leaf && nx <= leaf) return make_leaf(element_t::dense, ny, nx, r0, c0, level);
  const int y1 = ny > leaf ? ny / 2 : ny, x1 = nx > leaf ? nx / 2 : nx;
  std::vector<int> rows = (ny > leaf) ? std::vector<int>{y1, ny - y1} : std::vector<int>{ny};
  std::vector<int> cols = (nx > leaf) ? std::vector<int>{x1, nx - x1} : std::vector<int>{nx};
  h_block* b = make_hier(ny, nx, r0, c0, level, rows, cols);
  int ro = r0;
  for (int i = 0; i < b->py; ++i) {
    int co = c0;
    for (int j = 0; j < b->px; ++j) {
      b->children[(size_t)i * b->px + j] = dense_tree(rows[i], cols[j], ro, co, level + 1, leaf);
      co += cols[j];
    }
    ro += rows[i];
  }
  return b;
}

h_block* hodlr_tree(int n, int r0, int level, int levels, int rank) {
  if (level == levels) return make_leaf(element_t::dense, n, n, r0, r0, level);
  const int n0 = n / 2, n1 = n - n0;
  h_block* b = make_hier(n, n, r0, r0, level, {n0, n1}, {n0, n1});
  b->children[0] = hodlr_tree(n0, r0, level + 1, levels, rank);
  b->children[1] = make_leaf(element_t::low_rank, n0, n1, r0, r0 + n0, level + 1, rank);
  b->children[2] = make_leaf(element_t::low_rank, n1, n0, r0 + n0, r0, level + 1, rank);
  b->children[3] = hodlr_tree(n1, r0 + n0, level + 1, levels, rank);
  return b;
}

void collect_leaves(const h_block* b, std::vector<const h_block*>& out) {
  if (b->type == element_t::hierarchical) {
    for (const h_block* c : b->children) collect_leaves(c, out);
  } else {
    out.push_back(b);
  }
}

}  // namespace

int64_t Hierarchical::plan(h_block* b, int64_t cursor) {
  switch (b->type) {
    case element_t::dense:
      b->off = cursor;
      return align_up(cursor + (int64_t)b->ny * b->nx);
    case element_t::low_rank:
      b->off_u = cursor;
      cursor = align_up(cursor + (int64_t)b->ny * b->rank);
      b->off_v = cursor;
      return align_up(cursor + (int64_t)b->nx * b->rank);
    case element_t::hierarchical:
      for (h_block* c : b->children) cursor = plan(c, cursor);
      return cursor;
    default:
      return cursor;
  }
}

Hierarchical::Hierarchical(h_block* r, int id) : root(r), storage(0), ptr_id(id) {
  if (root == nullptr) throw std::invalid_argument("Hierarchical: null root");
  storage = plan(root, 0);
}

Hierarchical::~Hierarchical() { delete root; }

Hierarchical* Hierarchical::denseGrid(int n, int leaf, int id) {
  if (n <= 0 || leaf <= 0) throw std::invalid_argument("denseGrid: sizes must be positive");
  return new Hierarchical(dense_tree(n, n, 0, 0, 0, leaf), id);
}

Hierarchical* Hierarchical::hodlr(int n, int levels, int rank, int id) {
  if (n <= 0 || levels < 0 || rank <= 0) throw std::invalid_argument("hodlr: bad parameters");
  if ((n >> levels) < 1) throw std::invalid_argument("hodlr: too many levels for n");
  return new Hierarchical(hodlr_tree(n, 0, 0, levels, rank), id);
}

void Hierarchical::leaves(std::vector<const h_block*>& out) const {
  out.clear();
  collect_leaves(root, out);
}

int64_t Hierarchical::storedEntries() const {
  std::vector<const h_block*> lv;
  leaves(lv);
  int64_t s = 0;
  for (const h_block* b : lv) {
    if (b->type == element_t::dense) s += (int64_t)b->ny * b->nx;
    else if (b->type == element_t::low_rank) s += (int64_t)(b->ny + b->nx) * b->rank;
  }
  return s;
}

namespace {

// Dense leaves met along the diagonal, in the order the solve visits them.
void collect_diagonal(const h_block* b, std::vector<const h_block*>& out) {
  if (b->type == element_t::dense) {
    if (b->nx != b->ny) throw std::invalid_argument("Hierarchical: non-square diagonal leaf");
    out.push_back(b);
    return;
  }
  if (b->type != element_t::hierarchical) throw std::invalid_argument("Hierarchical: non-dense diagonal leaf");
  for (int i = 0; i < b->py; ++i) collect_diagonal(b->child(i, i), out);
}

int64_t round32(int64_t x) { return (x + 31) / 32 * 32; }

}  // namespace

void Hierarchical::inverseLayout(std::vector<inv_block>& out) const {
  std::vector<const h_block*> diag;
  collect_diagonal(root, diag);
  out.clear();
  int64_t cursor = 0;
  for (const h_block* b : diag) {
    out.push_back(inv_block{b->off, b->ny, cursor, b->row0});
    cursor += 2 * round32((int64_t)b->ny * b->ny);
  }
}

namespace {

bool hodlr_shaped(const h_block* b) {
  if (b->type == element_t::dense) return b->nx == b->ny;
  if (b->type != element_t::hierarchical || b->py != 2 || b->px != 2) return false;
  const h_block *l = b->child(1, 0), *u = b->child(0, 1);
  if (l->type != element_t::low_rank || u->type != element_t::low_rank) return false;
  return hodlr_shaped(b->child(0, 0)) && hodlr_shaped(b->child(1, 1));
}

void collect_nodes(const h_block* b, int64_t row0, std::vector<Hierarchical::inv_node>& out, int64_t& cursor) {
  if (b->type != element_t::hierarchical) return;
  const h_block *a = b->child(0, 0), *c = b->child(1, 1);
  collect_nodes(a, row0, out, cursor);
  collect_nodes(c, row0 + a->ny, out, cursor);
  Hierarchical::inv_node n{};
  n.node = b;
  n.row0 = row0;
  n.na = a->ny;
  n.nb = c->ny;
  n.rl = b->child(1, 0)->rank;
  n.ru = b->child(0, 1)->rank;
  n.l_u = cursor;
  cursor += round32((int64_t)n.nb * n.rl);
  n.l_v = cursor;
  cursor += round32((int64_t)n.na * n.rl);
  n.u_u = cursor;
  cursor += round32((int64_t)n.na * n.ru);
  n.u_v = cursor;
  cursor += round32((int64_t)n.nb * n.ru);
  out.push_back(n);
}

}  // namespace

bool Hierarchical::hodlrShaped() const { return hodlr_shaped(root); }

void Hierarchical::inverseNodes(std::vector<inv_node>& out) const {
  out.clear();
  if (!hodlrShaped()) return;
  std::vector<inv_block> lay;
  inverseLayout(lay);
  int64_t cursor = 0;
  for (const inv_block& e : lay) cursor += 2 * round32((int64_t)e.nb * e.nb);
  collect_nodes(root, 0, out, cursor);
}

int64_t Hierarchical::inverseStorage() const {
  std::vector<inv_block> lay;
  inverseLayout(lay);
  int64_t s = 0;
  for (const inv_block& e : lay) s += 2 * round32((int64_t)e.nb * e.nb);
  std::vector<inv_node> nodes;
  inverseNodes(nodes);
  for (const inv_node& n : nodes)
    s += round32((int64_t)n.nb * n.rl) + round32((int64_t)n.na * n.rl) + round32((int64_t)n.na * n.ru) +
         round32((int64_t)n.nb * n.ru);
  return s;
}
