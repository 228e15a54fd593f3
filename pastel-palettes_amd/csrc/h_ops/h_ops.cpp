// Leaf-op emitters and the H-LU / H-solve op generators.
//
// The reference's task-tree generator (h_ops/*, src/launcher.cu:138-139) is
// absent from its snapshot; docs/README.md:23-31 defines the h-ops GETRF,
// TRSM, GEMM (A = A - B*C) and ACCM (A = A + B).  This file fixes the
// decomposition used here (right-looking block LU without pivoting):
//
//   getrf(H):  for i: getrf(H_ii); trsml(H_ii, H_ij) for j > i;
//              trsmr(H_ii, H_ji) for j > i; gemm(H_jk, H_ji, H_ik) for j,k > i
//   trsml(L, B), trsmr(U, B): recurse over the hierarchical factor, applying
//              a low-rank B = U_B V_B^T to U_B (left) or V_B (right)
//   gemm(M, A, B) with A, B low rank and M hierarchical or low rank:
//              T = V_A^T U_B (once), W = -V_B T^T (once), then M += U_A W^T
//              distributed over M's leaves: dense leaves by gemm_plus,
//              low-rank leaves by accum.
//   dense M with low-rank operands uses gemm_3x / gemm_4x directly.
//
// Ops are emitted in program order; the DAG is derived from their operand
// maps alone.
#include "h_ops.h"

#include <algorithm>
#include <stdexcept>
#include <string>

#include "../matrix/Hierarchical.h"

namespace {

int64_t align32(int64_t x) { return (x + 31) / 32 * 32; }

void require(bool ok, const char* what) {
  if (!ok) throw std::invalid_argument(std::string("h_ops: ") + what);
}

h_op make_op(operation_t op, std::initializer_list<int> args) {
  h_op o;
  o.op = op;
  o.words.reserve(3 + args.size());
  o.words.push_back(execute);
  o.words.push_back(0);
  o.words.push_back(op);
  for (int a : args) o.words.push_back(a);
  return o;
}

int off32(int64_t off) {
  if (off < 0 || off > 0x7fffffffLL) throw std::out_of_range("h_ops: element offset exceeds int32");
  return (int)off;
}

}  // namespace

int operationLength(int op) {
  switch (op) {
    case nop: return nop_l;
    case getrf: return getrf_l;
    case trsml: return trsml_l;
    case trsmr: return trsmr_l;
    case gemm: return gemm_l;
    case gemm_plus: return gemm_plus_l;
    case gemm_3x: return gemm_3x_l;
    case gemm_4x: return gemm_4x_l;
    case accum: return accum_l;
    case trsmu: return trsmu_l;
    case trmm: return trmm_l;
    default: return -1;
  }
}

// ------------------------------------------------------------------ FLOP model

long long flops_getrf(long long nx, long long ny) {
  long long f = 0;
  const long long mn = std::min(nx, ny);
  for (long long i = 0; i < mn; ++i) f += (ny - i - 1) + 2 * (ny - i - 1) * (nx - i - 1);
  return f;
}

long long flops_trsml(long long nx_b, long long ny_b) { return nx_b * ny_b * (ny_b - 1); }
long long flops_trsmr(long long nx_b, long long ny_b) { return ny_b * nx_b * nx_b; }
long long flops_gemm(long long m, long long n, long long k) { return 2 * m * n * k; }

long long getFlops_GETRF(long long nx, long long ny) {
  const long long n = std::min(nx, ny);
  return n * (n - 1) / 2 + n * (n - 1) * (2 * n - 1) / 3;
}

void h_ops_tree::append(const h_ops_tree& other) {
  for (const h_op& o : other.ops) ops.push_back(o);
}

long long h_ops_tree::getFlops() const {
  long long f = 0;
  for (const h_op& o : ops) f += o.flops;
  return f;
}

int64_t h_ops_tree::maxScratch() const {
  int64_t s = 0;
  for (const h_op& o : ops) s = std::max(s, o.scratch);
  return s;
}

dview dev_temp::request(int ny, int nx) {
  dview v;
  v.ptr = ptr_id;
  v.off = used;
  v.ny = ny;
  v.nx = nx;
  v.ld = nx;
  used = align32(used + (int64_t)ny * nx);
  return v;
}

// Scratch of one accum: Y (ny x k1), Vn (nx x k1), Q (ny x k1), then the
// small area: S (k1 x kmax), the Gram matrix / Cholesky factor (k1 x k1), the
// inverses of its 32 x 32 diagonal blocks, and tau / w of the Householder
// fallback (2 kmax).  Every part starts 32-element aligned.
int64_t accum_small_elems(int k1, int k2) {
  const int64_t kmax = std::max(k1, k2);
  return align32(k1 * kmax) + align32((int64_t)k1 * k1) + (int64_t)((k1 + 31) / 32) * 1024 + align32(2 * kmax);
}

void accum_scratch(int nx, int ny, int k1, int k2, int64_t* offset1, int64_t* offset2, int64_t* total) {
  *offset1 = align32((int64_t)ny * k1);
  *offset2 = align32(*offset1 + (int64_t)nx * k1);
  *total = align32(*offset2 + (int64_t)ny * k1) + accum_small_elems(k1, k2);
}

// ------------------------------------------------------------------ emitters

// Device getrf / trsm keep inverses of the 32 x 32 diagonal blocks of an
// n x n triangular factor in the worker scratch.
int64_t tri_scratch_elems(int n) { return (int64_t)((n + 31) / 32) * 32 * 32; }

int64_t h_op_check_words(const int* w, int n_words, int n_slots, std::string* why) {
  auto fail = [&](const char* what) -> int64_t {
    if (why) *why = what;
    return -1;
  };
  if (n_words < 3 || w[0] != execute) return fail("not an execute entry");
  const int op = w[2];
  if (operationLength(op) != n_words) return fail("entry length differs from its operation's");
  auto slot_ok = [&](int i) { return w[i] >= 0 && w[i] < n_slots; };
  auto nonneg = [&](int i0, int i1) {
    for (int i = i0; i <= i1; ++i)
      if (w[i] < 0) return false;
    return true;
  };
  auto flag_ok = [&](int i) { return w[i] == 0 || w[i] == 1; };
  // a stored rows x cols block with pitch ld needs ld >= cols
  auto pitch_ok = [](int ld, int cols) { return ld >= cols; };
  switch (op) {
    case nop:
      return 0;
    case getrf:  // [3] P, [4] off, [5] nx, [6] ny, [7] ld
      if (!slot_ok(3) || !nonneg(4, 7) || !pitch_ok(w[7], w[5])) return fail("bad getrf operands");
      return tri_scratch_elems(32);
    case trsml:
    case trsmr:
    case trsmu: {  // [3] P(B), [4] P(F), [5,6] off, [7] nx_b, [8] ny_b, [9] n_f, [10,11] ld, [12] b_T
      if (!slot_ok(3) || !slot_ok(4) || !nonneg(5, 11) || !flag_ok(12)) return fail("bad trsm operands");
      const int nx_b = w[7], ny_b = w[8], nf = w[9];
      const bool bt = w[12] != 0;
      if (op == trsmu && bt) return fail("trsmu has no transposed form");
      const int n = op == trsmr ? nx_b : ny_b;  // order of the triangular factor
      if (nf != n || !pitch_ok(w[11], n) || !pitch_ok(w[10], bt ? ny_b : nx_b)) return fail("bad trsm shape");
      return tri_scratch_elems(n);
    }
    case trmm:  // [3] P(B), [4] P(T), [5,6] off, [7] nx_b, [8] ny_b, [9] ld_b, [10] ld_t
      if (!slot_ok(3) || !slot_ok(4) || !nonneg(5, 10) || !pitch_ok(w[9], w[7]) || !pitch_ok(w[10], w[8]))
        return fail("bad trmm operands");
      return (int64_t)w[7] * w[8];
    case gemm:
    case gemm_plus: {  // [3..5] P, [6..8] off, [9..11] m n k, [12..14] ld, [15,16] a_T b_T
      if (!slot_ok(3) || !slot_ok(4) || !slot_ok(5) || !nonneg(6, 14) || !flag_ok(15) || !flag_ok(16))
        return fail("bad gemm operands");
      const int m = w[9], n = w[10], k = w[11];
      if (!pitch_ok(w[12], n) || !pitch_ok(w[13], w[15] ? m : k) || !pitch_ok(w[14], w[16] ? k : n))
        return fail("bad gemm pitch");
      return 0;
    }
    case gemm_3x: {  // [3..6] P, [7..10] off, [11..14] m n k l, [15..18] ld, [19..21] T, [22] control
      for (int i = 3; i <= 6; ++i)
        if (!slot_ok(i)) return fail("bad gemm_3x slot");
      if (!nonneg(7, 18) || !flag_ok(19) || !flag_ok(20) || !flag_ok(21) || !flag_ok(22))
        return fail("bad gemm_3x operands");
      const int64_t m = w[11], n = w[12], k = w[13], l = w[14];
      if (!pitch_ok(w[15], (int)n) || !pitch_ok(w[16], (int)(w[19] ? m : k)) ||
          !pitch_ok(w[17], (int)(w[20] ? k : l)) || !pitch_ok(w[18], (int)(w[21] ? l : n)))
        return fail("bad gemm_3x pitch");
      return w[22] ? m * l : k * n;
    }
    case gemm_4x: {  // [3..7] P, [8..12] off, [13..17] m n k l o, [18..22] ld, [23..26] T, [27] control, [28] offset
      for (int i = 3; i <= 7; ++i)
        if (!slot_ok(i)) return fail("bad gemm_4x slot");
      if (!nonneg(8, 22) || !flag_ok(23) || !flag_ok(24) || !flag_ok(25) || !flag_ok(26) || w[27] < 0 || w[27] > 4 ||
          w[28] < 0)
        return fail("bad gemm_4x operands");
      const int64_t m = w[13], n = w[14], k = w[15], l = w[16], o = w[17];
      if (!pitch_ok(w[18], (int)n) || !pitch_ok(w[19], (int)(w[23] ? m : k)) ||
          !pitch_ok(w[20], (int)(w[24] ? k : l)) || !pitch_ok(w[21], (int)(w[25] ? l : o)) ||
          !pitch_ok(w[22], (int)(w[26] ? o : n)))
        return fail("bad gemm_4x pitch");
      // first and second temporaries of the five association orders
      const int64_t t1[5] = {m * l, k * o, k * o, l * n, m * l};
      const int64_t t2[5] = {m * o, m * o, k * n, k * n, l * n};
      const int c = w[27];
      if (w[28] < t1[c]) return fail("gemm_4x temporaries overlap");
      return w[28] + t2[c];
    }
    case accum: {  // [3..6] P, [7..10] off, [11] nx, [12] ny, [13] k1, [14] k2, [15..18] ld, [19,20] offsets
      for (int i = 3; i <= 6; ++i)
        if (!slot_ok(i)) return fail("bad accum slot");
      if (!nonneg(7, 20)) return fail("bad accum operands");
      const int nx = w[11], ny = w[12], k1 = w[13], k2 = w[14];
      if (k1 > 64) return fail("accum rank above 64");
      if (!pitch_ok(w[15], k1) || !pitch_ok(w[16], k1) || !pitch_ok(w[17], k2) || !pitch_ok(w[18], k2))
        return fail("bad accum pitch");
      // Y [0, ny k1), Vn [offset1, + nx k1), Q [offset2, + ny k1), then the small area
      if (w[19] < (int64_t)ny * k1 || w[20] < w[19] + (int64_t)nx * k1) return fail("accum temporaries overlap");
      return align32(w[20] + (int64_t)ny * k1) + accum_small_elems(k1, k2);
    }
    default:
      return fail("unknown operation");
  }
}

void emit_getrf(h_ops_tree* tree, const dview& m) {
  h_op o = make_op(getrf, {m.ptr, off32(m.off), m.nx, m.ny, m.ld});
  o.data.addData(false, m.map());
  o.flops = flops_getrf(m.nx, m.ny);
  o.scratch = tri_scratch_elems(32);
  tree->push(std::move(o));
}

void emit_trsml(h_ops_tree* tree, const mview& b, const dview& l) {
  const int ny_b = b.rows(), nx_b = b.cols();
  require(l.ny == ny_b && l.nx == ny_b, "trsml: L must be square with B's row count");
  h_op o = make_op(trsml, {b.v.ptr, l.ptr, off32(b.v.off), off32(l.off), nx_b, ny_b, l.nx, b.v.ld, l.ld, b.t ? 1 : 0});
  o.data.addData(false, b.v.map());
  o.data.addData(true, l.map());
  o.flops = flops_trsml(nx_b, ny_b);
  o.scratch = tri_scratch_elems(ny_b);
  tree->push(std::move(o));
}

void emit_trsmr(h_ops_tree* tree, const mview& b, const dview& u) {
  const int ny_b = b.rows(), nx_b = b.cols();
  require(u.ny == nx_b && u.nx == nx_b, "trsmr: U must be square with B's column count");
  h_op o = make_op(trsmr, {b.v.ptr, u.ptr, off32(b.v.off), off32(u.off), nx_b, ny_b, u.ny, b.v.ld, u.ld, b.t ? 1 : 0});
  o.data.addData(false, b.v.map());
  o.data.addData(true, u.map());
  o.flops = flops_trsmr(nx_b, ny_b);
  o.scratch = tri_scratch_elems(nx_b);
  tree->push(std::move(o));
}

void emit_trsmu(h_ops_tree* tree, const dview& b, const dview& u) {
  require(u.ny == b.ny && u.nx == b.ny, "trsmu: U must be square with B's row count");
  h_op o = make_op(trsmu, {b.ptr, u.ptr, off32(b.off), off32(u.off), b.nx, b.ny, u.ny, b.ld, u.ld, 0});
  o.data.addData(false, b.map());
  o.data.addData(true, u.map());
  o.flops = (long long)b.nx * b.ny * b.ny;
  o.scratch = tri_scratch_elems(b.ny);
  tree->push(std::move(o));
}

// The product goes through a copy of B in the worker scratch.
void emit_trmm(h_ops_tree* tree, const dview& b, const dview& t) {
  require(t.ny == b.ny && t.nx == b.ny, "trmm: T must be square with B's row count");
  h_op o = make_op(trmm, {b.ptr, t.ptr, off32(b.off), off32(t.off), b.nx, b.ny, b.ld, t.ld});
  o.data.addData(false, b.map());
  o.data.addData(true, t.map());
  o.flops = flops_gemm(b.ny, b.nx, b.ny);
  o.scratch = (int64_t)b.ny * b.nx;
  tree->push(std::move(o));
}

void emit_gemm(h_ops_tree* tree, mview m, mview a, mview b, bool plus) {
  if (m.t) {  // M^T -/+= A B  <=>  M -/+= B^T A^T
    mview a2 = b.T(), b2 = a.T();
    m = m.T();
    a = a2;
    b = b2;
  }
  const int M = m.rows(), N = m.cols(), K = a.cols();
  require(a.rows() == M && b.rows() == K && b.cols() == N, "gemm: shape mismatch");
  h_op o = make_op(plus ? gemm_plus : gemm,
                   {m.v.ptr, a.v.ptr, b.v.ptr, off32(m.v.off), off32(a.v.off), off32(b.v.off), M, N, K, m.v.ld, a.v.ld,
                    b.v.ld, a.t ? 1 : 0, b.t ? 1 : 0});
  o.data.addData(false, m.v.map());
  o.data.addData(true, a.v.map());
  o.data.addData(true, b.v.map());
  o.flops = flops_gemm(M, N, K);
  tree->push(std::move(o));
}

void emit_gemm3(h_ops_tree* tree, mview m, mview a, mview b, mview c) {
  if (m.t) {
    mview a2 = c.T(), b2 = b.T(), c2 = a.T();
    m = m.T();
    a = a2;
    b = b2;
    c = c2;
  }
  const long long M = m.rows(), N = m.cols(), K = a.cols(), L = b.cols();
  require(a.rows() == M && b.rows() == K && c.rows() == L && c.cols() == N, "gemm_3x: shape mismatch");
  const long long f_ab_c = 2 * M * K * L + 2 * M * L * N;  // (A B) C, control 1
  const long long f_a_bc = 2 * K * N * L + 2 * M * N * K;  // A (B C), control 0
  const int control = f_ab_c <= f_a_bc ? 1 : 0;
  h_op o = make_op(gemm_3x, {m.v.ptr, a.v.ptr, b.v.ptr, c.v.ptr, off32(m.v.off), off32(a.v.off), off32(b.v.off),
                             off32(c.v.off), (int)M, (int)N, (int)K, (int)L, m.v.ld, a.v.ld, b.v.ld, c.v.ld,
                             a.t ? 1 : 0, b.t ? 1 : 0, c.t ? 1 : 0, control});
  o.data.addData(false, m.v.map());
  o.data.addData(true, a.v.map());
  o.data.addData(true, b.v.map());
  o.data.addData(true, c.v.map());
  o.flops = std::min(f_ab_c, f_a_bc);
  o.scratch = control ? M * L : K * N;
  tree->push(std::move(o));
}

void emit_gemm4(h_ops_tree* tree, mview m, mview a, mview b, mview c, mview d) {
  if (m.t) {
    mview a2 = d.T(), b2 = c.T(), c2 = b.T(), d2 = a.T();
    m = m.T();
    a = a2;
    b = b2;
    c = c2;
    d = d2;
  }
  const long long M = m.rows(), N = m.cols(), K = a.cols(), L = b.cols(), O = c.cols();
  require(a.rows() == M && b.rows() == K && c.rows() == L && d.rows() == O && d.cols() == N, "gemm_4x: shape mismatch");
  // association orders of src/kernel.cu:914-953: FLOPs, first temp, second temp
  const long long f[5] = {2 * (M * K * L + M * L * O + M * O * N), 2 * (K * L * O + M * K * O + M * O * N),
                          2 * (K * L * O + K * O * N + M * K * N), 2 * (L * O * N + K * L * N + M * K * N),
                          2 * (M * K * L + L * O * N + M * L * N)};
  const long long t1[5] = {M * L, K * O, K * O, L * N, M * L};
  const long long t2[5] = {M * O, M * O, K * N, K * N, L * N};
  int control = 0;
  for (int i = 1; i < 5; ++i)
    if (f[i] < f[control]) control = i;
  const int64_t offset = align32(t1[control]);
  h_op o = make_op(gemm_4x, {m.v.ptr, a.v.ptr, b.v.ptr, c.v.ptr, d.v.ptr, off32(m.v.off), off32(a.v.off),
                             off32(b.v.off), off32(c.v.off), off32(d.v.off), (int)M, (int)N, (int)K, (int)L, (int)O,
                             m.v.ld, a.v.ld, b.v.ld, c.v.ld, d.v.ld, a.t ? 1 : 0, b.t ? 1 : 0, c.t ? 1 : 0,
                             d.t ? 1 : 0, control, off32(offset)});
  o.data.addData(false, m.v.map());
  o.data.addData(true, a.v.map());
  o.data.addData(true, b.v.map());
  o.data.addData(true, c.v.map());
  o.data.addData(true, d.v.map());
  o.flops = f[control];
  o.scratch = offset + t2[control];
  tree->push(std::move(o));
}

void emit_accum(h_ops_tree* tree, const dview& u1, const dview& v1, const dview& u2, const dview& v2) {
  const int ny = u1.ny, nx = v1.ny, k1 = u1.nx, k2 = u2.nx;
  require(v1.nx == k1 && u2.ny == ny && v2.ny == nx && v2.nx == k2, "accum: shape mismatch");
  int64_t o1, o2, total;
  accum_scratch(nx, ny, k1, k2, &o1, &o2, &total);
  h_op o = make_op(accum, {u1.ptr, v1.ptr, u2.ptr, v2.ptr, off32(u1.off), off32(v1.off), off32(u2.off),
                           off32(v2.off), nx, ny, k1, k2, u1.ld, v1.ld, u2.ld, v2.ld, off32(o1), off32(o2)});
  o.data.addData(false, u1.map());
  o.data.addData(false, v1.map());
  o.data.addData(true, u2.map());
  o.data.addData(true, v2.map());
  const long long NX = nx, NY = ny, K1 = k1, K2 = k2;
  o.flops = 2 * K1 * K1 * NX + 2 * NY * K1 * K1 + 2 * K1 * K2 * NX + 2 * NY * K1 * K2  // sketch
            + 4 * NY * K1 * K1                                                        // QR + explicit Q
            + 2 * K1 * K1 * NY + 2 * NX * K1 * K1 + 2 * K1 * K2 * NY + 2 * NX * K1 * K2;  // projection
  o.scratch = total;
  tree->push(std::move(o));
}

// ------------------------------------------------------------------ generator

namespace {

struct operand {
  element_t type = element_t::empty;
  int rows = 0, cols = 0;
  mview d;        // dense
  mview u, v;     // low rank: logical = u v^T
  const h_block* h = nullptr;
};

operand of_block(const h_block* b, int ptr) {
  operand o;
  o.type = b->type;
  o.rows = b->ny;
  o.cols = b->nx;
  if (b->type == element_t::dense) {
    o.d = mview{dview{ptr, b->off, b->ny, b->nx, b->nx}, false};
  } else if (b->type == element_t::low_rank) {
    o.u = mview{dview{ptr, b->off_u, b->ny, b->rank, b->rank}, false};
    o.v = mview{dview{ptr, b->off_v, b->nx, b->rank, b->rank}, false};
  } else if (b->type == element_t::hierarchical) {
    o.h = b;
  } else {
    throw std::invalid_argument("h_ops: empty block");
  }
  return o;
}

operand of_dense(const mview& m) {
  operand o;
  o.type = element_t::dense;
  o.rows = m.rows();
  o.cols = m.cols();
  o.d = m;
  return o;
}

// Row / column cut points of a hierarchical operand.
std::vector<int> prefix(const std::vector<int>& sizes) {
  std::vector<int> p(1, 0);
  for (int s : sizes) p.push_back(p.back() + s);
  return p;
}

class op_builder {
public:
  h_ops_tree* tree;
  dev_temp* tmp;
  int ptr;  // matrix slot

  op_builder(h_ops_tree* t, dev_temp* d, int p) : tree(t), tmp(d), ptr(p) {}

  operand child(const operand& o, int i, int j) const { return of_block(o.h->child(i, j), ptr); }

  // sub-block of a dense or low-rank operand
  static operand cut(const operand& o, int r0, int c0, int h, int w) {
    operand s;
    s.type = o.type;
    s.rows = h;
    s.cols = w;
    if (o.type == element_t::dense) {
      s.d = o.d.sub(r0, c0, h, w);
    } else if (o.type == element_t::low_rank) {
      s.u = o.u.sub(r0, 0, h, o.u.cols());
      s.v = o.v.sub(c0, 0, w, o.v.cols());
    } else {
      throw std::invalid_argument("h_ops: cannot cut a hierarchical operand");
    }
    return s;
  }

  // ---- LU
  void getrf(const operand& a) {
    if (a.type == element_t::dense) {
      require(!a.d.t, "getrf: transposed block");
      emit_getrf(tree, a.d.v);
      return;
    }
    require(a.type == element_t::hierarchical, "getrf: diagonal block must be dense or hierarchical");
    const h_block* h = a.h;
    require(h->py == h->px && h->row_sizes == h->col_sizes, "getrf: diagonal block must have a square partition");
    const int p = h->py;
    for (int i = 0; i < p; ++i) {
      getrf(child(a, i, i));
      for (int j = i + 1; j < p; ++j) trsml(child(a, i, i), child(a, i, j));
      for (int j = i + 1; j < p; ++j) trsmr(child(a, i, i), child(a, j, i));
      for (int j = i + 1; j < p; ++j)
        for (int k = i + 1; k < p; ++k) gemm(child(a, j, k), child(a, j, i), child(a, i, k));
    }
  }

  // B <- L^{-1} B with the unit-lower part of L
  void trsml(const operand& l, const operand& b) {
    if (b.type == element_t::low_rank) {
      trsml(l, of_dense(b.u));
      return;
    }
    if (l.type == element_t::dense) {
      if (b.type == element_t::dense) {
        require(!l.d.t, "trsml: transposed L");
        emit_trsml(tree, b.d, l.d.v);
        return;
      }
      // dense L against hierarchical B: cut L along B's rows
      const h_block* hb = b.h;
      const std::vector<int> rp = prefix(hb->row_sizes);
      for (int i = 0; i < hb->py; ++i) {
        const operand lii = cut(l, rp[i], rp[i], hb->row_sizes[i], hb->row_sizes[i]);
        for (int j = 0; j < hb->px; ++j) trsml(lii, child(b, i, j));
        for (int k = i + 1; k < hb->py; ++k) {
          const operand lki = cut(l, rp[k], rp[i], hb->row_sizes[k], hb->row_sizes[i]);
          for (int j = 0; j < hb->px; ++j) gemm(child(b, k, j), lki, child(b, i, j));
        }
      }
      return;
    }
    const h_block* hl = l.h;
    const std::vector<int> rp = prefix(hl->row_sizes);
    if (b.type == element_t::dense) {
      for (int i = 0; i < hl->py; ++i) {
        const operand bi = cut(b, rp[i], 0, hl->row_sizes[i], b.cols);
        trsml(child(l, i, i), bi);
        for (int k = i + 1; k < hl->py; ++k)
          gemm(cut(b, rp[k], 0, hl->row_sizes[k], b.cols), child(l, k, i), bi);
      }
      return;
    }
    const h_block* hb = b.h;
    require(hb->row_sizes == hl->row_sizes, "trsml: partition mismatch");
    for (int i = 0; i < hl->py; ++i) {
      for (int j = 0; j < hb->px; ++j) trsml(child(l, i, i), child(b, i, j));
      for (int k = i + 1; k < hl->py; ++k)
        for (int j = 0; j < hb->px; ++j) gemm(child(b, k, j), child(l, k, i), child(b, i, j));
    }
  }

  // B <- B U^{-1} with the upper part of U
  void trsmr(const operand& u, const operand& b) {
    if (b.type == element_t::low_rank) {
      trsmr(u, of_dense(b.v.T()));  // V^T <- V^T U^{-1}
      return;
    }
    if (u.type == element_t::dense) {
      if (b.type == element_t::dense) {
        require(!u.d.t, "trsmr: transposed U");
        emit_trsmr(tree, b.d, u.d.v);
        return;
      }
      const h_block* hb = b.h;
      const std::vector<int> cp = prefix(hb->col_sizes);
      for (int j = 0; j < hb->px; ++j) {
        const operand ujj = cut(u, cp[j], cp[j], hb->col_sizes[j], hb->col_sizes[j]);
        for (int i = 0; i < hb->py; ++i) trsmr(ujj, child(b, i, j));
        for (int k = j + 1; k < hb->px; ++k) {
          const operand ujk = cut(u, cp[j], cp[k], hb->col_sizes[j], hb->col_sizes[k]);
          for (int i = 0; i < hb->py; ++i) gemm(child(b, i, k), child(b, i, j), ujk);
        }
      }
      return;
    }
    const h_block* hu = u.h;
    const std::vector<int> cp = prefix(hu->col_sizes);
    if (b.type == element_t::dense) {
      for (int j = 0; j < hu->px; ++j) {
        const operand bj = cut(b, 0, cp[j], b.rows, hu->col_sizes[j]);
        trsmr(child(u, j, j), bj);
        for (int k = j + 1; k < hu->px; ++k)
          gemm(cut(b, 0, cp[k], b.rows, hu->col_sizes[k]), bj, child(u, j, k));
      }
      return;
    }
    const h_block* hb = b.h;
    require(hb->col_sizes == hu->col_sizes, "trsmr: partition mismatch");
    for (int j = 0; j < hu->px; ++j) {
      for (int i = 0; i < hb->py; ++i) trsmr(child(u, j, j), child(b, i, j));
      for (int k = j + 1; k < hu->px; ++k)
        for (int i = 0; i < hb->py; ++i) gemm(child(b, i, k), child(b, i, j), child(u, j, k));
    }
  }

  // M += U W^T, U rows x r, W cols x r
  void add_lr(const operand& m, const mview& u, const mview& w) {
    if (m.type == element_t::dense) {
      emit_gemm(tree, m.d, u, w.T(), true);
    } else if (m.type == element_t::low_rank) {
      require(!m.u.t && !m.v.t && !u.t && !w.t, "accum: transposed factor");
      emit_accum(tree, m.u.v, m.v.v, u.v, w.v);
    } else {
      const h_block* h = m.h;
      const std::vector<int> rp = prefix(h->row_sizes), cp = prefix(h->col_sizes);
      for (int i = 0; i < h->py; ++i)
        for (int j = 0; j < h->px; ++j)
          add_lr(child(m, i, j), u.sub(rp[i], 0, h->row_sizes[i], u.cols()),
                 w.sub(cp[j], 0, h->col_sizes[j], w.cols()));
    }
  }

  // M -= A B
  void gemm(const operand& m, const operand& a, const operand& b) {
    require(m.rows == a.rows && m.cols == b.cols && a.cols == b.rows, "gemm: operand shapes");
    const bool a_lr = a.type == element_t::low_rank, b_lr = b.type == element_t::low_rank;
    const bool a_d = a.type == element_t::dense, b_d = b.type == element_t::dense;

    if (m.type == element_t::dense) {
      if (a_d && b_d) return emit_gemm(tree, m.d, a.d, b.d, false);
      if (a_lr && b_d) return emit_gemm3(tree, m.d, a.u, a.v.T(), b.d);
      if (a_d && b_lr) return emit_gemm3(tree, m.d, a.d, b.u, b.v.T());
      if (a_lr && b_lr) return emit_gemm4(tree, m.d, a.u, a.v.T(), b.u, b.v.T());
    }

    if (m.type != element_t::dense && (a_lr || b_lr) && (a_lr || a_d) && (b_lr || b_d)) {
      // low-rank product P = A B = -X Y^T
      mview x, y;
      if (a_lr && b_lr) {
        const dview t = tmp->request(a.u.cols(), b.u.cols());
        emit_gemm(tree, mview{t, false}, a.v.T(), b.u, true);           // T = V_A^T U_B
        const dview w = tmp->request(b.cols, a.u.cols());
        emit_gemm(tree, mview{w, false}, b.v, mview{t, true}, false);   // W = -V_B T^T
        x = a.u;
        y = mview{w, false};
      } else if (a_d) {
        const dview w = tmp->request(a.rows, b.u.cols());
        emit_gemm(tree, mview{w, false}, a.d, b.u, false);               // W = -A U_B
        x = mview{w, false};
        y = b.v;
      } else {
        const dview w = tmp->request(b.cols, a.u.cols());
        emit_gemm(tree, mview{w, false}, b.d.T(), a.v, false);           // W = -B^T V_A
        x = a.u;
        y = mview{w, false};
      }
      add_lr(m, x, y);  // M -= P  <=>  M += X Y^T
      return;
    }

    // structural recursion: cut operands along the hierarchical partitions
    std::vector<int> rows, cols, inner;
    if (m.type == element_t::hierarchical) {
      rows = m.h->row_sizes;
      cols = m.h->col_sizes;
    } else {
      rows = a.type == element_t::hierarchical ? a.h->row_sizes : std::vector<int>{m.rows};
      cols = b.type == element_t::hierarchical ? b.h->col_sizes : std::vector<int>{m.cols};
    }
    if (a.type == element_t::hierarchical) inner = a.h->col_sizes;
    else if (b.type == element_t::hierarchical) inner = b.h->row_sizes;
    else inner = {a.cols};
    require(m.type != element_t::low_rank, "gemm: unsupported low-rank target with hierarchical operands");

    const std::vector<int> rp = prefix(rows), cp = prefix(cols), kp = prefix(inner);
    auto piece = [&](const operand& o, const std::vector<int>& rs, const std::vector<int>& cs, int i, int j,
                     const std::vector<int>& rpp, const std::vector<int>& cpp) -> operand {
      if (o.type == element_t::hierarchical) {
        require(o.h->row_sizes == rs && o.h->col_sizes == cs, "gemm: partition mismatch");
        return child(o, i, j);
      }
      return cut(o, rpp[i], cpp[j], rs[i], cs[j]);
    };
    for (size_t i = 0; i < rows.size(); ++i)
      for (size_t j = 0; j < cols.size(); ++j)
        for (size_t k = 0; k < inner.size(); ++k)
          gemm(piece(m, rows, cols, (int)i, (int)j, rp, cp), piece(a, rows, inner, (int)i, (int)k, rp, kp),
               piece(b, inner, cols, (int)k, (int)j, kp, cp));
  }

  // ---- solve: b <- L^{-1} b (unit lower), then b <- U^{-1} b.  With an
  // inverse table, a dense diagonal leaf is applied as one product with the
  // explicit inverse of its factor instead of a substitution.
  const std::vector<Hierarchical::inv_block>* inv = nullptr;
  int64_t inv_base = 0;

  // L^{-1} (upper == false) or U^{-1} of the dense diagonal leaf at `leaf_off`
  dview inverse_of(const dview& leaf, bool upper) const {
    for (const auto& e : *inv)
      if (e.leaf_off == leaf.off) {
        const int64_t off = inv_base + e.inv_off + (upper ? align32((int64_t)e.nb * e.nb) : 0);
        return dview{ptr, off, e.nb, e.nb, e.nb};
      }
    throw std::invalid_argument("solve: leaf without an inverse");
  }

  void forward(const operand& l, const mview& b) {
    if (l.type == element_t::dense) {
      if (inv) {
        require(!b.t, "trmm: transposed rhs");
        emit_trmm(tree, b.v, inverse_of(l.d.v, false));
      } else {
        emit_trsml(tree, b, l.d.v);
      }
      return;
    }
    const h_block* h = l.h;
    const std::vector<int> rp = prefix(h->row_sizes);
    for (int i = 0; i < h->py; ++i) {
      const mview bi = b.sub(rp[i], 0, h->row_sizes[i], b.cols());
      forward(child(l, i, i), bi);
      for (int k = i + 1; k < h->py; ++k)
        gemm(of_dense(b.sub(rp[k], 0, h->row_sizes[k], b.cols())), child(l, k, i), of_dense(bi));
    }
  }

  void backward(const operand& u, const mview& b) {
    if (u.type == element_t::dense) {
      require(!b.t, "trsmu: transposed rhs");
      if (inv) emit_trmm(tree, b.v, inverse_of(u.d.v, true));
      else emit_trsmu(tree, b.v, u.d.v);
      return;
    }
    const h_block* h = u.h;
    const std::vector<int> rp = prefix(h->row_sizes);
    for (int i = h->py - 1; i >= 0; --i) {
      const mview bi = b.sub(rp[i], 0, h->row_sizes[i], b.cols());
      backward(child(u, i, i), bi);
      for (int k = i - 1; k >= 0; --k)
        gemm(of_dense(b.sub(rp[k], 0, h->row_sizes[k], b.cols())), child(u, k, i), of_dense(bi));
    }
  }
};

}  // namespace

h_ops_tree* Hierarchical::generateOps_GETRF(const h_index* self, dev_temp* tmp, int slot) const {
  if (self == nullptr || tmp == nullptr) throw std::invalid_argument("generateOps_GETRF: null argument");
  const int ms = slot >= 0 ? slot : ptr_id;
  h_ops_tree* tree = new h_ops_tree();
  try {
    op_builder bld(tree, tmp, ms);
    bld.getrf(of_block(self->node, ms));
  } catch (...) {
    delete tree;
    throw;
  }
  return tree;
}

h_ops_tree* Hierarchical::generateOps_Solve(const h_index* self, int rhs_ptr_id, int64_t rhs_off, int nrhs, int rhs_ld,
                                            int rhs_split, int slot, int64_t inv_base) const {
  const int ms = slot >= 0 ? slot : ptr_id;
  if (self == nullptr) throw std::invalid_argument("generateOps_Solve: null index");
  if (nrhs <= 0 || rhs_ld < nrhs || rhs_split <= 0) throw std::invalid_argument("generateOps_Solve: bad rhs shape");
  if (inv_base >= 0 && self->node != root) throw std::invalid_argument("generateOps_Solve: inverses need the root");
  h_ops_tree* tree = new h_ops_tree();
  try {
    dev_temp unused(-1);
    op_builder bld(tree, &unused, ms);
    std::vector<inv_block> lay;
    if (inv_base >= 0) {
      inverseLayout(lay);
      bld.inv = &lay;
      bld.inv_base = inv_base;
    }
    const operand root = of_block(self->node, ms);
    const int groups = std::min(rhs_split, nrhs);
    int c0 = 0;
    for (int g = 0; g < groups; ++g) {
      const int w = nrhs / groups + (g < nrhs % groups ? 1 : 0);
      const mview b{dview{rhs_ptr_id, rhs_off + c0, self->node->ny, w, rhs_ld}, false};
      bld.forward(root, b);
      bld.backward(root, b);
      c0 += w;
    }
  } catch (...) {
    delete tree;
    throw;
  }
  return tree;
}

h_ops_tree* Hierarchical::generateOps_Invert(const h_index* self, int64_t inv_base, int slot) const {
  if (self == nullptr || self->node != root) throw std::invalid_argument("generateOps_Invert: needs the root index");
  if (inv_base < 0) throw std::invalid_argument("generateOps_Invert: negative base");
  const int ms = slot >= 0 ? slot : ptr_id;
  std::vector<inv_block> lay;
  inverseLayout(lay);
  h_ops_tree* tree = new h_ops_tree();
  try {
    // column slices of kInvSlice: L^{-1}'s slice at s0 is zero above row s0
    // and U^{-1}'s below row s0 + w, so each slice is a narrow solve with the
    // trailing (leading) part of the factor only, and the slices run in
    // parallel
    constexpr int kInvSlice = 64;
    for (const inv_block& e : lay) {
      const dview leaf{ms, e.leaf_off, e.nb, e.nb, e.nb};
      const dview linv{ms, inv_base + e.inv_off, e.nb, e.nb, e.nb};
      const dview uinv{ms, inv_base + e.inv_off + align32((int64_t)e.nb * e.nb), e.nb, e.nb, e.nb};
      for (int s0 = 0; s0 < e.nb; s0 += kInvSlice) {
        const int w = std::min(kInvSlice, e.nb - s0), hi = s0 + w;
        emit_trsml(tree, mview{linv.sub(s0, s0, e.nb - s0, w), false}, leaf.sub(s0, s0, e.nb - s0, e.nb - s0));
        emit_trsmu(tree, uinv.sub(0, s0, hi, w), leaf.sub(0, 0, hi, hi));
      }
    }
  } catch (...) {
    delete tree;
    throw;
  }
  return tree;
}

namespace {

// Emits R += op(Inv(X)) P for the triangular factor inverse of subtree X
// (lower: L^{-1}, upper: U^{-1}; op = transpose when `trans`), P and R with
// X's rows.  Diagonal leaves use the explicit leaf inverses; every internal
// node y inside X contributes t = F_in^T P[src]; R[dst] -= F_out t, split by
// the diagonal leaves of dst so that the row blocks proceed independently.
class inverse_emitter {
public:
  h_ops_tree* tree;
  dev_temp* tmp;
  int ms;
  int64_t base;
  std::vector<Hierarchical::inv_block> leaves;
  std::vector<Hierarchical::inv_node> nodes;

  const Hierarchical::inv_block& leaf_of(const h_block* b) const {
    for (const auto& e : leaves)
      if (e.leaf_off == b->off) return e;
    throw std::invalid_argument("inverse: unknown leaf");
  }
  const Hierarchical::inv_node& node_of(const h_block* b) const {
    for (const auto& n : nodes)
      if (n.node == b) return n;
    throw std::invalid_argument("inverse: unknown node");
  }
  dview leaf_inverse(const h_block* b, bool upper) const {
    const auto& e = leaf_of(b);
    return dview{ms, base + e.inv_off + (upper ? align32((int64_t)e.nb * e.nb) : 0), e.nb, e.nb, e.nb};
  }
  static void diagonal(const h_block* x, std::vector<const h_block*>& out) {
    if (x->type == element_t::dense) {
      out.push_back(x);
      return;
    }
    diagonal(x->child(0, 0), out);
    diagonal(x->child(1, 1), out);
  }
  static void internal(const h_block* x, std::vector<const h_block*>& out) {
    if (x->type != element_t::hierarchical) return;
    internal(x->child(0, 0), out);
    internal(x->child(1, 1), out);
    out.push_back(x);
  }

  void apply(const h_block* x, bool upper, bool trans, const dview& p, const dview& r) {
    const int cols = p.nx;
    std::vector<const h_block*> diag;
    diagonal(x, diag);
    for (const h_block* e : diag) {
      const int o = e->row0 - x->row0;
      emit_gemm(tree, mview{r.sub(o, 0, e->ny, cols), false}, mview{leaf_inverse(e, upper), trans},
                mview{p.sub(o, 0, e->ny, cols), false}, true);
    }
    std::vector<const h_block*> inner;
    internal(x, inner);
    for (const h_block* y : inner) {
      const auto& n = node_of(y);
      const int oa = (int)(n.row0 - x->row0), ob = oa + n.na;
      const int rank = upper ? n.ru : n.rl;
      const bool src_is_a = upper == trans;  // lower: A -> B; upper: B -> A; transposed swaps
      const int src_o = src_is_a ? oa : ob, src_n = src_is_a ? n.na : n.nb;
      const int dst_o = src_is_a ? ob : oa, dst_n = src_is_a ? n.nb : n.na;
      // the factor with the source rows and the one with the destination rows
      int64_t f_a, f_b;  // factor with A's rows, factor with B's rows
      if (upper) {
        f_a = n.u_u;
        f_b = n.u_v;
      } else {
        f_a = n.l_v;
        f_b = n.l_u;
      }
      const dview fin{ms, base + (src_is_a ? f_a : f_b), src_n, rank, rank};
      const dview fout{ms, base + (src_is_a ? f_b : f_a), dst_n, rank, rank};
      const dview t = tmp->request(rank, cols);
      emit_gemm(tree, mview{t, false}, mview{fin, true}, mview{p.sub(src_o, 0, src_n, cols), false}, true);
      std::vector<const h_block*> dleaves;
      diagonal(src_is_a ? y->child(1, 1) : y->child(0, 0), dleaves);
      for (const h_block* e : dleaves) {
        const int o = e->row0 - x->row0;
        emit_gemm(tree, mview{r.sub(o, 0, e->ny, cols), false}, mview{fout.sub(o - dst_o, 0, e->ny, rank), false},
                  mview{t, false}, false);
      }
    }
  }
};

}  // namespace

h_ops_tree* Hierarchical::generateOps_InvertH(const h_index* self, int64_t inv_base, dev_temp* tmp, int slot) const {
  if (tmp == nullptr) throw std::invalid_argument("generateOps_InvertH: null temp");
  h_ops_tree* tree = generateOps_Invert(self, inv_base, slot);
  try {
    inverse_emitter em;
    em.tree = tree;
    em.tmp = tmp;
    em.ms = slot >= 0 ? slot : ptr_id;
    em.base = inv_base;
    inverseLayout(em.leaves);
    inverseNodes(em.nodes);
    if (em.nodes.empty() && root->type == element_t::hierarchical)
      throw std::invalid_argument("generateOps_InvertH: the tree is not HODLR-shaped");
    for (const inv_node& n : em.nodes) {
      const h_block *a = n.node->child(0, 0), *b = n.node->child(1, 1);
      const h_block *lo = n.node->child(1, 0), *up = n.node->child(0, 1);
      const dview ul{em.ms, lo->off_u, n.nb, n.rl, n.rl}, vl{em.ms, lo->off_v, n.na, n.rl, n.rl};
      const dview uc{em.ms, up->off_u, n.na, n.ru, n.ru}, vc{em.ms, up->off_v, n.nb, n.ru, n.ru};
      em.apply(b, false, false, ul, dview{em.ms, inv_base + n.l_u, n.nb, n.rl, n.rl});
      em.apply(a, false, true, vl, dview{em.ms, inv_base + n.l_v, n.na, n.rl, n.rl});
      em.apply(a, true, false, uc, dview{em.ms, inv_base + n.u_u, n.na, n.ru, n.ru});
      em.apply(b, true, true, vc, dview{em.ms, inv_base + n.u_v, n.nb, n.ru, n.ru});
    }
  } catch (...) {
    delete tree;
    throw;
  }
  return tree;
}
