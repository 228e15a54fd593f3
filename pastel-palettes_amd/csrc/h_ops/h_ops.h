// H-ops: the leaf-level dense / low-rank operations an H-LU factorization and
// solve decompose into, each with its instruction words, operand data maps
// and FLOP count.
//
// Instruction word layout follows the decoder of the reference persistent
// kernel (src/kernel.cu:1384-1522; lengths also at src/kernel_cublas.cu:59,
// 70,95,105).  Every stream entry starts with [opcode, id]; execute entries
// continue with [operation, operands...]:
//   getrf     8  [ex, id, getrf, P(M), off, nx, ny, ld]
//   trsml    13  [ex, id, trsml, P(B), P(L), off_b, off_l, nx_b, ny_b, nx_l, ld_b, ld_l, b_T]
//   trsmr    13  [ex, id, trsmr, P(B), P(U), off_b, off_u, nx_b, ny_b, ny_u, ld_b, ld_u, b_T]
//   gemm     17  [ex, id, gemm|gemm_plus, P(M), P(A), P(B), off_m, off_a, off_b, m, n, k,
//                 ld_m, ld_a, ld_b, a_T, b_T]
//   gemm_3x  23  [ex, id, gemm_3x, P(M..C) x4, off x4, m, n, k, l, ld x4, a_T, b_T, c_T, control]
//   gemm_4x  29  [ex, id, gemm_4x, P(M..D) x5, off x5, m, n, k, l, o, ld x5, a_T..d_T, control, offset]
//   accum    21  [ex, id, accum, P(U1), P(VT1), P(U2), P(VT2), off x4, nx, ny, rank1, rank2,
//                 ld_u1, ld_vt1, ld_u2, ld_vt2, offset1, offset2]
//   trsmu    13  [ex, id, trsmu, P(B), P(U), off_b, off_u, nx_b, ny_b, ny_u, ld_b, ld_u, 0]
//                (addition: left solve with the non-unit upper factor, used by
//                 the backward substitution; the reference has no solve)
//   trmm     11  [ex, id, trmm, P(B), P(T), off_b, off_t, nx_b, ny_b, ld_b, ld_t]
//                (addition: B <- T B with a dense ny_b x ny_b T, used by the
//                 solve that applies explicit inverses of the leaf factors)
//   wait      3  [signal_wait, id, tick_hint]
//   finish    1  [finish]
// The numeric values of the opcodes and operations are not in the reference
// snapshot; the values below are this library's.
//
// Offsets are element offsets into the pointer-table slot P(.); matrices are
// row-major with row pitch ld; nx = columns, ny = rows.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../rts/ops.h"

enum opcode_t : int { finish = 0, execute = 1, signal_wait = 2 };

enum operation_t : int {
  nop = 0,
  getrf = 1,
  trsml = 2,
  trsmr = 3,
  gemm = 4,
  gemm_plus = 5,
  gemm_3x = 6,
  gemm_4x = 7,
  accum = 8,
  trsmu = 9,
  trmm = 10,
};

// Entry lengths in words.
constexpr int nop_l = 3, getrf_l = 8, trsml_l = 13, trsmr_l = 13, gemm_l = 17, gemm_plus_l = 17, gemm_3x_l = 23,
              gemm_4x_l = 29, accum_l = 21, trsmu_l = 13, trmm_l = 11, wait_l = 3, finish_l = 1;
constexpr int _MAX_INST_LENGTH = 32;

int operationLength(int op);

// Byte size used to lay out the virtual address space of the data maps.
constexpr int kMapElemBytes = 8;
// Data maps of different pointer-table slots never overlap: slot s occupies
// [(s + 1) << 44, (s + 2) << 44).
inline uintptr_t slot_base(int ptr_id) { return (uintptr_t)(ptr_id + 1) << 44; }

// Row-major view: element (i, j) at off + i * ld + j of slot `ptr`.
struct dview {
  int ptr = 0;
  int64_t off = 0;
  int ny = 0, nx = 0, ld = 0;
  dview sub(int r0, int c0, int h, int w) const {
    dview v = *this;
    v.off = off + (int64_t)r0 * ld + c0;
    v.ny = h;
    v.nx = w;
    return v;
  }
  DataMap2D map() const {
    return DataMap2D(reinterpret_cast<void*>(slot_base(ptr) + (uintptr_t)off * kMapElemBytes),
                     (size_t)ld * kMapElemBytes, (size_t)nx * kMapElemBytes, (size_t)ny);
  }
};

// Logical matrix = stored view (t == false) or its transpose (t == true).
struct mview {
  dview v;
  bool t = false;
  int rows() const { return t ? v.nx : v.ny; }
  int cols() const { return t ? v.ny : v.nx; }
  mview T() const { return mview{v, !t}; }
  // sub-block of the logical matrix
  mview sub(int r0, int c0, int h, int w) const {
    return t ? mview{v.sub(c0, r0, w, h), true} : mview{v.sub(r0, c0, h, w), false};
  }
};

struct h_op {
  operation_t op = nop;
  std::vector<int> words;  // full execute entry with id slot left 0
  Operation data;          // operands for dependency analysis
  long long flops = 0;
  int64_t scratch = 0;     // per-worker scratch elements this op needs
};

// Flattened h-op list in program order (the leaves of the reference's
// h_ops_tree, src/launcher.cu:139,146).
class h_ops_tree {
private:
  std::vector<h_op> ops;

public:
  int length() const { return (int)ops.size(); }
  const h_op& getOp(int i) const { return ops[i]; }
  h_op& getOp(int i) { return ops[i]; }
  void push(h_op&& o) { ops.push_back(std::move(o)); }
  // Append another tree's ops (used to merge independent matrices).
  void append(const h_ops_tree& other);
  long long getFlops() const;
  int64_t maxScratch() const;
};

// Temporary-buffer manager (src/launcher.cu:135,145): hands out zero-filled
// element ranges of the temp slot.  The runtime allocates `size()` elements
// and zero-fills them before every run.
class dev_temp {
private:
  int ptr_id;
  int64_t used;

public:
  explicit dev_temp(int id = 1) : ptr_id(id), used(0) {}
  int getPtrId() const { return ptr_id; }
  int64_t size() const { return used; }
  dview request(int ny, int nx);
};

// FLOP closed forms.
long long flops_getrf(long long nx, long long ny);
long long flops_trsml(long long nx_b, long long ny_b);
long long flops_trsmr(long long nx_b, long long ny_b);
long long flops_gemm(long long m, long long n, long long k);

// Dense-LU FLOPs n(n-1)/2 + n(n-1)(2n-1)/3 (reference h_ops::getFlops_GETRF,
// src/launcher.cu:174).
long long getFlops_GETRF(long long nx, long long ny);

// Leaf-op emitters (append one h_op to `tree`).
void emit_getrf(h_ops_tree* tree, const dview& m);
void emit_trsml(h_ops_tree* tree, const mview& b, const dview& l);
void emit_trsmr(h_ops_tree* tree, const mview& b, const dview& u);
void emit_trsmu(h_ops_tree* tree, const dview& b, const dview& u);
// B <- T B, T dense square with B's row count.
void emit_trmm(h_ops_tree* tree, const dview& b, const dview& t);
// M -= A B (plus == false) or M += A B (plus == true).
void emit_gemm(h_ops_tree* tree, mview m, mview a, mview b, bool plus);
// M -= A B C.
void emit_gemm3(h_ops_tree* tree, mview m, mview a, mview b, mview c);
// M -= A B C D.
void emit_gemm4(h_ops_tree* tree, mview m, mview a, mview b, mview c, mview d);
// U1 V1^T <- rank-k1 projection of U1 V1^T + U2 V2^T.
void emit_accum(h_ops_tree* tree, const dview& u1, const dview& v1, const dview& u2, const dview& v2);

// Scratch layout of accum for (nx, ny, k1, k2): Y at 0, V at offset1, Q at
// offset2, then small k1 x max(k1,k2) panels.
int64_t accum_small_elems(int k1, int k2);
void accum_scratch(int nx, int ny, int k1, int k2, int64_t* offset1, int64_t* offset2, int64_t* total);
int64_t tri_scratch_elems(int n);

// Checks the operand words of one execute entry (w[0..operationLength) with
// w[2] the operation) the way the device decoder will read them: every slot
// index in [0, n_slots), offsets, sizes and pitches non-negative, every pitch
// at least the stored row width, transpose flags 0 / 1, a gemm_3x / gemm_4x
// association order and accum's rank within the decoder's range, and the
// worker-scratch offsets consistent with the temporaries they hold.  Returns
// the per-worker scratch (elements) the device routine touches, or -1 with
// the reason in *why.
int64_t h_op_check_words(const int* w, int n_words, int n_slots, std::string* why);
