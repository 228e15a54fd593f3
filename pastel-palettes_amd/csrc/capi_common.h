// Internal definitions shared by the host and device halves of the C ABI.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "h_ops/h_ops.h"
#include "instructions/instructions_manager.h"
#include "matrix/Hierarchical.h"
#include "rts/h_ops_dag.h"
#include "rts/scheduler.h"

struct pp_program;
struct pp_hmatrix;

namespace ppc {
int set_error(int code, const std::string& msg);
// Factorization program followed by the inverse construction: the full
// hierarchical inverses of HODLR-shaped members when `hier`, else the leaf
// inverses only.
int factor_inverse_program(const pp_hmatrix* const* hs, int count, int workers, bool hier, pp_program** out);
// Convert the in-flight exception to an error code (call inside catch).
int from_exception();
}  // namespace ppc

struct pp_hmatrix {
  std::unique_ptr<Hierarchical> h;
  std::vector<const h_block*> leaves;
};

// The leaf-inverse area of a matrix slot starts after the arena, 64-element
// aligned.
inline int64_t inverse_base(const Hierarchical* h) { return (h->getStorage() + 63) / 64 * 64; }

// What a decoded program carries in place of its DAG, schedule and encoder
// (the accessors read it when `dag` is null).
struct pp_program_image {
  int32_t n_insts = 0;
  int64_t n_edges = 0, flops = 0, comm_waits = 0;
  double dag_density = 0;
  std::vector<std::vector<int32_t>> streams;  // per worker
  std::vector<std::vector<int32_t>> preds;    // per instruction: (inst, mask) pairs
  struct entry {
    int32_t inst, ex_w;
    int64_t start_flops;
  };
  std::vector<std::vector<entry>> sched;  // per worker
};

struct pp_program {
  h_ops_tree tree;
  std::unique_ptr<h_ops_dag> dag;
  std::unique_ptr<instructions_scheduler> sched;
  std::unique_ptr<instructions_manager> mgr;
  int workers = 0;
  int n_slots = 0;
  std::vector<int64_t> temp_elems;  // per batch member (slot 3*i+1)
  int64_t scratch = 0;              // per-worker scratch elements
  long long primary_flops = -1;     // FLOPs of the factorization ops alone (factor + inverse programs)
  bool zero_temps = true;           // false: a window after the first (temps carry over)
  double dag_ms = 0, schedule_ms = 0, encode_ms = 0;
  std::unique_ptr<pp_program_image> image;  // decoded programs only
  ~pp_program();
};

namespace ppc {
// FNV-1a 64 of the matrix structure as the byte-string format encodes it.
uint64_t tree_fingerprint(const pp_hmatrix* h);
int program_insts(const pp_program* p);
const std::vector<int32_t>& program_stream(const pp_program* p, int worker);
}  // namespace ppc
