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
  double dag_ms = 0, schedule_ms = 0, encode_ms = 0;
  ~pp_program();
};
