// Schedule -> instruction streams (see instructions_manager.h).
#include "instructions_manager.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>

instructions_manager::instructions_manager(const int workers, const h_ops_dag* dag,
                                           const instructions_scheduler* schedule, const h_ops_tree* tree,
                                           const int start_index)
    : streams(workers), n_insts(dag->getLength()) {
  // expected finish clock of every instruction under the scheduler's model
  std::vector<long long> done_at(n_insts, 0);
  for (int w = 0; w < workers; ++w)
    for (const instructions_queue* q = schedule->getSchedule(w); q != nullptr; q = q->getNext())
      if (q->getExW()) {
        const long long f = std::max(dag->getFlops_Trim(q->getInst()), (long long)_MIN_INST_FLOPS);
        done_at[q->getInst()] = q->getStartingFlops() + f;
      }

  for (int w = 0; w < workers; ++w) {
    std::vector<int32_t>& s = streams[w];
    for (const instructions_queue* q = schedule->getSchedule(w); q != nullptr; q = q->getNext()) {
      const int id = q->getInst();
      if (id < 0 || id >= n_insts) throw std::out_of_range("instructions_manager: instruction id");
      if (q->getExW()) {
        const h_op& op = tree->getOp(start_index + id);
        const size_t at = s.size();
        s.insert(s.end(), op.words.begin(), op.words.end());
        s[at + 1] = id;
      } else {
        const long long gap = done_at[id] - q->getStartingFlops();
        const int hint = (int)std::min<long long>(63, std::max<long long>(0, gap >> 18));
        s.push_back(signal_wait);
        s.push_back(id);
        s.push_back(hint);
      }
    }
    s.push_back(opcode_t::finish);
  }
}

int64_t instructions_manager::totalWords() const {
  int64_t t = 0;
  for (const auto& s : streams) t += (int64_t)s.size();
  return t;
}

void instructions_manager::print() const {
  for (size_t w = 0; w < streams.size(); ++w) {
    std::printf("worker %zu:", w);
    for (int32_t x : streams[w]) std::printf(" %d", x);
    std::printf("\n");
  }
}
