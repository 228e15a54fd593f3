// Encodes a schedule into per-worker int32 instruction streams.
//
// The reference encoder (instructions/instructions_manager.cuh, used at
// src/launcher.cu:72,82) is absent from its snapshot; its output format is
// pinned by the kernel decoder (src/kernel.cu:1384-1522).  Each worker's
// stream is its scheduler queue in order: execute entries carry the h-op's
// words with the id slot set to the DAG instruction id, wait entries are
// [signal_wait, id, tick_hint], and the stream ends with [finish].
//
// tick_hint (a polling back-off hint) is this library's choice:
//   clamp((finish_flops(src) - wait_start_flops) >> 18, 0, 63)
// where both FLOP clocks come from the scheduler's cost model.
#pragma once

#include <cstdint>
#include <vector>

#include "../h_ops/h_ops.h"
#include "../rts/scheduler.h"

class instructions_manager {
private:
  std::vector<std::vector<int32_t>> streams;
  int n_insts;

public:
  instructions_manager(const int workers, const h_ops_dag* dag, const instructions_scheduler* schedule,
                       const h_ops_tree* tree, const int start_index = 0);

  int getWorkers() const { return (int)streams.size(); }
  int getInstCount() const { return n_insts; }
  const std::vector<int32_t>& getStream(int w) const { return streams[w]; }
  int64_t totalWords() const;
  void print() const;
};
