// Persistent "virtual machine" kernel: one workgroup per scheduler worker
// interprets that worker's instruction stream (reference kernel_dynamic,
// src/kernel.cu:1375-1526).
//
// Cross-worker signalling: after an execute entry every thread issues an
// agent-scope release fence, the workgroup meets at a barrier, and thread 0
// stores `epoch` into flags[id] with an agent-scope release store.  A wait
// entry polls flags[id] with agent-scope acquire loads (with s_sleep back-off)
// until it reaches `epoch`, then every thread issues an agent-scope acquire
// fence.  Agent scope is required on MI355X because its 8 XCDs have private
// L2 caches.  Flags hold the launch epoch, so they never need clearing between
// launches: a launch with epoch e waits for flags >= e.
//
// The epoch is read from device memory (bumped by a one-thread kernel just
// before the VM launch) so that a captured hipGraph replays correctly.
//
// Watchdog: a wait gives up after `spin_limit` wall-clock ticks, records the worker and
// instruction in status[], and raises the abort flag that every other poller
// checks, so a lost signal ends the launch with an error instead of a hang.
#pragma once

#include <stdint.h>

namespace pp {

struct VmArgs {
  const int32_t* words;       // all worker streams, concatenated
  const int64_t* stream_off;  // [workers] start of each stream in `words`
  void* const* ptrs;          // pointer table
  uint32_t* flags;            // [n_insts] completion epochs
  void* scratch;              // per-worker scratch, worker w at w * scratch_stride elements
  int64_t scratch_stride;
  const void* rnd;            // sketch matrix Omega (storage precision)
  uint64_t* trace;            // optional: per-entry timestamps
  const int64_t* trace_off;   // [workers] start of each worker's trace
  int32_t* status;            // [0] code, [1] worker, [2] instruction, [3] abort flag, [4] entries executed,
                              // [5] Householder fallbacks of accum
  int64_t spin_limit;         // wall-clock ticks (100 MHz) a wait may last
  const uint32_t* epoch;      // device counter bumped before every launch
  int32_t n_insts;
};

// One execute entry for the single-op kernel, passed by value (no device
// copy of the words or the pointer table is needed).
constexpr int kSingleMaxWords = 32, kSingleMaxPtrs = 16;
struct SingleOp {
  int32_t words[kSingleMaxWords];
  void* ptrs[kSingleMaxPtrs];
};

// status codes
enum : int32_t { VM_OK = 0, VM_TIMEOUT = 1, VM_BAD_OPCODE = 2, VM_BAD_OPERATION = 3, VM_ABORTED = 4 };

}  // namespace pp
