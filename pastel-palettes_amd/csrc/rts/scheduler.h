// Static list scheduler: maps DAG instructions onto W persistent workers and
// inserts explicit waits for cross-worker dependencies.
//
// Drop-in for rts/scheduler.h:4-126 of the reference.  Signatures are kept;
// analyzeClocks returns hipError_t where the reference returns cudaError_t.
#pragma once

#include <hip/hip_runtime_api.h>

#include "h_ops_dag.h"

// One entry of a worker's output stream: execute (ex_w == true) or wait for
// (ex_w == false) instruction `inst`; `starting_flops_count` is the worker's
// FLOP clock when the entry was appended (rts/scheduler.cpp:5-66).
class instructions_queue {
private:
  int inst;
  bool ex_w;
  long long int starting_flops_count;
  instructions_queue* next;

public:
  instructions_queue(const int inst_in, const bool ex_w_in, const long long int flops_in);
  ~instructions_queue();

  int getInst() const;
  bool getExW() const;
  // FLOP clock advance between this entry and the next one; 0 at the tail
  // (the reference dereferences a null `next` there, rts/scheduler.cpp:22-23).
  long long int getElapsedFlops() const;
  long long int getStartingFlops() const { return starting_flops_count; }
  instructions_queue* getNext() const;
  instructions_queue* setNext(const int inst_in, const bool ex_w_in, const long long int flops_in);
  int getLength() const;
  // ex_w of every entry in order (the reference never advances its index,
  // rts/scheduler.cpp:44-54, so it writes only list[0]).
  bool* getExList(bool* list = nullptr) const;
  void print() const;
};

// A ready instruction together with its synchronisation needs
// (rts/scheduler.cpp:70-209).
class ready_queue {
private:
  int inst_number;
  int num_deps;
  long long int anticipated_flops;

  long long int max_sync;
  int* sync_with;
  ready_queue* next;

public:
  ready_queue(const int inst_in, const int workers, const h_ops_dag* dag, const long long int* flops_after_inst,
              const int* inst_executed_by, ready_queue* next_q = nullptr);
  ~ready_queue();

  int getInst() const;
  int getNumDeps() const;
  long long int getFlops(const long long int min_flops = _MIN_INST_FLOPS) const;
  long long int getMaxSync() const;
  int* getSyncWith() const;
  ready_queue* getNext() const { return next; }
  ready_queue* getLast();
  ready_queue* setNext(ready_queue* next_q);
  bool hookup(ready_queue* queue);
  ready_queue* deleteCriticalNode(ready_queue** deleted_ptr_out, const long long int flops_synced);
  void print() const;
};

class instructions_scheduler {
private:
  int length;
  int workers;

  ready_queue* working_queue;
  ready_queue* working_queue_tail;

  instructions_queue** result_queues;
  instructions_queue** result_queue_tails;

  int* inward_deps_counter;
  long long int* flops_after_inst;
  long long int* flops_worker;

  int* inst_executed_by;
  long long int* last_sync_flops;

  // summary figures of the last schedule() (printed when verbose)
  int comm_wait_counts;
  long long int flops_total;

  long long int getSmallestLoad();
  void loadWorkingQueue(const h_ops_dag* dag);
  long long int findLatestSyncs(const h_ops_dag* dag, int* sync_with, const int inst);
  int findWorkerWithMinimalWaiting(const long long int max_sync);
  void eliminateExtraSync(int* sync_with, const int worker_id);
  void addInstToWorker(const int inst, const long long int flops_anticipated, const int worker_id);
  void addWaitToWorker(const int inst, const int worker_id);
  void updateDepsCounts(const h_ops_dag* dag, const int inst_finished);
  void schedule(const h_ops_dag* dag);

public:
  // Print the "Scheduler Summary" block after each schedule (off by default).
  static bool verbose;

  instructions_scheduler(const h_ops_dag* dag, const int num_workers_limit);
  ~instructions_scheduler();
  instructions_scheduler(const instructions_scheduler&) = delete;
  instructions_scheduler& operator=(const instructions_scheduler&) = delete;

  instructions_queue* getSchedule(const int worker_id) const;
  int getLength(const int worker_id) const;
  // new[]-allocated array of W lengths; the caller releases it with delete[].
  int* getLengths() const;
  int getWorkers() const { return workers; }
  int getCommWaitCounts() const { return comm_wait_counts; }
  long long int getWorkerFlops(const int worker_id) const { return flops_worker[worker_id]; }

  // Download per-worker entry timestamps (device array of W device pointers,
  // each holding 1 + getLength(w) values) and print a Gantt chart, green for
  // execute and red for wait (rts/scheduler.cpp:494-544).  Unlike the
  // reference it does not free the device buffers: their owner does.
  hipError_t analyzeClocks(unsigned long long int** clocks_gpu, const unsigned long long int block_ticks = _TICKS,
                           const int row_blocks = _ROW_BLOCKS) const;

  void print() const;
};
