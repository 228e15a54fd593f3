// Static list scheduler (reference rts/scheduler.cpp).
//
// The output must equal the reference's for the same DAG, worker count and
// FLOP floor, so every tie-break below follows the reference:
//   * roots enter the ready list by prepending in index order, i.e. the list
//     starts in descending index order (rts/scheduler.cpp:224-236);
//   * a ready entry's per-worker sync target is the predecessor that finished
//     last on that worker, first maximum wins (strict <), and max_sync is the
//     latest finish over all predecessors (rts/scheduler.cpp:70-102);
//   * hookup() merges a sorted list after the head only, placing an entry
//     after all entries with equal max_sync (rts/scheduler.cpp:142-161);
//   * the critical-node rule and the worker choice follow
//     rts/scheduler.cpp:163-200 and 267-287;
//   * successors are released in from-list order and inserted as in
//     rts/scheduler.cpp:330-370.
#include "scheduler.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

bool instructions_scheduler::verbose = false;

// ---------------------------------------------------------------- instructions_queue

instructions_queue::instructions_queue(const int inst_in, const bool ex_w_in, const long long int flops_in)
    : inst(inst_in), ex_w(ex_w_in), starting_flops_count(flops_in), next(nullptr) {}

instructions_queue::~instructions_queue() {
  // iterative release: worker streams can be tens of thousands long
  instructions_queue* p = next;
  while (p != nullptr) {
    instructions_queue* n = p->next;
    p->next = nullptr;
    delete p;
    p = n;
  }
}

int instructions_queue::getInst() const { return inst; }
bool instructions_queue::getExW() const { return ex_w; }

long long int instructions_queue::getElapsedFlops() const {
  return next == nullptr ? 0 : next->starting_flops_count - starting_flops_count;
}

instructions_queue* instructions_queue::getNext() const { return next; }

instructions_queue* instructions_queue::setNext(const int inst_in, const bool ex_w_in, const long long int flops_in) {
  instructions_queue* q = new instructions_queue(inst_in, ex_w_in, flops_in);
  q->next = next;
  next = q;
  return q;
}

int instructions_queue::getLength() const {
  int l = 0;
  for (const instructions_queue* p = this; p != nullptr; p = p->next) ++l;
  return l;
}

bool* instructions_queue::getExList(bool* list) const {
  if (list == nullptr) list = new bool[getLength()];
  int i = 0;
  for (const instructions_queue* p = this; p != nullptr; p = p->next) list[i++] = p->ex_w;
  return list;
}

void instructions_queue::print() const {
  for (const instructions_queue* p = this; p != nullptr; p = p->next)
    std::printf("%c%d ", p->ex_w ? 'i' : 'w', p->inst);
  std::printf("\n");
}

// ---------------------------------------------------------------- ready_queue

namespace {
// Latest-finishing predecessor per worker and over all workers.
long long int latest_syncs(const h_ops_dag* dag, int inst, int workers, const long long int* flops_after_inst,
                           const int* inst_executed_by, int* sync_with) {
  std::vector<long long int> best(workers, 0);
  for (int w = 0; w < workers; ++w) sync_with[w] = -1;
  long long int max_sync = 0;
  for (const dependency_linked_list* p = dag->getDepList_To(inst); p != nullptr; p = p->getNext()) {
    if (!(p->getDep() > dependency_t::no_dep)) continue;
    const int src = p->getInst();
    const int w = inst_executed_by[src];
    const long long int f = flops_after_inst[src];
    if (best[w] < f) {
      best[w] = f;
      sync_with[w] = src;
    }
    if (max_sync < f) max_sync = f;
  }
  return max_sync;
}
}  // namespace

ready_queue::ready_queue(const int inst_in, const int workers, const h_ops_dag* dag,
                         const long long int* flops_after_inst, const int* inst_executed_by, ready_queue* next_q)
    : inst_number(inst_in),
      num_deps(dag->getDepCount_From(inst_in)),
      anticipated_flops(dag->getFlops_Trim(inst_in)),
      max_sync(0),
      sync_with(new int[workers]),
      next(next_q) {
  max_sync = latest_syncs(dag, inst_in, workers, flops_after_inst, inst_executed_by, sync_with);
}

ready_queue::~ready_queue() {
  delete[] sync_with;
  ready_queue* p = next;
  while (p != nullptr) {
    ready_queue* n = p->next;
    p->next = nullptr;
    delete p;
    p = n;
  }
}

int ready_queue::getInst() const { return inst_number; }
int ready_queue::getNumDeps() const { return num_deps; }

long long int ready_queue::getFlops(const long long int min_flops) const {
  return anticipated_flops > min_flops ? anticipated_flops : min_flops;
}

long long int ready_queue::getMaxSync() const { return max_sync; }
int* ready_queue::getSyncWith() const { return sync_with; }

ready_queue* ready_queue::getLast() {
  ready_queue* p = this;
  while (p->next != nullptr) p = p->next;
  return p;
}

ready_queue* ready_queue::setNext(ready_queue* next_q) {
  ready_queue* old = next;
  next = next_q;
  return old;
}

bool ready_queue::hookup(ready_queue* queue) {
  ready_queue* prev = this;   // node after which the next insertion goes
  ready_queue* cur = next;    // first node not yet passed
  while (queue != nullptr) {
    ready_queue* item = queue;
    queue = queue->next;
    while (cur != nullptr && cur->max_sync <= item->max_sync) {
      prev = cur;
      cur = cur->next;
    }
    item->next = cur;
    prev->next = item;
    prev = item;
  }
  return cur == nullptr;
}

ready_queue* ready_queue::deleteCriticalNode(ready_queue** deleted_ptr_out, const long long int flops_synced) {
  // Take the head unless it would stall the least-loaded worker and some
  // entry that would not stall it has strictly more successors.
  ready_queue* pick_prev = nullptr;
  if (!(flops_synced < max_sync)) {
    int most = num_deps;
    ready_queue* prev = this;
    for (ready_queue* p = next; p != nullptr && p->max_sync <= flops_synced; prev = p, p = p->next) {
      if (p->num_deps > most) {
        most = p->num_deps;
        pick_prev = prev;
      }
    }
  }
  if (pick_prev == nullptr) {
    ready_queue* rest = next;
    next = nullptr;
    *deleted_ptr_out = this;
    return rest;
  }
  ready_queue* target = pick_prev->next;
  pick_prev->next = target->next;
  target->next = nullptr;
  *deleted_ptr_out = target;
  return this;
}

void ready_queue::print() const {
  int n = 0;
  std::printf("Ready Queue: \n");
  for (const ready_queue* p = this; p != nullptr; p = p->next, ++n)
    std::printf("i%d: flops %lld start %lld. \n", p->inst_number, p->anticipated_flops, p->max_sync);
  std::printf("\nLength: %d. \n", n);
}

// ---------------------------------------------------------------- instructions_scheduler

long long int instructions_scheduler::getSmallestLoad() {
  long long int m = flops_worker[0];
  for (int w = 1; w < workers; ++w)
    if (flops_worker[w] < m) m = flops_worker[w];
  return m;
}

void instructions_scheduler::loadWorkingQueue(const h_ops_dag* dag) {
  for (int i = 0; i < length; ++i) {
    if (inward_deps_counter[i] != 0) continue;
    working_queue = new ready_queue(i, workers, dag, flops_after_inst, inst_executed_by, working_queue);
    if (working_queue_tail == nullptr) working_queue_tail = working_queue;
  }
}

long long int instructions_scheduler::findLatestSyncs(const h_ops_dag* dag, int* sync_with, const int inst) {
  return latest_syncs(dag, inst, workers, flops_after_inst, inst_executed_by, sync_with);
}

int instructions_scheduler::findWorkerWithMinimalWaiting(const long long int max_sync) {
  // Prefer the busiest worker that is already past max_sync-free time
  // (load <= max_sync); otherwise the least-loaded one, scanning from worker
  // 0 whose load seeds the comparison.
  int before = -1, after = 0;
  long long int before_load = -1, after_load = flops_worker[0];
  for (int w = 0; w < workers; ++w) {
    const long long int f = flops_worker[w];
    if (f <= max_sync) {
      if (f > before_load) {
        before = w;
        before_load = f;
      }
    } else if (f < after_load) {
      after = w;
      after_load = f;
    }
  }
  return before >= 0 ? before : after;
}

void instructions_scheduler::eliminateExtraSync(int* sync_with, const int worker_id) {
  sync_with[worker_id] = -1;
  const long long int* seen = &last_sync_flops[(size_t)worker_id * workers];
  for (int w = 0; w < workers; ++w) {
    const int src = sync_with[w];
    if (src >= 0 && flops_after_inst[src] <= seen[w]) sync_with[w] = -1;
  }
}

void instructions_scheduler::addInstToWorker(const int inst, const long long int flops_anticipated, const int worker_id) {
  if (result_queues[worker_id] == nullptr)
    result_queues[worker_id] = result_queue_tails[worker_id] = new instructions_queue(inst, true, flops_worker[worker_id]);
  else
    result_queue_tails[worker_id] = result_queue_tails[worker_id]->setNext(inst, true, flops_worker[worker_id]);
  flops_worker[worker_id] += flops_anticipated;
  flops_after_inst[inst] = flops_worker[worker_id];
  inst_executed_by[inst] = worker_id;
}

void instructions_scheduler::addWaitToWorker(const int inst, const int worker_id) {
  if (result_queues[worker_id] == nullptr)
    result_queues[worker_id] = result_queue_tails[worker_id] = new instructions_queue(inst, false, flops_worker[worker_id]);
  else
    result_queue_tails[worker_id] = result_queue_tails[worker_id]->setNext(inst, false, flops_worker[worker_id]);
  const long long int f = flops_after_inst[inst];
  if (flops_worker[worker_id] < f) flops_worker[worker_id] = f;
  last_sync_flops[(size_t)worker_id * workers + inst_executed_by[inst]] = f;
}

void instructions_scheduler::updateDepsCounts(const h_ops_dag* dag, const int inst_finished) {
  ready_queue* fresh = nullptr;  // newly ready entries, ascending max_sync
  ready_queue* fresh_tail = nullptr;
  for (const dependency_linked_list* p = dag->getDepList_From(inst_finished); p != nullptr; p = p->getNext()) {
    if (!(p->getDep() > dependency_t::no_dep)) continue;
    const int s = p->getInst();
    if (--inward_deps_counter[s] != 0) continue;
    ready_queue* e = new ready_queue(s, workers, dag, flops_after_inst, inst_executed_by);
    if (fresh == nullptr) {
      fresh = fresh_tail = e;
    } else if (e->getMaxSync() <= fresh->getMaxSync()) {
      e->setNext(fresh);
      fresh = e;
    } else if (fresh->hookup(e)) {
      fresh_tail = e;
    }
  }
  if (working_queue == nullptr) {
    working_queue = fresh;
    working_queue_tail = fresh_tail;
  } else if (fresh != nullptr && working_queue->hookup(fresh)) {
    working_queue_tail = fresh_tail;
  }
}

void instructions_scheduler::schedule(const h_ops_dag* dag) {
  comm_wait_counts = 0;
  flops_total = 0;
  loadWorkingQueue(dag);

  for (int done = 0; done < length; ++done) {
    if (working_queue == nullptr)
      throw std::runtime_error("instructions_scheduler: DAG has a cycle or an inconsistent in-degree");
    ready_queue* pick = nullptr;
    working_queue = working_queue->deleteCriticalNode(&pick, getSmallestLoad());

    const int inst = pick->getInst();
    int* sync_with = pick->getSyncWith();
    const long long int flops = pick->getFlops(_MIN_INST_FLOPS);
    const int worker = findWorkerWithMinimalWaiting(pick->getMaxSync());

    eliminateExtraSync(sync_with, worker);
    for (int w = 0; w < workers; ++w) {
      if (sync_with[w] >= 0) {
        addWaitToWorker(sync_with[w], worker);
        ++comm_wait_counts;
      }
    }
    delete pick;

    addInstToWorker(inst, flops, worker);
    updateDepsCounts(dag, inst);
    flops_total += flops;
  }

  if (verbose && length > 0) {
    long long int heaviest = 0;
    for (int w = 0; w < workers; ++w)
      if (flops_worker[w] > heaviest) heaviest = flops_worker[w];
    std::printf(
        "-- Scheduler Summary --\n"
        "Total # of Instructions: %d. \n"
        "Percent Heaviest Loaded Worker to Total Ratio: %f%%. \n"
        "Utilization Percent: %f%%. \n"
        "- Low FLOPS Instructions are trimmed.\n"
        "- Utilization may appear higher than actual.\n"
        "Avg. # of Communications per Instruction: %f. \n\n",
        length, 100. * heaviest / flops_total, 100. * flops_total / heaviest / workers,
        1. * comm_wait_counts / length);
  }
}

instructions_scheduler::instructions_scheduler(const h_ops_dag* dag, const int num_workers_limit)
    : length(dag->getLength()),
      workers(num_workers_limit),
      working_queue(nullptr),
      working_queue_tail(nullptr),
      comm_wait_counts(0),
      flops_total(0) {
  if (workers < 1) throw std::invalid_argument("instructions_scheduler: need at least one worker");
  result_queues = new instructions_queue*[workers]();
  result_queue_tails = new instructions_queue*[workers]();
  flops_after_inst = new long long int[length > 0 ? length : 1]();
  flops_worker = new long long int[workers]();
  inst_executed_by = new int[length > 0 ? length : 1];
  std::memset(inst_executed_by, 0xff, sizeof(int) * (length > 0 ? length : 1));
  last_sync_flops = new long long int[(size_t)workers * workers]();
  inward_deps_counter = dag->getDepCountList_To();
  schedule(dag);
}

instructions_scheduler::~instructions_scheduler() {
  delete working_queue;
  for (int w = 0; w < workers; ++w) delete result_queues[w];
  delete[] result_queues;
  delete[] result_queue_tails;
  delete[] inward_deps_counter;
  delete[] flops_after_inst;
  delete[] flops_worker;
  delete[] inst_executed_by;
  delete[] last_sync_flops;
}

instructions_queue* instructions_scheduler::getSchedule(const int worker_id) const {
  return (worker_id >= 0 && worker_id < workers) ? result_queues[worker_id] : nullptr;
}

int instructions_scheduler::getLength(const int worker_id) const {
  if (worker_id < 0 || worker_id >= workers || result_queues[worker_id] == nullptr) return 0;
  return result_queues[worker_id]->getLength();
}

int* instructions_scheduler::getLengths() const {
  int* l = new int[workers];
  for (int w = 0; w < workers; ++w) l[w] = getLength(w);
  return l;
}

hipError_t instructions_scheduler::analyzeClocks(unsigned long long int** clocks_gpu,
                                                 const unsigned long long int block_ticks,
                                                 const int row_blocks) const {
  std::vector<unsigned long long int*> rows(workers);
  hipError_t err = hipMemcpy(rows.data(), clocks_gpu, workers * sizeof(unsigned long long int*), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return err;
  for (int w = 0; w < workers; ++w) {
    const int n = getLength(w);
    std::vector<unsigned long long int> clk(n + 1);
    err = hipMemcpy(clk.data(), rows[w], (n + 1) * sizeof(unsigned long long int), hipMemcpyDeviceToHost);
    if (err != hipSuccess) return err;
    std::vector<char> ex(n > 0 ? n : 1);
    if (n > 0) {
      bool* tmp = new bool[n];
      result_queues[w]->getExList(tmp);
      for (int i = 0; i < n; ++i) ex[i] = tmp[i];
      delete[] tmp;
    }
    const unsigned long long int t0 = clk[0], t_end = clk[n];
    std::printf("worker: %d, ticks: %llu\n", w, t_end - t0);
    // one character per block_ticks; colour by the entry running at that time
    int entry = 0, col = 0;
    bool green = false;
    std::printf("\033[41m");
    for (unsigned long long int t = t0; t <= t_end && entry < n; t += block_ticks) {
      while (entry < n && clk[entry + 1] <= t) ++entry;
      if (entry >= n) break;
      const bool want = ex[entry] != 0;
      if (want != green) {
        green = want;
        std::printf(green ? "\033[42m" : "\033[41m");
      }
      std::printf(" ");
      if (++col == row_blocks) {
        col = 0;
        std::printf("\033[0m\n%s", green ? "\033[42m" : "\033[41m");
      }
    }
    std::printf("\033[0m\n");
  }
  return hipSuccess;
}

void instructions_scheduler::print() const {
  if (working_queue != nullptr) working_queue->print();
  for (int w = 0; w < workers; ++w) {
    std::printf("Worker #%d: ", w);
    if (result_queues[w] != nullptr) result_queues[w]->print(); else std::printf("\n");
    std::printf("flops: %lld \n", flops_worker[w]);
  }
}
