// Task DAG consumed by the list scheduler.
//
// The reference scheduler reads its DAG through h_ops_dag /
// dependency_linked_list (rts/scheduler.cpp:73-77,240,334; constructed at
// src/launcher.cu:53), whose definitions are not part of the reference
// snapshot.  This file specifies them:
//   * nodes are the flattened h-ops [start_index, start_index + length) of an
//     h_ops_tree, renumbered from 0;
//   * there is an edge i -> j (i < j) exactly when
//     op_j.checkDependencyFrom(op_i) != no, i.e. the all-pairs relation of
//     rts/ops.cpp, NOT transitively reduced;
//   * every from-list and to-list is in ascending instruction order;
//   * getFlops_Trim(i) is the algorithmic FLOP count of instruction i (the
//     scheduler itself applies the _MIN_INST_FLOPS floor).
// Candidate pairs are found with an address-interval sweep, so only operand
// pairs whose byte ranges intersect reach checkOverlap; the edge set equals
// the all-pairs one because disjoint ranges cannot overlap.
#pragma once

#include <cstdint>
#include <vector>

#include "dependency.h"
#include "ops.h"

// Floor applied to an instruction's FLOP estimate by the scheduler.
#ifndef _MIN_INST_FLOPS
#define _MIN_INST_FLOPS 65536LL
#endif
// Clock ticks per character and characters per row of analyzeClocks' chart.
#ifndef _TICKS
#define _TICKS 4096ULL
#endif
#ifndef _ROW_BLOCKS
#define _ROW_BLOCKS 96
#endif

class h_ops_tree;

class dependency_linked_list {
private:
  int inst;
  dependency_t dep;
  dependency_linked_list* next;

public:
  dependency_linked_list(int inst_in, dependency_t dep_in, dependency_linked_list* next_in = nullptr)
      : inst(inst_in), dep(dep_in), next(next_in) {}
  ~dependency_linked_list();  // frees the rest of the chain iteratively

  int getInst() const { return inst; }
  dependency_t getDep() const { return dep; }
  dependency_linked_list* getNext() const { return next; }
  int length() const;
};

class h_ops_dag {
private:
  int length;
  std::vector<long long> flops;
  std::vector<dependency_linked_list*> from_lists;  // successors of i
  std::vector<dependency_linked_list*> to_lists;    // predecessors of j
  std::vector<int> count_from, count_to;
  long long n_edges;

  void build(const std::vector<const Operation*>& ops);

public:
  h_ops_dag(const h_ops_tree* tree, int start_index = 0, int length_max = 0);
  // Build from bare operand sets and FLOP counts (used for scheduling DAGs
  // that do not come from an h_ops_tree, e.g. test graphs).
  h_ops_dag(const std::vector<const Operation*>& ops, const std::vector<long long>& flops_in);
  // Build from an explicit edge list (edges[2e] -> edges[2e+1], masks[e]);
  // lists are sorted ascending.
  h_ops_dag(int n, const long long* flops_in, int n_edges_in, const int* edges, const int* masks);
  ~h_ops_dag();
  h_ops_dag(const h_ops_dag&) = delete;
  h_ops_dag& operator=(const h_ops_dag&) = delete;

  int getLength() const { return length; }
  int getDepCount_From(int inst) const { return count_from[inst]; }
  int getDepCount_To(int inst) const { return count_to[inst]; }
  long long getFlops_Trim(int inst) const { return flops[inst]; }
  long long getFlops(int inst) const { return flops[inst]; }
  long long getFlops() const;  // total over all instructions
  const dependency_linked_list* getDepList_To(int inst) const { return to_lists[inst]; }
  const dependency_linked_list* getDepList_From(int inst) const { return from_lists[inst]; }
  // In-degree array, allocated with new[]; the scheduler takes ownership.
  int* getDepCountList_To() const;
  long long getEdgeCount() const { return n_edges; }
  // Edges divided by the all-pairs maximum L(L-1)/2.
  double dag_density() const;
  void print() const;
};
