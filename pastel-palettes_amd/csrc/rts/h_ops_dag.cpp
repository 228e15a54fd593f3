// DAG construction over flattened h-ops (see h_ops_dag.h for the contract).
#include "h_ops_dag.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>

#include "../h_ops/h_ops.h"

dependency_linked_list::~dependency_linked_list() {
  dependency_linked_list* p = next;
  while (p != nullptr) {
    dependency_linked_list* n = p->next;
    p->next = nullptr;
    delete p;
    p = n;
  }
}

int dependency_linked_list::length() const {
  int l = 0;
  for (const dependency_linked_list* p = this; p != nullptr; p = p->next) ++l;
  return l;
}

namespace {

struct Interval {
  uintptr_t lo, hi;
  int op;
};

// All (i, j), i < j, whose operands have intersecting byte ranges.
std::vector<std::pair<int, int>> candidate_pairs(const std::vector<const Operation*>& ops) {
  std::vector<Interval> iv;
  for (int k = 0; k < (int)ops.size(); ++k)
    for (const auto& d : ops[k]->getData())
      if (d.second.hi() > d.second.lo()) iv.push_back({d.second.lo(), d.second.hi(), k});
  std::sort(iv.begin(), iv.end(), [](const Interval& a, const Interval& b) {
    return a.lo != b.lo ? a.lo < b.lo : a.op < b.op;
  });
  std::vector<std::pair<int, int>> pairs;
  for (size_t a = 0; a < iv.size(); ++a) {
    for (size_t b = a + 1; b < iv.size() && iv[b].lo < iv[a].hi; ++b) {
      int i = iv[a].op, j = iv[b].op;
      if (i == j) continue;
      if (i > j) std::swap(i, j);
      pairs.emplace_back(i, j);
    }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  return pairs;
}

}  // namespace

void h_ops_dag::build(const std::vector<const Operation*>& ops) {
  length = (int)ops.size();
  from_lists.assign(length, nullptr);
  to_lists.assign(length, nullptr);
  count_from.assign(length, 0);
  count_to.assign(length, 0);
  n_edges = 0;

  std::vector<std::vector<std::pair<int, dependency_t>>> succ(length), pred(length);
  for (const auto& pr : candidate_pairs(ops)) {
    const dependency_t dep = ops[pr.second]->checkDependencyFrom(*ops[pr.first]);
    if (dep > dependency_t::no) {
      succ[pr.first].emplace_back(pr.second, dep);
      pred[pr.second].emplace_back(pr.first, dep);
      ++n_edges;
    }
  }
  // candidate pairs are sorted by (i, j), so succ lists are ascending; pred
  // lists are filled in ascending i for each j as well.
  for (int i = 0; i < length; ++i) {
    count_from[i] = (int)succ[i].size();
    count_to[i] = (int)pred[i].size();
    for (int k = (int)succ[i].size() - 1; k >= 0; --k)
      from_lists[i] = new dependency_linked_list(succ[i][k].first, succ[i][k].second, from_lists[i]);
    for (int k = (int)pred[i].size() - 1; k >= 0; --k)
      to_lists[i] = new dependency_linked_list(pred[i][k].first, pred[i][k].second, to_lists[i]);
  }
}

h_ops_dag::h_ops_dag(const h_ops_tree* tree, int start_index, int length_max) {
  const int total = tree->length();
  if (start_index < 0 || start_index > total) throw std::out_of_range("h_ops_dag: start_index");
  const int len = (length_max <= 0) ? total - start_index : std::min(length_max, total - start_index);
  std::vector<const Operation*> ops;
  flops.clear();
  for (int i = 0; i < len; ++i) {
    ops.push_back(&tree->getOp(start_index + i).data);
    flops.push_back(tree->getOp(start_index + i).flops);
  }
  build(ops);
}

h_ops_dag::h_ops_dag(const std::vector<const Operation*>& ops, const std::vector<long long>& flops_in)
    : flops(flops_in) {
  if (flops.size() != ops.size()) throw std::invalid_argument("h_ops_dag: flops/ops size");
  build(ops);
}

h_ops_dag::h_ops_dag(int n, const long long* flops_in, int n_edges_in, const int* edges, const int* masks) {
  length = n;
  flops.assign(flops_in, flops_in + n);
  from_lists.assign(n, nullptr);
  to_lists.assign(n, nullptr);
  count_from.assign(n, 0);
  count_to.assign(n, 0);
  n_edges = 0;
  std::vector<std::vector<std::pair<int, dependency_t>>> succ(n), pred(n);
  for (int e = 0; e < n_edges_in; ++e) {
    const int a = edges[2 * e], b = edges[2 * e + 1];
    if (a < 0 || b < 0 || a >= n || b >= n) throw std::out_of_range("h_ops_dag: edge endpoint");
    const dependency_t dep = (dependency_t)masks[e];
    succ[a].emplace_back(b, dep);
    pred[b].emplace_back(a, dep);
    if (dep > dependency_t::no) ++n_edges;
  }
  for (int i = 0; i < n; ++i) {
    std::stable_sort(succ[i].begin(), succ[i].end(),
                     [](const std::pair<int, dependency_t>& x, const std::pair<int, dependency_t>& y) { return x.first < y.first; });
    std::stable_sort(pred[i].begin(), pred[i].end(),
                     [](const std::pair<int, dependency_t>& x, const std::pair<int, dependency_t>& y) { return x.first < y.first; });
    for (const auto& s : succ[i]) if (s.second > dependency_t::no) ++count_from[i];
    for (const auto& p : pred[i]) if (p.second > dependency_t::no) ++count_to[i];
    for (int k = (int)succ[i].size() - 1; k >= 0; --k)
      from_lists[i] = new dependency_linked_list(succ[i][k].first, succ[i][k].second, from_lists[i]);
    for (int k = (int)pred[i].size() - 1; k >= 0; --k)
      to_lists[i] = new dependency_linked_list(pred[i][k].first, pred[i][k].second, to_lists[i]);
  }
}

h_ops_dag::~h_ops_dag() {
  for (auto* p : from_lists) delete p;
  for (auto* p : to_lists) delete p;
}

long long h_ops_dag::getFlops() const {
  long long s = 0;
  for (long long f : flops) s += f;
  return s;
}

int* h_ops_dag::getDepCountList_To() const {
  int* c = new int[length > 0 ? length : 1];
  for (int i = 0; i < length; ++i) c[i] = count_to[i];
  return c;
}

double h_ops_dag::dag_density() const {
  if (length < 2) return 0.;
  return (double)n_edges / (0.5 * (double)length * (double)(length - 1));
}

void h_ops_dag::print() const {
  for (int i = 0; i < length; ++i) {
    std::printf("%d:", i);
    for (const dependency_linked_list* p = from_lists[i]; p != nullptr; p = p->getNext())
      std::printf(" %d(%d)", p->getInst(), (int)p->getDep());
    std::printf("\n");
  }
}
