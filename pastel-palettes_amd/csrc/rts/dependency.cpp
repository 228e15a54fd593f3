// Edge table of the task DAG (reference rts/dependency.cpp).
#include "dependency.h"

#include <cstdio>

dependency_t operator+(dependency_t dep1, dependency_t dep2) {
  return static_cast<dependency_t>(static_cast<int>(dep1) | static_cast<int>(dep2));
}

bool operator>(dependency_t dep1, dependency_t dep2) {
  return static_cast<int>(dep1) > static_cast<int>(dep2);
}

namespace {
// Remove the single (key, value) entry from a multimap.  The reference erases
// inside its equal_range loop (rts/dependency.cpp:26-33), which invalidates
// the loop iterator; here the entry is located first and erased once.
void erase_pair(std::multimap<int64_t, int64_t>& mm, int64_t key, int64_t value) {
  auto range = mm.equal_range(key);
  for (auto it = range.first; it != range.second; ++it) {
    if (it->second == value) {
      mm.erase(it);
      return;
    }
  }
}
}  // namespace

void dependency_table::update(int64_t inst_from, int64_t inst_to, dependency_t dep) {
  const auto key = std::make_pair(inst_from, inst_to);
  if (dep > dependency_t::no) {
    if (table.insert(key).second) {
      table_f.emplace(inst_from, inst_to);
      table_t.emplace(inst_to, inst_from);
      type.emplace(key, dep);
    } else {
      type[key] = dep;
    }
    return;
  }
  auto it = table.find(key);
  if (it == table.end()) return;
  table.erase(it);
  erase_pair(table_f, inst_from, inst_to);
  erase_pair(table_t, inst_to, inst_from);
  type.erase(key);
}

dependency_t dependency_table::lookup(int64_t inst_from, int64_t inst_to) const {
  auto it = type.find(std::make_pair(inst_from, inst_to));
  return it == type.end() ? dependency_t::no : it->second;
}

void dependency_table::getDepFrom(int64_t from, std::vector<int64_t>& lis) const {
  lis.clear();
  auto range = table_f.equal_range(from);
  for (auto it = range.first; it != range.second; ++it) lis.push_back(it->second);
}

void dependency_table::getDepTo(int64_t to, std::vector<int64_t>& lis) const {
  lis.clear();
  auto range = table_t.equal_range(to);
  for (auto it = range.first; it != range.second; ++it) lis.push_back(it->second);
}

size_t dependency_table::countDepFrom(int64_t from) const { return table_f.count(from); }

size_t dependency_table::countDepTo(int64_t to) const { return table_t.count(to); }

// The reference leaves `type` populated after clear() (rts/dependency.cpp:67-71),
// and std::map::insert does not overwrite, so an edge re-inserted after a
// clear keeps its stale mask there; all four containers are emptied here.
void dependency_table::clear() {
  table.clear();
  table_f.clear();
  table_t.clear();
  type.clear();
}

void dependency_table::print() const {
  for (const auto& e : table) std::printf("%lld %lld\n", (long long)e.first, (long long)e.second);
}
