// C wrapper around the reference's own rts/ops.cpp and rts/dependency.cpp
// (compiled from /root/reference by oracle/Makefile) -- test infrastructure.
// Exposes DataMap2D::checkOverlap and the dependency_table container.
#include <ops.h>
#include <dependency.h>

#include <cstdint>
#include <vector>

extern "C" {

int ref_check_overlap(uint64_t pa, uint64_t pitch_a, uint64_t wa, uint64_t ha, uint64_t pb, uint64_t pitch_b,
                      uint64_t wb, uint64_t hb) {
  DataMap2D a(reinterpret_cast<void*>(pa), pitch_a, wa, ha), b(reinterpret_cast<void*>(pb), pitch_b, wb, hb);
  return a.checkOverlap(b) ? 1 : 0;
}

void* ref_table_new() { return new dependency_table(); }
void ref_table_free(void* t) { delete static_cast<dependency_table*>(t); }
void ref_table_update(void* t, int64_t f, int64_t to, int dep) {
  static_cast<dependency_table*>(t)->update(f, to, static_cast<dependency_t>(dep));
}
int ref_table_lookup(void* t, int64_t f, int64_t to) {
  return static_cast<int>(static_cast<dependency_table*>(t)->lookup(f, to));
}
int ref_table_from(void* t, int64_t f, int64_t* out, int cap) {
  std::vector<int64_t> v;
  static_cast<dependency_table*>(t)->getDepFrom(f, v);
  for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
  return (int)v.size();
}
int ref_table_to(void* t, int64_t to, int64_t* out, int cap) {
  std::vector<int64_t> v;
  static_cast<dependency_table*>(t)->getDepTo(to, v);
  for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
  return (int)v.size();
}
int64_t ref_table_count_from(void* t, int64_t f) { return (int64_t) static_cast<dependency_table*>(t)->countDepFrom(f); }
int64_t ref_table_count_to(void* t, int64_t to) { return (int64_t) static_cast<dependency_table*>(t)->countDepTo(to); }
}
