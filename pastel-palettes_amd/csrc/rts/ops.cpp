// Overlap test and dependency classification (reference rts/ops.cpp).
#include "ops.h"
#include "dependency.h"

#include <algorithm>

namespace {

// Does some row of `hi` (whose start address is >= row_begin's map start)
// intersect the byte interval [s, e)?  Rows of `hi` are scanned from the first
// row that starts at or before s, exactly like the candidate walk of
// rts/ops.cpp:24-33, but the walk is replaced by a direct index computation:
// rows of `hi` start at h0 + j*pitch, so the first row that can still reach
// past s is found with one division.
bool row_hits(uintptr_t s, uintptr_t e, const DataMap2D& hi) {
  const uintptr_t h0 = hi.lo();
  const uintptr_t h_end = hi.hi();
  if (hi.pitch == 0) return false;  // zero pitch spans no bytes (hi() == lo())
  // first candidate row: the row containing s (or row 0 when s < h0)
  size_t j = s > h0 ? (s - h0) / hi.pitch : 0;
  // walk forward while the candidate row starts before e and inside hi
  for (;; ++j) {
    const uintptr_t c = h0 + j * hi.pitch;
    if (!(c < e && c < h_end)) return false;
    if (c + hi.width > s && e > c) return true;
    // rows are sorted by start; once a row ends past s every later row
    // starting before e intersects, so at most two iterations are needed
    // when width <= pitch.  Keep walking for the width > pitch case.
  }
}

}  // namespace

bool DataMap2D::checkOverlap(const DataMap2D& data) const {
  // Order the pair by start address; equal addresses keep *this as `lo`.
  const bool swap = data.lo() < lo();
  const DataMap2D& a = swap ? data : *this;
  const DataMap2D& b = swap ? *this : data;

  const uintptr_t d = b.lo() - a.lo();
  if (d >= a.pitch * a.height) return false;  // b starts past a's last row
  if (d == 0) return true;                     // same start byte

  // Rows of `a` that can meet `b`: from the row holding b's first byte up to
  // the last row that starts before b's end.
  const uintptr_t b_end = b.hi();
  const size_t first = d / a.pitch;
  const size_t last = std::min(a.height, (size_t)((b_end - a.lo()) + a.pitch - 1) / a.pitch);
  for (size_t i = first; i < last; ++i) {
    const uintptr_t s = a.lo() + i * a.pitch;
    if (row_hits(s, s + a.width, b)) return true;
  }
  return false;
}

dependency_t Operation::checkDependencyFrom(const Operation& op_from) const {
  int mask = 0;
  for (const auto& f : op_from.data) {
    for (const auto& t : data) {
      if (!f.second.checkOverlap(t.second)) continue;
      const bool f_writes = !f.first, t_writes = !t.first;
      if (f_writes && !t_writes) mask |= (int)dependency_t::flow;
      if (!f_writes && t_writes) mask |= (int)dependency_t::anti;
      if (f_writes && t_writes) mask |= (int)dependency_t::out;
    }
  }
  return (dependency_t)mask;
}

dependency_t Operation::checkDependencyTo(const Operation& op_to) const {
  return op_to.checkDependencyFrom(*this);
}
