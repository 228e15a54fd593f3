// Strided 2-D data maps and read/write operand sets of one h-op.
//
// Drop-in for rts/ops.h:9-36 of the reference.  Public members and the two
// dependency queries keep their reference signatures; `addData` / `getData`
// are additive (the reference keeps `data` private with no way to fill it,
// rts/ops.h:24).
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

enum class dependency_t : int;

// A region of `height` rows, each `width` bytes long, successive rows `pitch`
// bytes apart, starting at `ptr`.  All quantities are in BYTES
// (rts/ops.cpp:9-12 does its arithmetic on unsigned char*).
class DataMap2D {
public:
  void* ptr;
  size_t pitch;
  size_t width;
  size_t height;

  DataMap2D(void* ptr, size_t pitch, size_t width, size_t height)
      : ptr(ptr), pitch(pitch), width(width), height(height) {}

  // True when some byte lies in both regions (rts/ops.cpp:7-40).
  bool checkOverlap(const DataMap2D& data) const;

  // First and one-past-last byte address the map can touch.
  uintptr_t lo() const { return reinterpret_cast<uintptr_t>(ptr); }
  uintptr_t hi() const { return lo() + pitch * height; }
};

class Operation {
private:
  // (read_only, map): read_only == true means the operand is only read.
  std::vector<std::pair<bool, DataMap2D>> data;

public:
  Operation() : data() {}
  ~Operation() {}

  // Dependency of *this on an earlier op_from (rts/ops.cpp:42-61):
  // from-write & this-read -> flow, from-read & this-write -> anti,
  // both write -> out; every overlapping pair is OR-ed in.
  dependency_t checkDependencyFrom(const Operation& op_from) const;

  // Same relation seen from the producer side (rts/ops.cpp:63-64).
  dependency_t checkDependencyTo(const Operation& op_to) const;

  // Additive: register an operand.
  void addData(bool read_only, const DataMap2D& map) { data.emplace_back(read_only, map); }
  const std::vector<std::pair<bool, DataMap2D>>& getData() const { return data; }
};
