// Internal state of the device runtime shared by runtime.cpp and solver.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/pastel.h"
#include "../capi_common.h"
#include "../kernels/kernels.h"

namespace ppr {

inline int hip_fail(hipError_t e, const char* where) {
  return ppc::set_error(PP_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                       \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ppr::hip_fail(_e, #expr);  \
  } while (0)

// splitmix64 stream -> uniform [-1, 1); the sketch matrix of accum.
inline uint64_t splitmix64(uint64_t& state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline void sketch_values(std::vector<double>& v, int64_t count, uint64_t seed) {
  v.resize(count);
  uint64_t st = seed;
  for (int64_t i = 0; i < count; ++i) v[i] = 2.0 * ((double)(splitmix64(st) >> 11) * 0x1.0p-53) - 1.0;
}

constexpr uint64_t kSketchSeed = 0x5EEDULL;

inline int elem_bytes(int precision) { return precision == PP_FP64 ? 8 : 4; }

}  // namespace ppr

struct pp_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  int cus = 0;
  hipStream_t stream = nullptr;
  std::unordered_map<void*, uint64_t> allocs;
  std::mutex mu;
};

struct pp_stream {
  pp_ctx* ctx = nullptr;
  int workers = 0, n_insts = 0, precision = 0, n_ptrs = 0;
  int32_t* d_words = nullptr;
  int64_t* d_off = nullptr;
  void** d_ptrs = nullptr;
  uint32_t* d_flags = nullptr;
  uint32_t* d_epoch = nullptr;
  int32_t* d_status = nullptr;
  void* d_scratch = nullptr;
  int64_t scratch_stride = 0;
  void* d_rnd = nullptr;
  uint64_t* d_trace = nullptr;
  int64_t* d_trace_off = nullptr;
  int64_t trace_total = 0;
  bool trace_on = false;
  std::vector<void*> ptrs;
  std::vector<std::pair<int, int64_t>> zero_slots;  // slot, elements
  std::vector<void*> owned;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t spin_limit = 20LL * 100000000LL;  // 20 s of 100 MHz ticks
  hipStream_t last = nullptr;
  bool launched = false;
  int role = 0;  // 1: a solver's factorization program (own kernel symbol)
  int last_fallbacks = 0;  // Householder fallbacks of accum in the last waited launch
};

namespace ppr {

int ctx_alloc(pp_ctx* c, uint64_t bytes, void** p);
int ctx_free(pp_ctx* c, void* p);

template <typename T>
int ctx_alloc_t(pp_ctx* c, int64_t count, T** p) {
  void* v = nullptr;
  int rc = ctx_alloc(c, (uint64_t)std::max<int64_t>(count, 1) * sizeof(T), &v);
  *p = static_cast<T*>(v);
  return rc;
}

inline hipStream_t pick(pp_ctx* c, void* s) { return s ? static_cast<hipStream_t>(s) : c->stream; }

// Checks one execute entry for the single-op kernel and packs its words and
// pointer table (at most kSingleMaxPtrs slots) by value.
int single_op_args(int precision, const int32_t* words, int n_words, const void* const* ptr_table, int n_ptrs,
                   pp::SingleOp* op);
// Sketch values the entry reads (nx * k1 for accum, else 1).
int64_t single_op_sketch(const int32_t* words);
// Resolves this device's allocation-free scratch (the device must be
// current); called once per device by pp_ctx_create.
int noalloc_init(int device);

// Interprets the 8 status words of a finished launch of s (records the
// accum fallbacks; on a VM error optionally finds the first unsignalled
// instruction) and returns PP_OK or the error.
int stream_status(pp_stream* s, const int32_t* st, int32_t* first_unsignalled);

// Enqueue the launch sequence of a stream (temp fills, epoch bump, VM) on st
// without waiting; usable under hipGraph capture.
int enqueue_launch(pp_ctx* ctx, pp_stream* s, int wavefronts_per_inst, hipStream_t st, bool record_events);

}  // namespace ppr
