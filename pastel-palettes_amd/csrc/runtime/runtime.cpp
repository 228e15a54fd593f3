// C ABI, device half: context, uploaded streams and VM launches.
#include "runtime_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <thread>

using namespace ppr;
using ppc::from_exception;
using ppc::set_error;

namespace ppr {

int ctx_alloc(pp_ctx* c, uint64_t bytes, void** p) {
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->budget != 0 && c->in_use + bytes > c->budget)
    return set_error(PP_ERR_NOMEM, "HBM budget exceeded: need " + std::to_string(bytes) + " bytes, " +
                                       std::to_string(c->budget - c->in_use) + " left");
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, bytes > 0 ? bytes : 256);
  if (e != hipSuccess) return set_error(PP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs[d] = bytes;
  c->in_use += bytes;
  *p = d;
  return PP_OK;
}

int ctx_free(pp_ctx* c, void* p) {
  if (p == nullptr) return PP_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->allocs.find(p);
  if (it == c->allocs.end()) return set_error(PP_ERR_INVALID, "pp_free: pointer not allocated by this context");
  c->in_use -= it->second;
  c->allocs.erase(it);
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return hip_fail(e, "hipFree");
  return PP_OK;
}

int enqueue_launch(pp_ctx* ctx, pp_stream* s, int wavefronts_per_inst, hipStream_t st, bool record_events) {
  const int threads = 64 * wavefronts_per_inst;
  if (threads != 256) return set_error(PP_ERR_INVALID, "wavefronts_per_inst must be 4");
  const int per_cu = pp::vm_blocks_per_cu(s->precision, threads);
  if (per_cu < 1 || s->workers > per_cu * ctx->cus)
    return set_error(PP_ERR_INVALID, "workers (" + std::to_string(s->workers) + ") exceed co-resident capacity " +
                                         std::to_string(per_cu * ctx->cus));
  for (const auto& z : s->zero_slots) {
    hipError_t e = pp::launch_fill_zero(s->ptrs[z.first], z.second, elem_bytes(s->precision), st);
    if (e != hipSuccess) return hip_fail(e, "fill_zero");
  }
  pp::VmArgs a{};
  a.words = s->d_words;
  a.stream_off = s->d_off;
  a.ptrs = s->d_ptrs;
  a.flags = s->d_flags;
  a.scratch = s->d_scratch;
  a.scratch_stride = s->scratch_stride;
  a.rnd = s->d_rnd;
  a.trace = s->trace_on ? s->d_trace : nullptr;
  a.trace_off = s->d_trace_off;
  a.status = s->d_status;
  a.spin_limit = s->spin_limit;
  a.epoch = s->d_epoch;
  a.n_insts = s->n_insts;
  if (record_events) HIP_TRY(hipEventRecord(s->ev0, st));
  hipError_t e = pp::launch_vm(a, s->precision, s->workers, threads, s->role == 1, s->d_epoch, s->d_status, st);
  if (e != hipSuccess) return hip_fail(e, "launch_vm");
  if (record_events) HIP_TRY(hipEventRecord(s->ev1, st));
  s->last = st;
  s->launched = true;
  return PP_OK;
}

int single_op_args(int precision, const int32_t* words, int n_words, const void* const* ptr_table, int n_ptrs,
                   pp::SingleOp* op) {
  if (words == nullptr || n_words < 3 || ptr_table == nullptr || n_ptrs < 1 || n_ptrs > pp::kSingleMaxPtrs ||
      precision < 0 || precision > 2)
    return set_error(PP_ERR_INVALID, "bad single-op args");
  const int len = operationLength(words[2]);
  if (words[0] != execute || len <= 0 || n_words < len || len > pp::kSingleMaxWords)
    return set_error(PP_ERR_INVALID, "bad instruction words");
  *op = pp::SingleOp{};
  for (int i = 0; i < len; ++i) op->words[i] = words[i];
  for (int i = 0; i < n_ptrs; ++i) op->ptrs[i] = const_cast<void*>(ptr_table[i]);
  return PP_OK;
}

int64_t single_op_sketch(const int32_t* words) {
  return words[2] == accum ? std::max<int64_t>((int64_t)words[11] * words[13], 1) : 1;
}

}  // namespace ppr

namespace {

int upload_sketch(pp_stream* s, int64_t count, uint64_t seed) {
  std::vector<double> v;
  sketch_values(v, std::max<int64_t>(count, 1), seed);
  int rc = ctx_alloc(s->ctx, v.size() * elem_bytes(s->precision), &s->d_rnd);
  if (rc) return rc;
  s->owned.push_back(s->d_rnd);
  if (s->precision == PP_FP64) {
    HIP_TRY(hipMemcpy(s->d_rnd, v.data(), v.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(v.begin(), v.end());
    HIP_TRY(hipMemcpy(s->d_rnd, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return PP_OK;
}

// Sketch size a program needs: max over accum ops of nx * k1.
int64_t sketch_need(const pp_program* p) {
  int64_t need = 1;
  for (int i = 0; i < p->tree.length(); ++i) {
    const h_op& o = p->tree.getOp(i);
    if (o.op == accum) need = std::max<int64_t>(need, (int64_t)o.words[11] * o.words[13]);
  }
  return need;
}

int stream_common(pp_ctx* ctx, pp_stream* s, const std::vector<const std::vector<int32_t>*>& streams, int n_insts,
                  uint64_t scratch_bytes_per_worker, int64_t sketch_count) {
  const int W = s->workers;
  std::vector<int64_t> off(W + 1, 0), toff(W + 1, 0);
  for (int w = 0; w < W; ++w) {
    off[w + 1] = off[w] + (int64_t)streams[w]->size();
    int64_t entries = 0;
    for (size_t i = 0; i < streams[w]->size();) {
      const int op = (*streams[w])[i];
      ++entries;
      if (op == finish) break;
      if (op == signal_wait) i += wait_l;
      else if (op == execute) {
        const int len = operationLength((*streams[w])[i + 2]);
        if (len <= 0) return set_error(PP_ERR_INVALID, "stream holds an unknown operation");
        i += len;
      } else {
        return set_error(PP_ERR_INVALID, "stream holds an unknown opcode");
      }
    }
    toff[w + 1] = toff[w] + entries + 1;
  }
  std::vector<int32_t> all;
  all.reserve(off[W]);
  for (int w = 0; w < W; ++w) all.insert(all.end(), streams[w]->begin(), streams[w]->end());
  int rc;
  if ((rc = ctx_alloc_t(ctx, off[W], &s->d_words))) return rc;
  s->owned.push_back(s->d_words);
  if ((rc = ctx_alloc_t(ctx, W, &s->d_off))) return rc;
  s->owned.push_back(s->d_off);
  if ((rc = ctx_alloc_t(ctx, std::max(n_insts, 1), &s->d_flags))) return rc;
  s->owned.push_back(s->d_flags);
  if ((rc = ctx_alloc_t(ctx, 16, &s->d_epoch))) return rc;
  s->owned.push_back(s->d_epoch);
  if ((rc = ctx_alloc_t(ctx, 16, &s->d_status))) return rc;
  s->owned.push_back(s->d_status);
  if ((rc = ctx_alloc_t(ctx, toff[W], &s->d_trace))) return rc;
  s->owned.push_back(s->d_trace);
  if ((rc = ctx_alloc_t(ctx, W, &s->d_trace_off))) return rc;
  s->owned.push_back(s->d_trace_off);
  s->trace_total = toff[W];
  const int eb = elem_bytes(s->precision);
  s->scratch_stride = (int64_t)((scratch_bytes_per_worker + eb - 1) / eb + 63) / 64 * 64;
  if ((rc = ctx_alloc(ctx, (uint64_t)std::max<int64_t>(s->scratch_stride, 64) * eb * W, &s->d_scratch))) return rc;
  s->owned.push_back(s->d_scratch);
  if ((rc = ctx_alloc_t(ctx, std::max(s->n_ptrs, 1), &s->d_ptrs))) return rc;
  s->owned.push_back(s->d_ptrs);
  HIP_TRY(hipMemcpy(s->d_words, all.data(), all.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_off, off.data(), W * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_trace_off, toff.data(), W * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(s->d_flags, 0, sizeof(uint32_t) * std::max(n_insts, 1)));
  HIP_TRY(hipMemset(s->d_epoch, 0, 64));
  HIP_TRY(hipMemset(s->d_status, 0, 64));
  HIP_TRY(hipEventCreate(&s->ev0));
  HIP_TRY(hipEventCreate(&s->ev1));
  return upload_sketch(s, sketch_count, kSketchSeed);
}

int push_ptrs(pp_stream* s) {
  HIP_TRY(hipMemcpy(s->d_ptrs, s->ptrs.data(), sizeof(void*) * s->ptrs.size(), hipMemcpyHostToDevice));
  return PP_OK;
}

void destroy_stream(pp_stream* s) {
  if (s == nullptr) return;
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  for (void* p : s->owned) (void)ctx_free(s->ctx, p);
  delete s;
}

}  // namespace

namespace ppr {

int stream_status(pp_stream* s, const int32_t* st, int32_t* first_unsignalled) {
  s->last_fallbacks = st[5];
  if (st[0] == pp::VM_OK) return PP_OK;
  if (first_unsignalled) {
    uint32_t epoch = 0;
    std::vector<uint32_t> flags(std::max(s->n_insts, 1));
    HIP_TRY(hipMemcpy(&epoch, s->d_epoch, 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(flags.data(), s->d_flags, 4 * flags.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < s->n_insts; ++i)
      if ((int32_t)(flags[i] - epoch) < 0) {
        *first_unsignalled = i;
        break;
      }
  }
  const char* what = st[0] == pp::VM_TIMEOUT ? "watchdog: wait timed out"
                     : st[0] == pp::VM_BAD_OPCODE ? "bad opcode"
                     : st[0] == pp::VM_BAD_OPERATION ? "bad operation"
                                                     : "aborted";
  return set_error(st[0] == pp::VM_TIMEOUT ? PP_ERR_TIMEOUT : PP_ERR_DEVICE,
                   std::string("VM ") + what + " (worker " + std::to_string(st[1]) + ", id " + std::to_string(st[2]) +
                       ")");
}

}  // namespace ppr

extern "C" {

int pp_ctx_create(int device, uint64_t budget_bytes, pp_ctx** out) {
  try {
    if (out == nullptr) return set_error(PP_ERR_INVALID, "null out");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(PP_ERR_INVALID, "no such device");
    HIP_TRY(hipSetDevice(device));
    auto c = std::make_unique<pp_ctx>();
    c->device = device;
    c->budget = budget_bytes;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    // the allocation-free entry point's scratch on this device (once per device)
    const int nrc = noalloc_init(device);
    if (nrc != PP_OK) {
      (void)hipStreamDestroy(c->stream);
      return nrc;
    }
    *out = c.release();
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_ctx_destroy(pp_ctx* ctx) {
  if (ctx == nullptr) return PP_OK;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& a : ctx->allocs) (void)hipFree(a.first);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PP_OK;
}

int pp_ctx_cu_count(const pp_ctx* ctx) { return ctx ? ctx->cus : PP_ERR_INVALID; }
uint64_t pp_ctx_bytes_in_use(const pp_ctx* ctx) { return ctx ? ctx->in_use : 0; }

int pp_alloc(pp_ctx* ctx, uint64_t bytes, void** dptr) {
  if (ctx == nullptr || dptr == nullptr) return set_error(PP_ERR_INVALID, "null");
  return ctx_alloc(ctx, bytes, dptr);
}

int pp_free(pp_ctx* ctx, void* dptr) {
  if (ctx == nullptr) return set_error(PP_ERR_INVALID, "null ctx");
  return ctx_free(ctx, dptr);
}

int pp_copy_h2d(pp_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (ctx == nullptr) return set_error(PP_ERR_INVALID, "null ctx");
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return PP_OK;
}

int pp_copy_d2h(pp_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (ctx == nullptr) return set_error(PP_ERR_INVALID, "null ctx");
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pp_stream_upload(pp_ctx* ctx, const int32_t* const* words, const int32_t* lengths, int workers,
                     const void* const* ptr_table, int n_ptrs, uint64_t scratch_bytes_per_worker, int n_insts,
                     int precision, pp_stream** out) {
  try {
    if (ctx == nullptr || words == nullptr || lengths == nullptr || workers < 1 || out == nullptr || n_ptrs < 0 ||
        precision < 0 || precision > 2)
      return set_error(PP_ERR_INVALID, "bad upload args");
    std::vector<std::vector<int32_t>> ws(workers);
    std::vector<const std::vector<int32_t>*> refs;
    int64_t need = 1;
    for (int w = 0; w < workers; ++w) {
      ws[w].assign(words[w], words[w] + lengths[w]);
      if (ws[w].empty() || ws[w].back() != finish) ws[w].push_back(finish);
      for (size_t i = 0; i + 2 < ws[w].size();) {
        if (ws[w][i] == execute && ws[w][i + 2] == accum && i + 13 < ws[w].size())
          need = std::max<int64_t>(need, (int64_t)ws[w][i + 11] * ws[w][i + 13]);
        if (ws[w][i] == signal_wait) i += wait_l;
        else if (ws[w][i] == execute && operationLength(ws[w][i + 2]) > 0) i += operationLength(ws[w][i + 2]);
        else break;
      }
      refs.push_back(&ws[w]);
    }
    auto s = new pp_stream();
    s->ctx = ctx;
    s->workers = workers;
    s->n_insts = n_insts;
    s->precision = precision;
    s->n_ptrs = n_ptrs;
    for (int i = 0; i < n_ptrs; ++i) s->ptrs.push_back(const_cast<void*>(ptr_table[i]));
    int rc = stream_common(ctx, s, refs, n_insts, scratch_bytes_per_worker, need);
    if (rc == PP_OK) rc = push_ptrs(s);
    if (rc != PP_OK) {
      destroy_stream(s);
      return rc;
    }
    *out = s;
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_stream_upload_program(pp_ctx* ctx, const pp_program* prog, const void* const* ptr_table, int n_ptrs,
                             int precision, pp_stream** out) {
  try {
    if (ctx == nullptr || prog == nullptr || out == nullptr || precision < 0 || precision > 2)
      return set_error(PP_ERR_INVALID, "bad upload args");
    if (n_ptrs < prog->n_slots) return set_error(PP_ERR_INVALID, "pointer table shorter than the program's slots");
    auto s = new pp_stream();
    s->ctx = ctx;
    s->workers = prog->workers;
    s->n_insts = ppc::program_insts(prog);
    s->precision = precision;
    s->n_ptrs = n_ptrs;
    for (int i = 0; i < n_ptrs; ++i) s->ptrs.push_back(const_cast<void*>(ptr_table[i]));
    std::vector<const std::vector<int32_t>*> refs;
    for (int w = 0; w < prog->workers; ++w) refs.push_back(&ppc::program_stream(prog, w));
    const int eb = elem_bytes(precision);
    int rc = stream_common(ctx, s, refs, s->n_insts, (uint64_t)prog->scratch * eb, sketch_need(prog));
    // temp slots the caller left empty are owned (and zeroed per launch) here
    for (size_t i = 0; rc == PP_OK && i < prog->temp_elems.size(); ++i) {
      const int slot = 3 * (int)i + 1;
      const int64_t n = prog->temp_elems[i];
      if (n <= 0) continue;
      if (s->ptrs[slot] == nullptr) {
        if (!prog->zero_temps) {
          rc = set_error(PP_ERR_INVALID, "a window after the first needs the temp buffer of the first window");
          break;
        }
        rc = ctx_alloc(ctx, (uint64_t)n * eb, &s->ptrs[slot]);
        if (rc == PP_OK) s->owned.push_back(s->ptrs[slot]);
      }
      if (prog->zero_temps) s->zero_slots.emplace_back(slot, n);
    }
    if (rc == PP_OK) rc = push_ptrs(s);
    if (rc != PP_OK) {
      destroy_stream(s);
      return rc;
    }
    *out = s;
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_stream_qr_fallbacks(const pp_stream* s) { return s ? s->last_fallbacks : set_error(PP_ERR_INVALID, "null"); }

int pp_stream_set_slot(pp_ctx* ctx, pp_stream* s, int slot, void* dptr) {
  if (ctx == nullptr || s == nullptr || slot < 0 || slot >= s->n_ptrs) return set_error(PP_ERR_INVALID, "bad slot");
  s->ptrs[slot] = dptr;
  return push_ptrs(s);
}

int pp_stream_set_watchdog(pp_ctx* ctx, pp_stream* s, int ms) {
  if (ctx == nullptr || s == nullptr || ms <= 0) return set_error(PP_ERR_INVALID, "bad watchdog");
  s->spin_limit = (int64_t)ms * 100000LL;
  return PP_OK;
}

int pp_stream_enable_trace(pp_ctx* ctx, pp_stream* s, int enable) {
  if (ctx == nullptr || s == nullptr) return set_error(PP_ERR_INVALID, "null");
  s->trace_on = enable != 0;
  return PP_OK;
}

int pp_max_workers(pp_ctx* ctx, int precision, int wavefronts_per_inst) {
  if (ctx == nullptr) return set_error(PP_ERR_INVALID, "null ctx");
  const int per_cu = pp::vm_blocks_per_cu(precision, 64 * wavefronts_per_inst);
  return per_cu * ctx->cus;
}

int pp_launch(pp_ctx* ctx, pp_stream* s, int wavefronts_per_inst, void* hip_stream) {
  try {
    if (ctx == nullptr || s == nullptr) return set_error(PP_ERR_INVALID, "null");
    return enqueue_launch(ctx, s, wavefronts_per_inst, pick(ctx, hip_stream), true);
  } catch (...) {
    return from_exception();
  }
}

int pp_wait(pp_ctx* ctx, pp_stream* s, int timeout_ms, int32_t* first_unsignalled) {
  if (ctx == nullptr || s == nullptr) return set_error(PP_ERR_INVALID, "null");
  if (first_unsignalled) *first_unsignalled = -1;
  if (!s->launched) return PP_OK;
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    hipError_t q = hipEventQuery(s->ev1);
    if (q == hipSuccess) break;
    if (q != hipErrorNotReady) return hip_fail(q, "hipEventQuery");
    const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (timeout_ms > 0 && el > timeout_ms) return set_error(PP_ERR_TIMEOUT, "host wait timed out");
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  int32_t st[8];
  HIP_TRY(hipMemcpy(st, s->d_status, sizeof(st), hipMemcpyDeviceToHost));
  return ppr::stream_status(s, st, first_unsignalled);
}

double pp_stream_last_ms(const pp_stream* s) {
  if (s == nullptr || !s->launched) return -1.;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, s->ev0, s->ev1) != hipSuccess) return -1.;
  return ms;
}

int64_t pp_read_trace(pp_ctx* ctx, pp_stream* s, uint64_t* out, int64_t cap) {
  if (ctx == nullptr || s == nullptr) return set_error(PP_ERR_INVALID, "null");
  if (out != nullptr && cap > 0) {
    hipError_t e = hipMemcpy(out, s->d_trace, sizeof(uint64_t) * std::min(cap, s->trace_total), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "trace copy");
  }
  return s->trace_total;
}

int pp_stream_destroy(pp_ctx* ctx, pp_stream* s) {
  if (ctx == nullptr) return set_error(PP_ERR_INVALID, "null ctx");
  if (s && s->launched) (void)hipEventSynchronize(s->ev1);
  destroy_stream(s);
  return PP_OK;
}

int pp_fill_sketch(pp_ctx* ctx, void* dptr, int64_t count, int precision, uint64_t seed) {
  if (ctx == nullptr || dptr == nullptr || count < 0) return set_error(PP_ERR_INVALID, "bad args");
  std::vector<double> v;
  sketch_values(v, count, seed);
  if (precision == PP_FP64) {
    HIP_TRY(hipMemcpy(dptr, v.data(), v.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(v.begin(), v.end());
    HIP_TRY(hipMemcpy(dptr, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return PP_OK;
}

int pp_exec_single_timed(pp_ctx* ctx, int precision, const int32_t* words, int n_words, const void* const* ptr_table,
                         int n_ptrs, uint64_t scratch_bytes, uint64_t seed, int wavefronts, int reps, double* ms) {
  pp::SingleOp op{};
  int rc = single_op_args(precision, words, n_words, ptr_table, n_ptrs, &op);
  if (ctx == nullptr || reps < 1 || wavefronts != 4) return set_error(PP_ERR_INVALID, "bad args");
  if (rc) return rc;
  void* d_s = nullptr;
  void* d_r = nullptr;
  int32_t* d_st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int eb = elem_bytes(precision);
  const int64_t need = single_op_sketch(words);
  do {
    if ((rc = ctx_alloc(ctx, std::max<uint64_t>(scratch_bytes, 256), &d_s))) break;
    if ((rc = ctx_alloc(ctx, need * eb, &d_r))) break;
    if ((rc = ctx_alloc_t(ctx, 4, &d_st))) break;
    hipError_t e = pp::launch_sketch_fill(d_r, need, eb, seed, ctx->stream);
    if (e == hipSuccess) e = pp::launch_clear_i32(d_st, 4, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    if (e == hipSuccess) e = pp::launch_single_op(precision, 64 * wavefronts, reps, op, d_s, d_r, d_st, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    int32_t st = 0;
    if (e == hipSuccess) e = hipMemcpy(&st, d_st, 4, hipMemcpyDeviceToHost);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e != hipSuccess) {
      rc = hip_fail(e, "exec_single");
      break;
    }
    if (ms) *ms = t / reps;
    if (st != operationLength(words[2])) rc = set_error(PP_ERR_DEVICE, "device rejected the instruction");
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  ctx_free(ctx, d_s);
  ctx_free(ctx, d_r);
  ctx_free(ctx, d_st);
  return rc;
}

int pp_exec_single(pp_ctx* ctx, int precision, const int32_t* words, int n_words, const void* const* ptr_table,
                   int n_ptrs, uint64_t scratch_bytes, uint64_t seed) {
  return pp_exec_single_timed(ctx, precision, words, n_words, ptr_table, n_ptrs, scratch_bytes, seed, 4, 1, nullptr);
}

int pp_mfma_probe(pp_ctx* ctx, int mode, const double* a, int na, const double* b, int nb, double* out) {
  if (ctx == nullptr || a == nullptr || b == nullptr || out == nullptr) return set_error(PP_ERR_INVALID, "null");
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  int rc = PP_OK;
  do {
    if ((rc = ctx_alloc_t(ctx, na, &da))) break;
    if ((rc = ctx_alloc_t(ctx, nb, &db))) break;
    if ((rc = ctx_alloc_t(ctx, 256, &dout))) break;
    hipError_t e = hipMemcpy(da, a, 8 * na, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, 8 * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = pp::launch_mfma_probe(mode, da, db, dout, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * 256, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "mfma_probe");
  } while (false);
  ctx_free(ctx, da);
  ctx_free(ctx, db);
  ctx_free(ctx, dout);
  return rc;
}

}  // extern "C"

