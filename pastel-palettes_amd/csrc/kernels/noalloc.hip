// The allocation-free variant of the single-op entry point (pp_exec_single):
// its worker scratch, sketch values and status word live in one file-scope
// device array, so a call makes no hipMalloc / hipFree and no other
// allocating call.  Each device holds its own copy of the array, at its own
// address, resolved once per device when a context is created on it
// (pp_ctx_create -> noalloc_init) and kept in a per-device table; host code
// uses only those addresses, never the array's name.  Calls on one device
// share its array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "../runtime/runtime_internal.h"
#include "kernels.h"

// Maximum bytes of one call's scratch (worker scratch, then sketch values,
// then the status words, each part 256-byte aligned).
static constexpr uint64_t kNoAllocBytes = 4ull << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

uint64_t align256(uint64_t x) { return (x + 255) / 256 * 256; }

}  // namespace

int ppr::noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return ppc::set_error(PP_ERR_INVALID, "device index beyond the table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_addr[device] != nullptr) return PP_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    return ppc::set_error(PP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_addr[device] = p;
  return PP_OK;
}

extern "C" {

uint64_t pp_exec_single_noalloc_max_bytes(void) { return kNoAllocBytes; }

int pp_exec_single_noalloc(pp_ctx* ctx, int precision, const int32_t* words, int n_words, const void* const* ptr_table,
                           int n_ptrs, uint64_t scratch_bytes, uint64_t seed, void* hip_stream) {
  try {
    if (ctx == nullptr) return ppc::set_error(PP_ERR_INVALID, "null ctx");
    pp::SingleOp op{};
    int rc = ppr::single_op_args(precision, words, n_words, ptr_table, n_ptrs, &op);
    if (rc) return rc;
    void* base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_addr[ctx->device] : nullptr;
    }
    if (base == nullptr) return ppc::set_error(PP_ERR_INVALID, "no allocation-free scratch on this device");
    const int eb = ppr::elem_bytes(precision);
    const int64_t need = ppr::single_op_sketch(words);
    const uint64_t off_rnd = align256(std::max<uint64_t>(scratch_bytes, 256));
    const uint64_t off_st = off_rnd + align256((uint64_t)need * eb);
    if (off_st + 256 > kNoAllocBytes)
      return ppc::set_error(PP_ERR_INVALID, "allocation-free call needs " + std::to_string(off_st + 256) +
                                                " bytes of scratch, above its maximum of " +
                                                std::to_string(kNoAllocBytes));
    unsigned char* b = static_cast<unsigned char*>(base);
    int32_t* st = reinterpret_cast<int32_t*>(b + off_st);
    hipStream_t s = ppr::pick(ctx, hip_stream);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = pp::launch_sketch_fill(b + off_rnd, need, eb, seed, s);
    if (e == hipSuccess) e = pp::launch_clear_i32(st, 4, s);
    if (e == hipSuccess) e = pp::launch_single_op(precision, 256, 1, op, b, b + off_rnd, st, s);
    int32_t status = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&status, st, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ppr::hip_fail(e, "exec_single_noalloc");
    if (status != operationLength(words[2])) return ppc::set_error(PP_ERR_DEVICE, "device rejected the instruction");
    return PP_OK;
  } catch (...) {
    return ppc::from_exception();
  }
}

}  // extern "C"
