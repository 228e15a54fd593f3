// Partial-pivoted dense LU and row permutation (C ABI).  Restates the
// reference's pivot helpers (src/kernel.cu:958-1051), which no opcode of the
// reference dispatches: a pivoting variant for diagonal blocks that are not
// diagonally dominant.
#include <vector>

#include "runtime_internal.h"

using namespace ppr;
using ppc::from_exception;
using ppc::set_error;

extern "C" {

int pp_dense_getrf_pivoted(pp_ctx* ctx, int precision, void* a, int ld, int ny, int nx, int count, int64_t stride,
                           int32_t* pivots, int32_t* info, void* hip_stream) {
  try {
    if (ctx == nullptr || a == nullptr || pivots == nullptr || info == nullptr)
      return set_error(PP_ERR_INVALID, "pp_dense_getrf_pivoted: null argument");
    if (precision != PP_FP64 && precision != PP_FP32)
      return set_error(PP_ERR_INVALID, "pp_dense_getrf_pivoted: precision must be PP_FP64 or PP_FP32");
    if (ny < 1 || nx < 1 || ld < nx || count < 1 || (count > 1 && stride < (int64_t)(ny - 1) * ld + nx))
      return set_error(PP_ERR_INVALID, "pp_dense_getrf_pivoted: bad shape, pitch, count or stride");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(pp::launch_getrf_pivoted(precision == PP_FP64, a, ld, ny, nx, count, stride, pivots, info,
                                     pick(ctx, hip_stream)));
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

int pp_dense_apply_pivot(pp_ctx* ctx, int precision, void* m, int ld, int nx, int ny, const int32_t* pivots,
                         int recover, void* hip_stream) {
  try {
    if (ctx == nullptr || m == nullptr || pivots == nullptr)
      return set_error(PP_ERR_INVALID, "pp_dense_apply_pivot: null argument");
    if (precision != PP_FP64 && precision != PP_FP32)
      return set_error(PP_ERR_INVALID, "pp_dense_apply_pivot: precision must be PP_FP64 or PP_FP32");
    if (ny < 1 || nx < 1 || ld < nx)
      return set_error(PP_ERR_INVALID, "pp_dense_apply_pivot: bad shape or pitch");
    if (pp::apply_pivot_chunk(precision == PP_FP64, ny) < 1)
      return set_error(PP_ERR_INVALID, "pp_dense_apply_pivot: more rows than one LDS column chunk holds");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = pick(ctx, hip_stream);
    // the permutation is checked before any row moves: an entry out of range
    // or repeated fails the call instead of reading outside M
    std::vector<int32_t> h(ny);
    HIP_TRY(hipMemcpyAsync(h.data(), pivots, sizeof(int32_t) * ny, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<char> seen(ny, 0);
    for (int r = 0; r < ny; ++r) {
      if (h[r] < 0 || h[r] >= ny || seen[h[r]])
        return set_error(PP_ERR_INVALID, "pp_dense_apply_pivot: pivots is not a permutation of [0, ny)");
      seen[h[r]] = 1;
    }
    HIP_TRY(pp::launch_apply_pivot(precision == PP_FP64, m, ld, nx, ny, pivots, recover, st));
    return PP_OK;
  } catch (...) {
    return from_exception();
  }
}

}  // extern "C"
