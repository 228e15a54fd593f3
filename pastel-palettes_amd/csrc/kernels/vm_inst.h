// Per-precision entry points of the persistent VM, each defined in its own
// translation unit: vm_fp64.hip (FP64 storage and MFMA), vm_fp32.hip (FP32)
// and vm_mixed.hip (FP32 storage and leaf routines, BF16 products).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vm.h"

namespace pp {

#define PP_VM_DECLARE(name)                                                                                 \
  hipError_t vm_launch_##name(const VmArgs& a, int workers, int threads, bool factor, hipStream_t s);      \
  int vm_blocks_##name(int threads);                                                                        \
  hipError_t vm_single_##name(int threads, int reps, const SingleOp& op, void* scratch, const void* rnd,      \
                              int32_t* status, hipStream_t s);
PP_VM_DECLARE(fp64)
PP_VM_DECLARE(fp32)
PP_VM_DECLARE(mixed)
#undef PP_VM_DECLARE

}  // namespace pp
