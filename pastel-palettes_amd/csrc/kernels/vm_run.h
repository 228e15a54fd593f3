// The persistent VM interpreter (reference kernel_dynamic,
// src/kernel.cu:1375-1526) and its launch helpers, as templates over the
// storage type and the precision modes of the LU routines (FMODE) and of the
// products (GMODE).  Each precision mode is instantiated in a translation
// unit of its own (PP_VM_INSTANTIATE below) so that the three large kernels
// compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_steps.h"
#include "kernels.h"
#include "vm.h"
#include "vm_inst.h"

namespace pp {

static __device__ void record_error(int32_t* status, int32_t code, int32_t worker, int32_t inst) {
  if (atomicCAS(&status[0], VM_OK, code) == VM_OK) {
    status[1] = worker;
    status[2] = inst;
  }
  __hip_atomic_store(&status[3], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int FMODE, int GMODE>
__device__ __forceinline__ void vm_run(const VmArgs& a) {
  __shared__ int32_t ctl[2];
  const int w = blockIdx.x;
  const int32_t* pc = a.words + a.stream_off[w];
  T* tmp = reinterpret_cast<T*>(a.scratch) + (int64_t)w * a.scratch_stride;
  uint64_t* trace = a.trace ? a.trace + a.trace_off[w] : nullptr;
  const uint32_t epoch = __builtin_amdgcn_readfirstlane(*a.epoch);
  int entry = 0;
  for (;; ++entry) {
    const int32_t op = word(pc, 0);
    if (trace != nullptr && threadIdx.x == 0) trace[entry] = wall_clock64();
    if (op == finish) break;
    if (op == signal_wait) {
      const int32_t id = word(pc, 1);
      const int hint = word(pc, 2);
      if (threadIdx.x == 0) {
        int ok = 1;
        const uint64_t t0 = wall_clock64();
        int64_t polls = 0;
        // Poll with relaxed agent-scope loads (they read the coherent level
        // without invalidating caches); one acquire fence after the flag is
        // seen makes the producer's data visible.  An acquire load per poll
        // would invalidate this XCD's L2 on every iteration.
        while ((int32_t)(__hip_atomic_load(&a.flags[id], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - epoch) < 0) {
          if ((++polls & 31) == 0) {
            if (__hip_atomic_load(&a.status[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
              ok = 0;
              break;
            }
            if ((int64_t)(wall_clock64() - t0) > a.spin_limit) {
              record_error(a.status, VM_TIMEOUT, w, id);
              ok = 0;
              break;
            }
          }
          if (hint > 8) __builtin_amdgcn_s_sleep(8);
          else __builtin_amdgcn_s_sleep(1);
        }
        // one acquire for the workgroup: invalidates this CU's L1 and the
        // XCD's L2 copies so the producer's written-back data is read fresh
        if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ctl[entry & 1] = ok;
      }
      __syncthreads();
      if (!ctl[entry & 1]) break;
      pc += wait_l;
      continue;
    }
    if (op != execute) {
      if (threadIdx.x == 0) record_error(a.status, VM_BAD_OPCODE, w, op);
      break;
    }
    const int32_t id = word(pc, 1);
    const int len = exec_entry<T, FMODE, GMODE>(pc, a.ptrs, a.rnd, tmp, a.status + 5);
    if (len < 0) {
      if (threadIdx.x == 0) record_error(a.status, VM_BAD_OPERATION, w, id);
      break;
    }
    // every wave waits until its stores are acknowledged by L2; after the
    // barrier one release (L2 write-back) covers the whole workgroup
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(&a.flags[id], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    pc += len;
  }
}

template <typename T, int FMODE, int GMODE, int NT>
__global__ __launch_bounds__(NT) void vm_kernel(VmArgs a) {
  vm_run<T, FMODE, GMODE>(a);
}

// The same interpreter under a symbol of its own, launched for a solver's
// factorization program, so that kernel profiles list factorization and
// solve launches separately.
template <typename T, int FMODE, int GMODE, int NT>
__global__ __launch_bounds__(NT) void vm_factor_kernel(VmArgs a) {
  vm_run<T, FMODE, GMODE>(a);
}

template <typename T, int FMODE, int GMODE, int NT>
static hipError_t launch_vm_t(const VmArgs& a, int workers, bool factor, hipStream_t s) {
  if (factor) hipLaunchKernelGGL((vm_factor_kernel<T, FMODE, GMODE, NT>), dim3(workers), dim3(NT), 0, s, a);
  else hipLaunchKernelGGL((vm_kernel<T, FMODE, GMODE, NT>), dim3(workers), dim3(NT), 0, s, a);
  return hipGetLastError();
}

template <typename T, int FMODE, int GMODE>
static hipError_t launch_vm_nt(const VmArgs& a, int workers, int nt, bool factor, hipStream_t s) {
  switch (nt) {
    case 256: return launch_vm_t<T, FMODE, GMODE, 256>(a, workers, factor, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int FMODE, int GMODE, int NT>
static int max_blocks_t() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, vm_kernel<T, FMODE, GMODE, NT>, NT, 0) != hipSuccess) return 0;
  return nb;
}

// Single-op harness: one workgroup runs one execute entry `reps` times
// (tests use reps = 1; latency measurements use more, in place).  The words
// and the pointer table arrive by value and are read from LDS.
template <typename T, int FMODE, int GMODE, int NT>
__global__ __launch_bounds__(NT) void single_op_kernel(SingleOp op, void* scratch, const void* rnd, int32_t* status,
                                                       int reps) {
  __shared__ int32_t w[kSingleMaxWords];
  __shared__ void* p[kSingleMaxPtrs];
  if (threadIdx.x < kSingleMaxWords) w[threadIdx.x] = op.words[threadIdx.x];
  if (threadIdx.x < kSingleMaxPtrs) p[threadIdx.x] = op.ptrs[threadIdx.x];
  __syncthreads();
  int len = -1;
  for (int r = 0; r < reps; ++r) {
    len = exec_entry<T, FMODE, GMODE>(w, p, rnd, reinterpret_cast<T*>(scratch));
    __syncthreads();
  }
  if (threadIdx.x == 0) status[0] = len;
}

template <typename T, int FMODE, int GMODE>
static void launch_single_t(int threads, const SingleOp& op, void* scratch, const void* rnd, int32_t* status, int reps,
                            hipStream_t s) {
  (void)threads;  // one workgroup shape: 256 threads (4 wavefronts)
  hipLaunchKernelGGL((single_op_kernel<T, FMODE, GMODE, 256>), dim3(1), dim3(256), 0, s, op, scratch, rnd, status,
                     reps);
}

}  // namespace pp

// Defines the vm_*_<name> entry points of vm_inst.h for one precision mode.
#define PP_VM_INSTANTIATE(name, T, FMODE, GMODE)                                                              \
  namespace pp {                                                                                              \
  hipError_t vm_launch_##name(const VmArgs& a, int workers, int threads, bool factor, hipStream_t s) {        \
    return launch_vm_nt<T, FMODE, GMODE>(a, workers, threads, factor, s);                                     \
  }                                                                                                           \
  int vm_blocks_##name(int threads) { return threads == 256 ? max_blocks_t<T, FMODE, GMODE, 256>() : 0; }     \
  hipError_t vm_single_##name(int threads, int reps, const SingleOp& op, void* scratch, const void* rnd,        \
                              int32_t* status, hipStream_t s) {                                               \
    launch_single_t<T, FMODE, GMODE>(threads, op, scratch, rnd, status, reps, s);                             \
    return hipGetLastError();                                                                                 \
  }                                                                                                           \
  }
