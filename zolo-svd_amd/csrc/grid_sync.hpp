// Barrier among the workgroups of one cooperative launch (or of one y-slice
// of it), for kernels that keep their data in registers across many steps.
#pragma once
#include <hip/hip_runtime.h>

namespace zolo {

// All workgroups sharing ctr must be resident (cooperative launch).  The
// counter only grows: arrival k of every workgroup waits until it reads
// k * (number of workgroups).  The arrival is a device-scope release (it
// publishes what the workgroup wrote before the barrier); the spin reads the
// counter relaxed and one device-scope acquire after it makes the other
// workgroups' writes visible.
__device__ inline void grid_barrier(unsigned* ctr, unsigned target) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

}  // namespace zolo
