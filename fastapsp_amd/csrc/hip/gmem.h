// Address-space helpers.  Pointers read from descriptor tables are generic
// ("flat") to the compiler; casting them to the global address space gives
// global_load / global_store, and a wave-uniform pointer in the constant address
// space gives scalar (s_load) reads of read-only data.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define FAP_GLOBAL __attribute__((address_space(1)))
#define FAP_CONST __attribute__((address_space(4)))

namespace fap_hip {

template <typename T>
__device__ __forceinline__ FAP_GLOBAL T *gptr(T *p) {
  return (FAP_GLOBAL T *)p;
}

// Read-only, wave-uniform pointer -> constant address space (scalar loads).
template <typename T>
__device__ __forceinline__ const FAP_CONST T *uniform_cptr(const T *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const FAP_CONST T *)(((uint64_t)hi << 32) | lo);
}

// Workgroup barrier that orders LDS accesses only: s_waitcnt lgkmcnt(0) + s_barrier.
// __syncthreads() also waits for the wave's global stores to complete (vmcnt(0)),
// which a loop storing to global memory between barriers pays on every step.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace fap_hip
