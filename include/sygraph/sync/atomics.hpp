/*
 * sync/atomics.hpp -- relaxed, device-scope atomics for user functors
 * (reference sync/atomics.hpp:24-100).  min returns the previous value, cas
 * updates `expected` on failure, like atomic_ref.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace sygraph {
namespace sync {

template<typename T>
__host__ __device__ inline T atomicFetchAdd(T* ptr, T val) {
  return __hip_atomic_fetch_add(ptr, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__host__ __device__ inline T load(T* ptr) {
  return __hip_atomic_load(ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__host__ __device__ inline void store(T* ptr, T val) {
  __hip_atomic_store(ptr, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__host__ __device__ inline T min(T* v1, T* v2) {
  const T x = load(v2);
  if constexpr (std::is_floating_point_v<T>) {
    T old = load(v1);
    while (x < old && !__hip_atomic_compare_exchange_strong(v1, &old, x, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
    }
    return old;
  } else {
    return __hip_atomic_fetch_min(v1, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template<typename T>
__host__ __device__ inline T max(T* v1, T* v2) {
  const T x = load(v2);
  if constexpr (std::is_floating_point_v<T>) {
    T old = load(v1);
    while (old < x && !__hip_atomic_compare_exchange_strong(v1, &old, x, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
    }
    return old;
  } else {
    return __hip_atomic_fetch_max(v1, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template<typename T>
__host__ __device__ inline bool cas(T* ptr, T& expected, T desired) {
  return __hip_atomic_compare_exchange_strong(ptr, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}

} // namespace sync
} // namespace sygraph
