/*
 * hip_device.hpp -- wave64 / CDNA4 device primitives shared by the C-ABI
 * library (sygraph_amd/csrc) and the functor-templated operator kernels.
 *
 * gfx950: 64-lane wavefronts, 64-bit ballots, relaxed agent-scope atomics on
 * coarse-grained HBM.  One bitmap word (uint64) == one wavefront: lane i owns
 * bit i, so a wave reads or writes a whole word with one ballot.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sygraph::hipdev {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

/* Number of set bits of `mask` strictly below this lane. */
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template<typename T>
__device__ __forceinline__ T shfl(T v, int src) {
  return __shfl(v, src, kWave);
}

/* Wave-wide sum / max of a 64-bit value (butterfly over 64 lanes). */
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint64_t t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}

/* ---- relaxed device(agent)-scope atomics (sync/atomics.hpp semantics) ---- */
template<typename T, typename U>
__device__ __forceinline__ T atomic_or(T* p, U v) {
  return __hip_atomic_fetch_or(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T, typename U>
__device__ __forceinline__ T atomic_and(T* p, U v) {
  return __hip_atomic_fetch_and(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T, typename U>
__device__ __forceinline__ T atomic_add(T* p, U v) {
  return __hip_atomic_fetch_add(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T, typename U>
__device__ __forceinline__ T atomic_min(T* p, U v) {
  return __hip_atomic_fetch_min(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T, typename U>
__device__ __forceinline__ T atomic_max(T* p, U v) {
  return __hip_atomic_fetch_max(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T>
__device__ __forceinline__ T atomic_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T, typename U>
__device__ __forceinline__ void atomic_store(T* p, U v) {
  __hip_atomic_store(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T>
__device__ __forceinline__ bool atomic_cas(T* p, T& expected, T desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}

/* Wave-aggregated append: every lane with `want` gets a unique slot in
 * [*counter, *counter + popcount), one atomic per wave. Returns the slot
 * (undefined for lanes without `want`). Must be called by the whole wave. */
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
  uint64_t m = ballot(want);
  if (m == 0) return 0;
  int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomic_add(counter, static_cast<uint32_t>(__popcll(m)));
  base = shfl(base, leader);
  return base + lanes_below(m);
}

__device__ __forceinline__ bool test_bit(const uint64_t* bm, uint32_t v) { return (bm[v >> 6] >> (v & 63)) & 1ull; }

/* Test-then-set of bit v; returns true iff this call flipped it from 0 to 1. */
__device__ __forceinline__ bool claim_bit(uint64_t* bm, uint32_t v) {
  const uint64_t b = 1ull << (v & 63);
  if (bm[v >> 6] & b) return false;
  return (atomic_or(&bm[v >> 6], b) & b) == 0;
}

} // namespace sygraph::hipdev
