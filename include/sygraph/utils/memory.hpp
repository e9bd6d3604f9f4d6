/*
 * utils/memory.hpp -- memory spaces and allocation (reference
 * utils/memory.hpp:12-48), backed by sg_alloc / sg_free.
 */
#pragma once

#include <stdexcept>
#include <sycl/sycl.hpp>

namespace sygraph {
namespace memory {

enum class space { host, device, shared };

namespace detail {

template<typename T, space V>
inline T* memoryAlloc(size_t n, sycl::queue& q) {
  if constexpr (V == space::host) {
    return sycl::malloc_host<T>(n, q);
  } else if constexpr (V == space::device) {
    return sycl::malloc_device<T>(n, q);
  } else {
    T* ptr = sycl::malloc_shared<T>(n, q);
#ifdef ENABLE_PREFETCH
    q.prefetch(ptr, n * sizeof(T)).wait(); /* utils/memory.hpp:24-26 */
#endif
    return ptr;
  }
}

template<typename T>
inline T* memoryAlloc(size_t n, sycl::queue& q, space s) {
  if (s == space::host) return memoryAlloc<T, space::host>(n, q);
  if (s == space::device) return memoryAlloc<T, space::device>(n, q);
  if (s == space::shared) return memoryAlloc<T, space::shared>(n, q);
  throw std::runtime_error("Unknown memory space");
}

template<typename T>
inline void releaseUSM(T*& ptr, sycl::queue& q) {
  if (ptr == nullptr) return;
  sycl::free(const_cast<std::remove_const_t<T>*>(ptr), q);
  ptr = nullptr;
}

} // namespace detail
} // namespace memory
} // namespace sygraph
