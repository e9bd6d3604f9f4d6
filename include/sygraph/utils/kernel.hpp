/*
 * utils/kernel.hpp -- launch-size helpers (reference utils/kernel.hpp:10-25).
 */
#pragma once

#include <algorithm>
#include <cstddef>

namespace sygraph {
namespace detail {
namespace kernel {

struct LaunchConfig {
  size_t global = 0; /* work items */
  size_t local = 0;  /* workgroup size */
};

inline size_t roundUpToMultiple(size_t value, size_t factor) {
  if (factor == 0) return value;
  return ((value + factor - 1) / factor) * factor;
}

inline size_t ensureLocalMultiple(size_t requested, size_t local_size) {
  if (requested == 0) return local_size;
  return std::max(local_size, roundUpToMultiple(requested, local_size));
}

} // namespace kernel
} // namespace detail
} // namespace sygraph
