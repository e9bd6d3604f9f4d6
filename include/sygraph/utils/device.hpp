/*
 * utils/device.hpp -- device queries (reference utils/device.hpp:16-34).
 */
#pragma once

#include <sycl/sycl.hpp>

namespace sygraph {
namespace detail {
namespace device {

inline uint32_t getMaxNumSubgroups(const sycl::queue& queue) {
  return queue.get_device().get_info<sycl::info::device::max_num_sub_groups>();
}

inline uint32_t getSubgroupSize(const sycl::queue& queue) {
  return static_cast<uint32_t>(queue.get_device().get_info<sycl::info::device::sub_group_sizes>()[0]);
}

inline uint32_t getWorkgroupSize(const sycl::queue& queue) {
  return static_cast<uint32_t>(queue.get_device().get_info<sycl::info::device::max_work_group_size>());
}

inline uint32_t getNumComputeUnits(const sycl::queue& queue) {
  return queue.get_device().get_info<sycl::info::device::max_compute_units>();
}

} // namespace device
} // namespace detail
} // namespace sygraph
