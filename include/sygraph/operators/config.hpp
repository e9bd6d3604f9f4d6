/*
 * operators/config.hpp -- operator enums (same names and meaning as the
 * reference's include/sygraph/operators/config.hpp:10-37).
 */
#pragma once

namespace sygraph {
namespace operators {

/* Work distribution of an advance.
 *  workitem_mapped : one lane per frontier vertex, arcs walked serially
 *  subgroup_mapped : one wavefront per frontier vertex
 *  workgroup_mapped: edge-balanced merge path over each 256-vertex chunk
 *  bucketing       : lane / wavefront / workgroup bins by degree */
enum class load_balancer {
  workitem_mapped,
  subgroup_mapped,
  workgroup_mapped,
  bucketing,
};

enum class direction {
  push,     // out-arcs of frontier vertices
  pull,     // in-arcs of non-frontier vertices, stop at the first success
  pull_all, // in-arcs of non-frontier vertices, all of them
};

template<direction D>
constexpr bool is_pull() {
  return D == direction::pull || D == direction::pull_all;
}

template<direction D>
constexpr bool is_short_circuit() {
  return D == direction::pull;
}
} // namespace operators

namespace frontier::size {

using frontier_size_t = int;

constexpr frontier_size_t fetch_from_memory = -1; // read the active count back to the host and size the grid to it
constexpr frontier_size_t infer_from_device = 0;  // device-sized grid; the count never leaves the device

} // namespace frontier::size

} // namespace sygraph
