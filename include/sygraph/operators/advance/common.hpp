/*
 * operators/advance/common.hpp -- shared launch logic and per-arc semantics of
 * the advance operator (reference operators/advance/common.hpp:23-158).
 *
 * Semantics kept from the reference (common.hpp:49-75):
 *  - push: every out-arc (u, v) of a frontier vertex u; if functor(u, v, e, w)
 *    holds, v is inserted into the output frontier;
 *  - pull / pull_all: every vertex v NOT in the input frontier scans its
 *    in-arcs (inverse CSR) in order; an in-neighbour u qualifies if it is in
 *    the input frontier; if functor(v, u, e, w) holds, v is inserted into the
 *    output frontier, and `pull` stops at that first success;
 *  - graph view: every vertex of the graph is a source, push semantics.
 * The input frontier is compacted on the device first (ordered scan), so all
 * lanes of a push kernel hold real sources and the count never needs to
 * reach the host unless the caller asks for fetch_from_memory.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "sygraph_hip.h"
#include <sycl/sycl.hpp>
#include <sygraph/detail/hip_device.hpp>
#include <sygraph/frontier/frontier.hpp>
#include <sygraph/graph/graph.hpp>
#include <sygraph/operators/config.hpp>
#include <sygraph/sycl/event.hpp>
#include <sygraph/utils/device.hpp>
#include <sygraph/utils/types.hpp>

namespace sygraph {
namespace operators {
namespace advance {
namespace detail {

constexpr int kBlock = 256;

/* Source of a push: either a compacted element list with a device-side count,
 * or the whole vertex range [0, n) (graph view). */
struct SourceList {
  const int32_t* list;    /* nullptr: identity */
  const uint32_t* count;  /* device count when list != nullptr */
  uint32_t n;             /* count for the identity source */
  __device__ inline uint32_t size() const { return list ? *count : n; }
  __device__ inline uint32_t at(uint32_t i) const { return list ? (uint32_t)list[i] : i; }
};

/* Output frontier wrapper: insert when the view is a vertex frontier. */
template<typename OutDev, bool kInsert>
struct OutSink {
  OutDev dev;
  __device__ inline void insert(uint32_t v) const {
    if constexpr (kInsert) dev.insert(v);
  }
};

template<typename F, typename G>
__device__ inline bool call_functor(F& f, const G& g, uint32_t src, uint32_t dst, typename G::edge_t e) {
  return f(static_cast<typename G::vertex_t>(src), static_cast<typename G::vertex_t>(dst), e, g.getEdgeWeight(e));
}

/* Grid sizing for a push over a compacted frontier (buildAdvanceLaunchConfig,
 * common.hpp:91-135). */
inline unsigned push_grid(sycl::queue& q, int expected_size, const uint32_t* dev_count) {
  const unsigned cus = sygraph::detail::device::getNumComputeUnits(q);
  const unsigned persistent = cus * 8;
  if (expected_size == frontier::size::fetch_from_memory) {
    uint32_t count = 0;
    if (sg_memcpy(q.handle(), &count, dev_count, sizeof(count)) != SG_OK) throw std::runtime_error(sg_last_error());
    const unsigned need = (count + kBlock - 1) / kBlock;
    return std::max(1u, std::min(need, persistent));
  }
  if (expected_size == frontier::size::infer_from_device) return persistent;
  if (expected_size > 0) {
    const unsigned need = ((unsigned)expected_size + kBlock - 1) / kBlock;
    return std::max(1u, std::min(need, persistent));
  }
  throw std::runtime_error("Invalid expected_size value");
}

inline void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
}

template<typename FrontierT>
inline const uint32_t* compact_sources(const FrontierT& in) {
  if (sg_frontier_compact_async(in.handle(), 0) != SG_OK) throw std::runtime_error(sg_last_error());
  return in.getDeviceFrontier().getOffsetsSize();
}

} // namespace detail
} // namespace advance
} // namespace operators
} // namespace sygraph
