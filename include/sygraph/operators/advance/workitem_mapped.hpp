/*
 * operators/advance/workitem_mapped.hpp -- one lane per source, arcs walked
 * serially (reference operators/advance/workitem_mapped.hpp:24-62, whose
 * template cannot be instantiated from advance.hpp:130), and the
 * subgroup_mapped balancer the reference declares but never implements
 * (config.hpp:12): one wavefront per source.
 */
#pragma once

#include <sygraph/operators/advance/common.hpp>

namespace sygraph {
namespace operators {
namespace advance {
namespace detail {

template<typename G, typename Out, typename F>
__global__ void __launch_bounds__(kBlock) push_workitem_kernel(G g, SourceList src, Out out, F f) {
  using edge_t = typename G::edge_t;
  const uint32_t count = src.size();
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
    const uint32_t u = src.at(i);
    const edge_t b = g.getFirstNeighbor(u);
    const uint64_t d = g.getDegree(u);
    for (uint64_t j = 0; j < d; ++j) {
      const edge_t e = b + static_cast<edge_t>(j);
      const uint32_t dst = static_cast<uint32_t>(g.getDestinationVertex(e));
      if (call_functor(f, g, u, dst, e)) out.insert(dst);
    }
  }
}

template<typename G, typename Out, typename F>
__global__ void __launch_bounds__(kBlock) push_subgroup_kernel(G g, SourceList src, Out out, F f) {
  using edge_t = typename G::edge_t;
  const uint32_t count = src.size();
  const int lane = sygraph::hipdev::lane_id();
  const uint32_t gwave = (blockIdx.x * kBlock + threadIdx.x) >> 6, nwaves = (gridDim.x * kBlock) >> 6;
  for (uint32_t i = gwave; i < count; i += nwaves) {
    const uint32_t u = src.at(i);
    const edge_t b = g.getFirstNeighbor(u);
    const uint64_t d = g.getDegree(u);
    for (uint64_t j = lane; j < d; j += 64) {
      const edge_t e = b + static_cast<edge_t>(j);
      const uint32_t dst = static_cast<uint32_t>(g.getDestinationVertex(e));
      if (call_functor(f, g, u, dst, e)) out.insert(dst);
    }
  }
}

/* Pull over the inverse CSR: one lane per candidate vertex (bit of an
 * input-frontier word), in-arcs in order; stops at the first success when
 * kShort.  Vertices already in the input frontier are skipped. */
template<bool kShort, typename G, typename InDev, typename Out, typename F>
__global__ void __launch_bounds__(kBlock) pull_kernel(G ig, InDev in, uint32_t n, Out out, F f) {
  using edge_t = typename G::edge_t;
  const uint32_t words = (n + 63) / 64;
  const int lane = sygraph::hipdev::lane_id();
  const uint32_t gwave = (blockIdx.x * kBlock + threadIdx.x) >> 6, nwaves = (gridDim.x * kBlock) >> 6;
  for (uint32_t w = gwave; w < words; w += nwaves) {
    const uint64_t word = in.getData()[w];
    if (word == ~0ull) continue;
    const uint32_t v = w * 64 + lane;
    if (v >= n || ((word >> lane) & 1ull)) continue;
    const edge_t b = ig.getFirstNeighbor(v);
    const uint64_t d = ig.getDegree(v);
    for (uint64_t j = 0; j < d; ++j) {
      const edge_t e = b + static_cast<edge_t>(j);
      const uint32_t u = static_cast<uint32_t>(ig.getDestinationVertex(e));
      if (!in.check(u)) continue;
      if (call_functor(f, ig, v, u, e)) {
        out.insert(v);
        if constexpr (kShort) break;
      }
    }
  }
}

} // namespace detail
} // namespace advance
} // namespace operators
} // namespace sygraph
