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
 * kShort.  Vertices already in the input frontier are skipped.  In-lists
 * longer than kPullCoop arcs are walked by the whole wavefront, 64 arcs per
 * step in order, and with kShort the wave stops after the first step with a
 * success (the sub-group bin of operators/advance/bucketing.hpp:163-178 does
 * the same), so a hub does not serialize one lane. */
constexpr uint64_t kPullCoop = 64;

template<bool kShort, typename G, typename InDev, typename Out, typename F>
__global__ void __launch_bounds__(kBlock) pull_kernel(G ig, InDev in, uint32_t n, Out out, F f) {
  using edge_t = typename G::edge_t;
  const uint32_t words = (n + 63) / 64;
  const int lane = sygraph::hipdev::lane_id();
  const uint32_t gwave = (blockIdx.x * kBlock + threadIdx.x) >> 6, nwaves = (gridDim.x * kBlock) >> 6;
  for (uint32_t w = gwave; w < words; w += nwaves) {
    const uint64_t word = in.getData()[w];
    if (word == ~0ull) continue; /* the same word in every lane */
    const uint32_t v = w * 64 + lane;
    const bool cand = v < n && !((word >> lane) & 1ull);
    edge_t b = 0;
    uint64_t d = 0;
    if (cand) {
      b = ig.getFirstNeighbor(v);
      d = ig.getDegree(v);
    }
    const bool coop = d > kPullCoop;
    if (cand && !coop) {
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
    uint64_t heavy = sygraph::hipdev::ballot(coop);
    while (heavy) {
      const int k = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
      heavy &= heavy - 1;
      const uint32_t vk = w * 64 + static_cast<uint32_t>(k);
      const edge_t bk = sygraph::hipdev::shfl(b, k);
      const uint64_t dk = sygraph::hipdev::shfl(d, k);
      for (uint64_t j0 = 0; j0 < dk; j0 += 64) {
        const uint64_t j = j0 + static_cast<uint64_t>(lane);
        bool ok = false;
        if (j < dk) {
          const edge_t e = bk + static_cast<edge_t>(j);
          const uint32_t u = static_cast<uint32_t>(ig.getDestinationVertex(e));
          if (in.check(u)) ok = call_functor(f, ig, vk, u, e);
        }
        const uint64_t hits = sygraph::hipdev::ballot(ok);
        if (hits) {
          if (lane == __ffsll(static_cast<unsigned long long>(hits)) - 1) out.insert(vk);
          if constexpr (kShort) break;
        }
      }
    }
  }
}

} // namespace detail
} // namespace advance
} // namespace operators
} // namespace sygraph
