/*
 * operators/advance/bucketing.hpp -- three-bin push (reference
 * operators/advance/bucketing.hpp:21-293).
 *
 * For each chunk of 256 sources the workgroup sorts vertices into bins by
 * degree: >= 1024 arcs -> the whole workgroup walks them 256 at a time;
 * >= 64 -> one wavefront, 64 at a time; smaller -> the owning lane alone.
 * Bin lists are compacted with ballots into LDS (no tail races: the
 * reference's tail counters are reset without a barrier, :96-97).
 */
#pragma once

#include <sygraph/operators/advance/common.hpp>

namespace sygraph {
namespace operators {
namespace advance {
namespace detail {

constexpr uint64_t kBucketWave = 64;    /* degree for the wavefront bin */
constexpr uint64_t kBucketGroup = 1024; /* degree for the workgroup bin */

template<typename G, typename Out, typename F>
__global__ void __launch_bounds__(kBlock) push_bucketing_kernel(G g, SourceList src, Out out, F f) {
  using edge_t = typename G::edge_t;
  namespace hd = sygraph::hipdev;
  __shared__ uint32_t s_big[kBlock], s_med[kBlock];
  __shared__ uint32_t s_nbig, s_nmed;
  const uint32_t count = src.size();
  const int tid = threadIdx.x, lane = hd::lane_id(), wv = tid >> 6;
  for (uint32_t base = blockIdx.x * kBlock; base < count; base += gridDim.x * kBlock) {
    if (tid == 0) {
      s_nbig = 0;
      s_nmed = 0;
    }
    __syncthreads();
    const uint32_t i = base + tid;
    uint32_t v = 0;
    uint64_t deg = 0;
    if (i < count) {
      v = src.at(i);
      deg = g.getDegree(v);
    }
    const bool big = deg >= kBucketGroup, med = !big && deg >= kBucketWave;
    const uint64_t mb = hd::ballot(big), mm = hd::ballot(med);
    uint32_t bb = 0, bm = 0;
    if (lane == 0) {
      if (mb) bb = atomicAdd(&s_nbig, (uint32_t)__popcll(mb));
      if (mm) bm = atomicAdd(&s_nmed, (uint32_t)__popcll(mm));
    }
    bb = __shfl(bb, 0, 64);
    bm = __shfl(bm, 0, 64);
    if (big) s_big[bb + hd::lanes_below(mb)] = v;
    if (med) s_med[bm + hd::lanes_below(mm)] = v;
    __syncthreads();
    const uint32_t nbig = s_nbig, nmed = s_nmed;
    for (uint32_t k = 0; k < nbig; ++k) { /* workgroup bin */
      const uint32_t u = s_big[k];
      const edge_t b = g.getFirstNeighbor(u);
      const uint64_t d = g.getDegree(u);
      for (uint64_t j = tid; j < d; j += kBlock) {
        const edge_t e = b + static_cast<edge_t>(j);
        const uint32_t dst = static_cast<uint32_t>(g.getDestinationVertex(e));
        if (call_functor(f, g, u, dst, e)) out.insert(dst);
      }
    }
    for (uint32_t k = wv; k < nmed; k += kBlock / 64) { /* wavefront bin */
      const uint32_t u = s_med[k];
      const edge_t b = g.getFirstNeighbor(u);
      const uint64_t d = g.getDegree(u);
      for (uint64_t j = lane; j < d; j += 64) {
        const edge_t e = b + static_cast<edge_t>(j);
        const uint32_t dst = static_cast<uint32_t>(g.getDestinationVertex(e));
        if (call_functor(f, g, u, dst, e)) out.insert(dst);
      }
    }
    if (i < count && !big && !med) { /* lane bin */
      const edge_t b = g.getFirstNeighbor(v);
      for (uint64_t j = 0; j < deg; ++j) {
        const edge_t e = b + static_cast<edge_t>(j);
        const uint32_t dst = static_cast<uint32_t>(g.getDestinationVertex(e));
        if (call_functor(f, g, v, dst, e)) out.insert(dst);
      }
    }
    __syncthreads();
  }
}

} // namespace detail
} // namespace advance
} // namespace operators
} // namespace sygraph
