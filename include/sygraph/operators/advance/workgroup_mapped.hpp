/*
 * operators/advance/workgroup_mapped.hpp -- edge-balanced push (reference
 * operators/advance/workgroup_mapped.hpp:20-213).
 *
 * A workgroup takes 256 sources, scans their degrees (wave shuffles + LDS)
 * and then walks the concatenated arc ranges 256 arcs at a time; each thread
 * finds the owner of its arc by a binary search over the scan in LDS.  Every
 * load of the column array is coalesced, whatever the degree mix.  Arc ids
 * stay OffsetT (the reference truncates them to 32 bits, :108,129).
 */
#pragma once

#include <sygraph/operators/advance/common.hpp>

namespace sygraph {
namespace operators {
namespace advance {
namespace detail {

template<typename G, typename Out, typename F>
__global__ void __launch_bounds__(kBlock) push_workgroup_mapped_kernel(G g, SourceList src, Out out, F f) {
  using edge_t = typename G::edge_t;
  __shared__ uint64_t s_incl[kBlock];
  __shared__ edge_t s_start[kBlock];
  __shared__ uint32_t s_vertex[kBlock];
  __shared__ uint64_t s_wave[kBlock / 64];
  const uint32_t count = src.size();
  const int tid = threadIdx.x, lane = sygraph::hipdev::lane_id(), wv = tid >> 6;
  for (uint32_t base = blockIdx.x * kBlock; base < count; base += gridDim.x * kBlock) {
    const uint32_t i = base + tid;
    uint64_t deg = 0;
    uint32_t v = 0;
    edge_t st = 0;
    if (i < count) {
      v = src.at(i);
      st = g.getFirstNeighbor(v);
      deg = g.getDegree(v);
    }
    uint64_t x = deg;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t t = __shfl_up(x, o, 64);
      if (lane >= o) x += t;
    }
    if (lane == 63) s_wave[wv] = x;
    __syncthreads();
    uint64_t before = 0, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wv) before += s_wave[w];
      total += s_wave[w];
    }
    s_incl[tid] = before + x;
    s_start[tid] = st;
    s_vertex[tid] = v;
    __syncthreads();
    for (uint64_t r = tid; r < total; r += kBlock) {
      int lo = 0, hi = kBlock - 1; /* first slot with inclusive scan > r */
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_incl[mid] > r) hi = mid;
        else lo = mid + 1;
      }
      const uint64_t excl = lo ? s_incl[lo - 1] : 0;
      const edge_t e = s_start[lo] + static_cast<edge_t>(r - excl);
      const uint32_t dst = static_cast<uint32_t>(g.getDestinationVertex(e));
      if (call_functor(f, g, s_vertex[lo], dst, e)) out.insert(dst);
    }
    __syncthreads();
  }
}

} // namespace detail
} // namespace advance
} // namespace operators
} // namespace sygraph
