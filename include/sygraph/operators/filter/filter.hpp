/*
 * operators/filter/filter.hpp -- frontier filters (reference
 * operators/filter/filter.hpp:41-69, bitmap_filter_impl.hpp:27-121):
 *   external(G, in, out, f): out is cleared, then out = { v in in : f(v) }
 *   inplace(G, frontier, f): removes every v of the frontier with f(v)
 * The input is compacted on the device and one lane handles one element.
 */
#pragma once

#include <sygraph/operators/advance/common.hpp>

namespace sygraph {
namespace operators {
namespace filter {
namespace detail {

template<bool kExternal, typename OutDev, typename InDev, typename F>
__global__ void __launch_bounds__(256) filter_kernel(InDev in, OutDev out, F f) {
  const uint32_t count = *in.getOffsetsSize();
  const int* list = in.getOffsets();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) {
    const uint32_t v = (uint32_t)list[i];
    if (f(v)) {
      if constexpr (kExternal) out.insert(v);
      else in.remove(v);
    }
  }
}

template<bool kExternal, typename GraphT, typename InF, typename OutF, typename LambdaT>
sygraph::Event launch(GraphT& graph, const InF& in, OutF* out, LambdaT&& functor) {
  sycl::queue& q = graph.getQueue();
  auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
  const uint32_t* dcount = operators::advance::detail::compact_sources(in);
  const unsigned grid = operators::advance::detail::push_grid(q, frontier::size::fetch_from_memory, dcount);
  std::decay_t<LambdaT> f = functor;
  if constexpr (kExternal) {
    filter_kernel<true><<<grid, 256, 0, q.stream()>>>(in.getDeviceFrontier(), out->getDeviceFrontier(), f);
  } else {
    filter_kernel<false><<<grid, 256, 0, q.stream()>>>(in.getDeviceFrontier(), in.getDeviceFrontier(), f);
  }
  operators::advance::detail::check_launch();
  ev.finish();
  return ev;
}

} // namespace detail

template<typename GraphT, typename T, sygraph::frontier::frontier_type FrontierType, typename LambdaT>
sygraph::Event inplace(GraphT& graph, const sygraph::frontier::Frontier<T, FrontierType>& frontier, LambdaT&& functor) {
  static_assert(FrontierType != sygraph::frontier::frontier_type::none, "Invalid frontier type");
  return detail::launch<false>(graph, frontier, (sygraph::frontier::Frontier<T, FrontierType>*)nullptr,
                               std::forward<LambdaT>(functor));
}

template<typename GraphT, typename T, sygraph::frontier::frontier_type FrontierType, typename LambdaT>
sygraph::Event external(GraphT& graph,
                        const sygraph::frontier::Frontier<T, FrontierType>& in,
                        sygraph::frontier::Frontier<T, FrontierType>& out,
                        LambdaT&& functor) {
  static_assert(FrontierType != sygraph::frontier::frontier_type::none, "Invalid frontier type");
  out.clear();
  return detail::launch<true>(graph, in, &out, std::forward<LambdaT>(functor));
}

} // namespace filter
} // namespace operators
} // namespace sygraph
