/*
 * operators/advance/advance.hpp -- the advance operator (reference
 * operators/advance/advance.hpp:46-218):
 *   advance::frontier<direction, load_balancer, InView, OutView>(G, in, out, f[, expected_size])
 *   advance::frontier<load_balancer, InView, OutView>(G, in, out, f[, expected_size])   (push)
 *   advance::frontier<direction, load_balancer, InView>(G, in, f[, expected_size])      (no output)
 *   advance::vertices<load_balancer, OutView>(G, out, f)                               (graph view)
 *   advance::vertices<load_balancer>(G, f)
 * The functor is fused into the kernel: f(src, dst, edge, weight) -> bool.
 * All work is queued on the graph's queue; the returned Event brackets it.
 */
#pragma once

#include <sygraph/frontier/frontier.hpp>
#include <sygraph/frontier/frontier_settings.hpp>
#include <sygraph/graph/graph.hpp>
#include <sygraph/operators/advance/bucketing.hpp>
#include <sygraph/operators/advance/common.hpp>
#include <sygraph/operators/advance/workgroup_mapped.hpp>
#include <sygraph/operators/advance/workitem_mapped.hpp>
#include <sygraph/operators/config.hpp>
#include <sygraph/sycl/event.hpp>

namespace sygraph {
namespace operators {
namespace advance {

namespace detail {

template<load_balancer Lb, typename G, typename Out, typename F>
void launch_push(sycl::queue& q, unsigned grid, const G& g, SourceList src, Out out, F f) {
  hipStream_t s = q.stream();
  if constexpr (Lb == load_balancer::workgroup_mapped) {
    push_workgroup_mapped_kernel<<<grid, kBlock, 0, s>>>(g, src, out, f);
  } else if constexpr (Lb == load_balancer::bucketing) {
    push_bucketing_kernel<<<grid, kBlock, 0, s>>>(g, src, out, f);
  } else if constexpr (Lb == load_balancer::workitem_mapped) {
    push_workitem_kernel<<<grid, kBlock, 0, s>>>(g, src, out, f);
  } else {
    push_subgroup_kernel<<<grid, kBlock, 0, s>>>(g, src, out, f);
  }
  check_launch();
}

template<direction Dir, load_balancer Lb, frontier::frontier_view InView, frontier::frontier_view OutView,
         typename GraphT, typename InF, typename OutF, typename LambdaT>
sygraph::Event run(GraphT& graph, const InF& in, OutF& out, LambdaT&& functor, int expected_size) {
  sycl::queue& q = graph.getQueue();
  static_assert(InView != frontier::frontier_view::edge, "edge frontiers are not supported by advance");
  constexpr bool kInsert = OutView == frontier::frontier_view::vertex;
  using OutDev = std::decay_t<decltype(out.getDeviceFrontier())>;
  OutSink<OutDev, kInsert> sink{out.getDeviceFrontier()};
  std::decay_t<LambdaT> f = functor;
  auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
  const unsigned cus = sygraph::detail::device::getNumComputeUnits(q);
  if constexpr (InView == frontier::frontier_view::graph) {
    SourceList src{nullptr, nullptr, static_cast<uint32_t>(graph.getVertexCount())};
    const unsigned need = (unsigned)((graph.getVertexCount() + kBlock - 1) / kBlock);
    launch_push<Lb>(q, std::max(1u, std::min(need, cus * 8)), graph.getDeviceGraph(), src, sink, f);
  } else if constexpr (is_pull<Dir>()) {
    const auto& ig = graph.getInverseDeviceGraph();
    const unsigned words = (unsigned)((graph.getVertexCount() + 63) / 64);
    const unsigned grid = std::max(1u, std::min((words + 3) / 4, cus * 8));
    pull_kernel<is_short_circuit<Dir>()><<<grid, kBlock, 0, q.stream()>>>(
        ig, in.getDeviceFrontier(), static_cast<uint32_t>(graph.getVertexCount()), sink, f);
    check_launch();
  } else {
    const uint32_t* dcount = compact_sources(in);
    const unsigned grid = push_grid(q, expected_size, dcount);
    SourceList src{in.getDeviceFrontier().getOffsets(), dcount, 0};
    launch_push<Lb>(q, grid, graph.getDeviceGraph(), src, sink, f);
  }
  ev.finish();
  return ev;
}

} // namespace detail

/* graph view: every vertex is a source (advance.hpp:46-66) */
template<load_balancer Lb, frontier::frontier_view FW, typename GraphT, typename LambdaT, typename T,
         frontier::frontier_type FrontierType>
sygraph::Event vertices(GraphT& graph, sygraph::frontier::Frontier<T, FrontierType>& out, LambdaT&& functor) {
  auto in = sygraph::frontier::Frontier<bool, sygraph::frontier::frontier_type::none>{};
  return detail::run<direction::push, Lb, frontier::frontier_view::graph, FW>(graph, in, out, std::forward<LambdaT>(functor),
                                                                              frontier::size::fetch_from_memory);
}

template<load_balancer Lb, typename GraphT, typename LambdaT>
sygraph::Event vertices(GraphT& graph, LambdaT&& functor) {
  auto out = sygraph::frontier::Frontier<void, sygraph::frontier::frontier_type::none>{};
  return vertices<Lb, frontier::frontier_view::none>(graph, out, std::forward<LambdaT>(functor));
}

template<direction Direction, load_balancer Lb, frontier::frontier_view InView, frontier::frontier_view OutView,
         typename GraphT, typename LambdaT, typename T, frontier::frontier_type FrontierType>
sygraph::Event frontier(GraphT& graph,
                        sygraph::frontier::Frontier<T, FrontierType>& in,
                        sygraph::frontier::Frontier<T, FrontierType>& out,
                        LambdaT&& functor,
                        frontier::size::frontier_size_t expected_size = frontier::size::fetch_from_memory) {
  return detail::run<Direction, Lb, InView, OutView>(graph, in, out, std::forward<LambdaT>(functor), expected_size);
}

template<load_balancer Lb, frontier::frontier_view InView, frontier::frontier_view OutView, typename GraphT,
         typename LambdaT, typename T, frontier::frontier_type FrontierType>
sygraph::Event frontier(GraphT& graph,
                        sygraph::frontier::Frontier<T, FrontierType>& in,
                        sygraph::frontier::Frontier<T, FrontierType>& out,
                        LambdaT&& functor,
                        frontier::size::frontier_size_t expected_size = frontier::size::fetch_from_memory) {
  return detail::run<direction::push, Lb, InView, OutView>(graph, in, out, std::forward<LambdaT>(functor), expected_size);
}

/* no output frontier (advance.hpp:190-218; there the output type cannot be
 * deduced, here it is a placeholder) */
template<direction Direction, load_balancer Lb, frontier::frontier_view InView, typename GraphT, typename LambdaT,
         typename T, frontier::frontier_type FrontierType>
sygraph::Event frontier(GraphT& graph,
                        sygraph::frontier::Frontier<T, FrontierType>& in,
                        LambdaT&& functor,
                        frontier::size::frontier_size_t expected_size = frontier::size::fetch_from_memory) {
  auto out = sygraph::frontier::Frontier<void, sygraph::frontier::frontier_type::none>{};
  return detail::run<Direction, Lb, InView, frontier::frontier_view::none>(graph, in, out, std::forward<LambdaT>(functor),
                                                                            expected_size);
}

template<load_balancer Lb, frontier::frontier_view InView, typename GraphT, typename LambdaT, typename T,
         frontier::frontier_type FrontierType>
sygraph::Event frontier(GraphT& graph,
                        sygraph::frontier::Frontier<T, FrontierType>& in,
                        LambdaT&& functor,
                        frontier::size::frontier_size_t expected_size = frontier::size::fetch_from_memory) {
  auto out = sygraph::frontier::Frontier<void, sygraph::frontier::frontier_type::none>{};
  return detail::run<direction::push, Lb, InView, frontier::frontier_view::none>(graph, in, out,
                                                                                  std::forward<LambdaT>(functor), expected_size);
}

} // namespace advance
} // namespace operators
} // namespace sygraph
