/*
 * operators/intersection/intersection.hpp -- frontier intersection
 * (reference operators/intersection/intersection.hpp:44-55 and
 * bitmap_intersect_impl.hpp:21-30): out = in1 & in2 in one pass, both levels.
 * The functor is accepted for interface compatibility and not called, as in
 * the reference.  Works for bitmap and mlb frontiers.
 */
#pragma once

#include <stdexcept>

#include <sygraph/frontier/frontier.hpp>
#include <sygraph/sycl/event.hpp>

namespace sygraph {
namespace operators {
namespace intersection {

template<typename GraphT, typename T, sygraph::frontier::frontier_type FrontierType, typename LambdaT>
sygraph::Event execute(GraphT& graph,
                       const sygraph::frontier::Frontier<T, FrontierType>& in1,
                       const sygraph::frontier::Frontier<T, FrontierType>& in2,
                       sygraph::frontier::Frontier<T, FrontierType>& out,
                       LambdaT&&) {
  static_assert(FrontierType != sygraph::frontier::frontier_type::none, "Frontier type not implemented");
  sycl::queue& q = graph.getQueue();
  auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
  if (sg_frontier_and(out.handle(), in1.handle(), in2.handle()) != SG_OK) throw std::runtime_error(sg_last_error());
  ev.finish();
  return ev;
}

} // namespace intersection
} // namespace operators
} // namespace sygraph
