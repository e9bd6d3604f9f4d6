/*
 * frontier/frontier.hpp -- Frontier<T, Type>, makeFrontier and swap
 * (reference frontier/frontier.hpp:14-101).
 */
#pragma once

#include <stdexcept>

#include <sygraph/frontier/frontier_settings.hpp>
#include <sygraph/frontier/impls/bitmap_frontier.hpp>
#include <sygraph/frontier/impls/mlb_frontier.hpp>

namespace sygraph {
namespace frontier {
namespace detail {

template<typename T, frontier_type Type>
class frontier_impl_t;

/* placeholder for "no frontier" arguments of the operators */
template<typename T>
class frontier_impl_t<T, frontier_type::none> {
public:
  frontier_impl_t() = default;
  struct device_frontier_type {
    __device__ inline bool insert(uint32_t) const { return false; }
    __device__ inline bool check(uint32_t) const { return false; }
  };
  device_frontier_type getDeviceFrontier() const { return {}; }
};

template<typename T>
class frontier_impl_t<T, frontier_type::bitmap> : public FrontierBitmap<T> {
public:
  using FrontierBitmap<T>::FrontierBitmap;
};

template<typename T>
class frontier_impl_t<T, frontier_type::mlb> : public FrontierMLB<T> {
public:
  using FrontierMLB<T>::FrontierMLB;
};
} // namespace detail

template<typename T, frontier_type Type = frontier_type::mlb>
class Frontier : public detail::frontier_impl_t<T, Type> {
public:
  using detail::frontier_impl_t<T, Type>::frontier_impl_t;
  using type_t = T;
  static constexpr frontier_type kind = Type;
};

/* A frontier sized by the graph's vertex (or edge) count. */
template<frontier_view View, frontier_type Type, typename GraphType>
auto makeFrontier(sycl::queue& q, const GraphType& graph) {
  if constexpr (View == frontier_view::vertex) {
    return Frontier<typename GraphType::vertex_t, Type>(q, graph.getVertexCount());
  } else if constexpr (View == frontier_view::edge) {
    return Frontier<typename GraphType::edge_t, Type>(q, graph.getEdgeCount());
  } else {
    static_assert(View == frontier_view::vertex || View == frontier_view::edge, "Invalid frontier view");
  }
}

/* Swap the storage of two frontiers (frontier.hpp:92-99). */
template<typename T, frontier_type FT>
void swap(Frontier<T, FT>& a, Frontier<T, FT>& b) {
  if constexpr (FT == frontier_type::bitmap) {
    detail::FrontierBitmap<T>::swap(a, b);
  } else if constexpr (FT == frontier_type::mlb) {
    detail::FrontierMLB<T>::swap(a, b);
  }
}

} // namespace frontier
} // namespace sygraph
