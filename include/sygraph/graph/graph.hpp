/*
 * graph/graph.hpp -- graph base class and concepts (reference
 * graph/graph.hpp:26-78).
 */
#pragma once

#include <concepts>
#include <cstddef>
#include <functional>
#include <utility>

#include <sycl/sycl.hpp>
#include <sygraph/graph/properties.hpp>

namespace sygraph {
namespace graph {

template<typename VertexT, typename EdgeT, typename WeightT>
class Graph {
public:
  Graph(Properties properties) : _properties(properties) {}
  Graph() = default;
  virtual ~Graph() = default;

  const Properties& getProperties() const { return _properties; }

  virtual size_t getVertexCount() const = 0;
  virtual size_t getEdgeCount() const = 0;
  virtual size_t getDegree(VertexT vertex) const = 0;
  virtual EdgeT getFirstNeighbor(VertexT vertex) const = 0;
  virtual VertexT getSourceVertex(EdgeT edge) const = 0;
  virtual VertexT getDestinationVertex(EdgeT edge) const = 0;
  virtual WeightT getEdgeWeight(EdgeT edge) const = 0;

private:
  Properties _properties;
};

namespace detail {

template<typename DeviceGraphT>
concept DeviceGraphConcept = requires(DeviceGraphT g) {
  { g.getVertexCount() } -> std::convertible_to<size_t>;
  { g.getEdgeCount() } -> std::convertible_to<size_t>;
  { g.getDegree(std::declval<typename DeviceGraphT::vertex_t>()) } -> std::convertible_to<size_t>;
  { g.getFirstNeighbor(std::declval<typename DeviceGraphT::vertex_t>()) } -> std::convertible_to<typename DeviceGraphT::edge_t>;
  { g.getSourceVertex(std::declval<typename DeviceGraphT::edge_t>()) } -> std::convertible_to<typename DeviceGraphT::vertex_t>;
  { g.getDestinationVertex(std::declval<typename DeviceGraphT::edge_t>()) } -> std::convertible_to<typename DeviceGraphT::vertex_t>;
  { g.getEdgeWeight(std::declval<typename DeviceGraphT::edge_t>()) } -> std::convertible_to<typename DeviceGraphT::weight_t>;
};

template<typename GraphT>
concept GraphConcept = requires(GraphT g) {
  { g.getQueue() } -> std::convertible_to<sycl::queue&>;
  { g.getDeviceGraph() };
  { g.getInverseDeviceGraph() };
  { g.getProperties() } -> std::convertible_to<Properties>;
  { g.getVertexCount() } -> std::convertible_to<size_t>;
  { g.getEdgeCount() } -> std::convertible_to<size_t>;
};

} // namespace detail

} // namespace graph
} // namespace sygraph
