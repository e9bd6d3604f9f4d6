/*
 * graph/build.hpp -- build a device graph from a host CSR (reference
 * graph/build.hpp:31-41).
 */
#pragma once

#include <sycl/sycl.hpp>
#include <sygraph/formats/csr.hpp>
#include <sygraph/graph/graph.hpp>
#include <sygraph/graph/impls/graph_csr.hpp>
#include <sygraph/graph/properties.hpp>
#include <sygraph/utils/memory.hpp>

namespace sygraph {
namespace graph {
namespace build {

template<memory::space Space, typename IndexT, typename OffsetT, typename ValueT>
auto fromCSR(sycl::queue& q, const sygraph::formats::CSR<ValueT, IndexT, OffsetT>& csr, graph::Properties properties = graph::Properties()) {
  return detail::GraphCSR<Space, IndexT, OffsetT, ValueT>{q, csr, properties};
}

template<memory::space Space, typename IndexT, typename OffsetT, typename ValueT>
auto fromCSR(sycl::queue& q, sygraph::formats::CSR<ValueT, IndexT, OffsetT>&& csr, graph::Properties properties = graph::Properties()) {
  return detail::GraphCSR<Space, IndexT, OffsetT, ValueT>{q, std::move(csr), properties};
}

} // namespace build
} // namespace graph
} // namespace sygraph
