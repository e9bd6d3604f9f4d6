/*
 * graph/impls/graph_csr.hpp -- device CSR view and owning CSR graph
 * (reference graph/impls/graph_csr.hpp:15-439).
 *
 * GraphCSR uploads through sg_graph_create, which keeps the arrays in the
 * caller's widths (32- or 64-bit offsets, 32-bit ids, 4-byte weights) and, for
 * a directed graph, builds the inverse CSR on the device; an undirected graph
 * aliases it (graph_csr.hpp:423-425).  Edge ids are OffsetT end to end (the
 * reference truncates them to the vertex type, graph_csr.hpp:71).
 *
 * Host accessors follow the reference (graph_csr.hpp:306-371): with
 * memory::space::device they return the host CSR kept by the graph; with
 * shared or host space they return the graph's own arrays (managed or pinned
 * memory), the ones the kernels use, so host and device code see the same
 * data.  A values array always exists (all ones when the input has none), as
 * the reference allocates one of m entries unconditionally.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "sygraph_hip.h"
#include <sycl/sycl.hpp>
#include <sygraph/formats/csr.hpp>
#include <sygraph/graph/graph.hpp>
#include <sygraph/utils/memory.hpp>

namespace sygraph {
namespace graph {
namespace detail {

template<typename IndexT, typename OffsetT, typename ValueT>
class GraphCSRDevice {
public:
  using vertex_t = IndexT;
  using edge_t = OffsetT;
  using weight_t = ValueT;

  struct NeighborIterator {
    __host__ __device__ NeighborIterator(IndexT* start_ptr, IndexT* ptr) : _ptr(ptr), _start_ptr(start_ptr) {}
    __host__ __device__ inline IndexT operator*() const { return *_ptr; }
    __host__ __device__ inline NeighborIterator& operator++() {
      ++_ptr;
      return *this;
    }
    __host__ __device__ inline NeighborIterator operator+(int64_t n) const {
      NeighborIterator t = *this;
      t._ptr += n;
      return t;
    }
    __host__ __device__ inline bool operator==(const NeighborIterator& o) const { return _ptr == o._ptr; }
    __host__ __device__ inline bool operator!=(const NeighborIterator& o) const { return _ptr != o._ptr; }
    __host__ __device__ inline edge_t getIndex() const { return static_cast<edge_t>(_ptr - _start_ptr); }
    IndexT* _ptr;
    IndexT* _start_ptr;
  };

  __host__ __device__ inline size_t getVertexCount() const { return _n_rows; }
  __host__ __device__ inline size_t getEdgeCount() const { return _n_nonzeros; }
  __host__ __device__ inline size_t getDegree(vertex_t v) const { return _row_offsets[v + 1] - _row_offsets[v]; }
  __host__ __device__ inline edge_t getFirstNeighbor(vertex_t v) const { return _row_offsets[v]; }
  __host__ __device__ inline IndexT* getColumnIndices() const { return _column_indices; }
  __host__ __device__ inline OffsetT* getRowOffsets() const { return _row_offsets; }
  __host__ __device__ inline ValueT* getValues() const { return _nnz_values; }

  /* row owning arc `edge` (binary search over the offsets) */
  __host__ __device__ inline vertex_t getSourceVertex(edge_t edge) const {
    size_t lo = 0, hi = _n_rows;
    while (lo < hi) {
      const size_t mid = lo + (hi - lo) / 2;
      if (_row_offsets[mid + 1] <= edge) lo = mid + 1;
      else hi = mid;
    }
    return static_cast<vertex_t>(lo);
  }
  __host__ __device__ inline vertex_t getDestinationVertex(edge_t edge) const { return _column_indices[edge]; }
  __host__ __device__ inline weight_t getEdgeWeight(edge_t edge) const {
    return _nnz_values ? _nnz_values[edge] : static_cast<weight_t>(1);
  }
  __host__ __device__ inline NeighborIterator begin(vertex_t v) const {
    return NeighborIterator(_column_indices, _column_indices + _row_offsets[v]);
  }
  __host__ __device__ inline NeighborIterator end(vertex_t v) const {
    return NeighborIterator(_column_indices, _column_indices + _row_offsets[v + 1]);
  }

  /* |N(src) n N(dst)| by a merge of the two sorted lists; func(w) per common w */
  template<typename Func>
  __host__ __device__ inline size_t getIntersectionCount(const vertex_t& src, const vertex_t& dst, Func&& func) const {
    size_t count = 0;
    auto a = begin(src), ae = end(src), b = begin(dst), be = end(dst);
    while (a != ae && b != be) {
      if (*a == *b) {
        func(*a);
        ++a;
        ++b;
        ++count;
      } else if (*a < *b) {
        ++a;
      } else {
        ++b;
      }
    }
    return count;
  }

  IndexT _n_rows = 0;
  OffsetT _n_nonzeros = 0;
  IndexT* _column_indices = nullptr;
  OffsetT* _row_offsets = nullptr;
  ValueT* _nnz_values = nullptr;
};

template<memory::space Space, typename IndexT, typename OffsetT, typename ValueT>
class GraphCSR : public Graph<IndexT, OffsetT, ValueT> {
  static_assert(sizeof(IndexT) == 4, "vertex ids are 32-bit");
  static_assert(sizeof(OffsetT) == 4 || sizeof(OffsetT) == 8, "offsets are 32- or 64-bit");
  static_assert(sizeof(ValueT) == 4, "edge values are 4-byte (uint32 / int32 / float)");

public:
  using vertex_t = IndexT;
  using edge_t = OffsetT;
  using weight_t = ValueT;
  using device_graph_t = GraphCSRDevice<IndexT, OffsetT, ValueT>;
  static_assert(sizeof(IndexT) == 4, "vertex ids are stored as 32-bit integers");
  static_assert(sizeof(OffsetT) == 4 || sizeof(OffsetT) == 8, "row offsets are stored as 32- or 64-bit integers");
  static_assert(sizeof(ValueT) == 4, "edge values are stored as 4-byte integers or floats");

  GraphCSR(sycl::queue& q, const formats::CSR<ValueT, IndexT, OffsetT>& csr, Properties properties)
      : Graph<IndexT, OffsetT, ValueT>(properties), _queue(q), _csr(csr) {
    upload(properties);
  }
  GraphCSR(sycl::queue& q, formats::CSR<ValueT, IndexT, OffsetT>&& csr, Properties properties)
      : Graph<IndexT, OffsetT, ValueT>(properties), _queue(q), _csr(std::move(csr)) {
    upload(properties);
  }

  GraphCSR(const GraphCSR&) = delete;
  GraphCSR& operator=(const GraphCSR&) = delete;
  GraphCSR(GraphCSR&& o) noexcept
      : Graph<IndexT, OffsetT, ValueT>(o.getProperties()), _queue(o._queue), _csr(std::move(o._csr)), _handle(o._handle),
        _device_graph(o._device_graph), _inverse_device_graph(o._inverse_device_graph) {
    o._handle = nullptr;
    o._device_graph = {};
    o._inverse_device_graph = {};
  }
  GraphCSR& operator=(GraphCSR&&) = delete;

  ~GraphCSR() override {
    if (_handle) sg_graph_destroy(_handle);
  }

  device_graph_t& getDeviceGraph() { return _device_graph; }
  device_graph_t& getInverseDeviceGraph() { return _inverse_device_graph; }
  const device_graph_t& getDeviceGraph() const { return _device_graph; }
  const device_graph_t& getInverseDeviceGraph() const { return _inverse_device_graph; }

  /* the C-ABI handle, for fixed-functor algorithms */
  sg_graph_t handle() const { return _handle; }

  size_t getVertexCount() const override { return _device_graph.getVertexCount(); }
  size_t getEdgeCount() const override { return _device_graph.getEdgeCount(); }
  size_t getDegree(vertex_t v) const override { return getRowOffsets()[v + 1] - getRowOffsets()[v]; }
  edge_t getFirstNeighbor(vertex_t v) const override { return getRowOffsets()[v]; }
  vertex_t getSourceVertex(edge_t edge) const override {
    const OffsetT* off = getRowOffsets();
    size_t lo = 0, hi = getVertexCount();
    while (lo < hi) {
      const size_t mid = lo + (hi - lo) / 2;
      if (off[mid + 1] <= edge) lo = mid + 1;
      else hi = mid;
    }
    return static_cast<vertex_t>(lo);
  }
  vertex_t getDestinationVertex(edge_t edge) const override { return getColumnIndices()[edge]; }
  weight_t getEdgeWeight(edge_t edge) const override { return getValues()[edge]; }

  IndexT getOffsetsSize() const { return static_cast<IndexT>(getVertexCount() + 1); }
  OffsetT getValuesSize() const { return static_cast<OffsetT>(getEdgeCount()); }

  /* host-readable CSR arrays: the host copy (device space) or the graph's own
   * managed / pinned arrays (shared, host space) */
  IndexT* getColumnIndices() const {
    if constexpr (Space == memory::space::device) return const_cast<IndexT*>(_csr.getColumnIndices().data());
    else return _device_graph.getColumnIndices();
  }
  OffsetT* getRowOffsets() const {
    if constexpr (Space == memory::space::device) return const_cast<OffsetT*>(_csr.getRowOffsets().data());
    else return _device_graph.getRowOffsets();
  }
  ValueT* getValues() const {
    if constexpr (Space == memory::space::device) return const_cast<ValueT*>(_csr.getValues().data());
    else return _device_graph.getValues();
  }

  size_t getIntersectionCount(const vertex_t& src, const vertex_t& dst) const {
    size_t count = 0;
    const OffsetT* off = getRowOffsets();
    const IndexT* col = getColumnIndices();
    OffsetT a = off[src], ae = off[src + 1], b = off[dst], be = off[dst + 1];
    while (a < ae && b < be) {
      if (col[a] == col[b]) {
        ++count;
        ++a;
        ++b;
      } else if (col[a] < col[b]) {
        ++a;
      } else {
        ++b;
      }
    }
    return count;
  }

  sycl::queue& getQueue() const { return _queue; }

private:
  void upload(const Properties& properties) {
    const size_t n = _csr.getRowOffsetsSize();
    const size_t m = _csr.getNumNonzeros();
    int flags = 0;
    if (properties.directed) flags |= SG_GRAPH_DIRECTED;
    if (std::is_floating_point_v<ValueT>) flags |= SG_GRAPH_WEIGHT_FLOAT;
    if constexpr (Space == memory::space::shared) flags |= SG_GRAPH_SPACE_SHARED;
    if constexpr (Space == memory::space::host) flags |= SG_GRAPH_SPACE_HOST;
    if (_csr.getValues().size() != m) _csr.getValues().assign(m, static_cast<ValueT>(1));
    if (_csr.getRowOffsets().empty()) _csr.getRowOffsets().push_back(0);
    flags |= SG_GRAPH_WEIGHTED;
    int st = sg_graph_create(_queue.handle(), (int64_t)n, (int64_t)m, _csr.getRowOffsets().data(), (int)sizeof(OffsetT),
                             reinterpret_cast<const int32_t*>(_csr.getColumnIndices().data()),
                             m > 0 ? _csr.getValues().data() : nullptr, flags, &_handle);
    if (st != SG_OK) throw std::runtime_error(sg_last_error());
    sg_graph_view v{};
    sg_graph_get_view(_handle, &v);
    _device_graph = {static_cast<IndexT>(n), static_cast<OffsetT>(m), reinterpret_cast<IndexT*>(v.cols),
                     static_cast<OffsetT*>(v.offsets), static_cast<ValueT*>(v.weights)};
    _inverse_device_graph = {static_cast<IndexT>(n), static_cast<OffsetT>(m), reinterpret_cast<IndexT*>(v.inv_cols),
                             static_cast<OffsetT*>(v.inv_offsets), static_cast<ValueT*>(v.inv_weights)};
  }

  sycl::queue& _queue;
  formats::CSR<ValueT, IndexT, OffsetT> _csr;
  sg_graph_t _handle = nullptr;
  device_graph_t _device_graph{};
  device_graph_t _inverse_device_graph{};
};

} // namespace detail
} // namespace graph
} // namespace sygraph
