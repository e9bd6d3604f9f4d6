/*
 * formats/csr.hpp -- host CSR container (reference formats/csr.hpp:32-215),
 * including the transpose used for directed graphs (invert, :179-209).
 */
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

#include <sygraph/utils/types.hpp>

namespace sygraph {
namespace formats {

template<typename ValueT, typename IndexT = types::index_t, typename OffsetT = types::offset_t>
class CSR {
public:
  CSR() = default;

  CSR(std::vector<OffsetT> row_offsets, std::vector<IndexT> column_indices, std::vector<ValueT> nnz_values)
      : _row_offsets(std::move(row_offsets)), _column_indices(std::move(column_indices)), _nnz_values(std::move(nnz_values)) {}

  CSR(IndexT n_rows, OffsetT n_nonzeros) {
    _row_offsets.resize(static_cast<size_t>(n_rows) + 1);
    _column_indices.resize(n_nonzeros);
    _nnz_values.resize(n_nonzeros);
  }

  IndexT getRowOffsetsSize() const { return _row_offsets.empty() ? 0 : static_cast<IndexT>(_row_offsets.size() - 1); }
  OffsetT getNumNonzeros() const { return static_cast<OffsetT>(_column_indices.size()); }

  const std::vector<OffsetT>& getRowOffsets() const { return _row_offsets; }
  std::vector<OffsetT>& getRowOffsets() { return _row_offsets; }
  const std::vector<IndexT>& getColumnIndices() const { return _column_indices; }
  std::vector<IndexT>& getColumnIndices() { return _column_indices; }
  const std::vector<ValueT>& getValues() const { return _nnz_values; }
  std::vector<ValueT>& getValues() { return _nnz_values; }

  void setRowOffsets(const std::vector<OffsetT>& offsets) { _row_offsets = offsets; }
  void setColumnIndices(const std::vector<IndexT>& indices) { _column_indices = indices; }
  void setNnzValues(const std::vector<ValueT>& values) { _nnz_values = values; }

  /* Transpose: counting sort of the arcs by destination; each in-list keeps
   * its sources in ascending row order. */
  CSR<ValueT, IndexT, OffsetT> invert() const {
    const size_t n = getRowOffsetsSize();
    const size_t m = _column_indices.size();
    std::vector<OffsetT> off(n + 1, 0);
    std::vector<IndexT> cols(m);
    std::vector<ValueT> vals(m);
    for (size_t e = 0; e < m; ++e) ++off[static_cast<size_t>(_column_indices[e]) + 1];
    for (size_t r = 0; r < n; ++r) off[r + 1] += off[r];
    std::vector<OffsetT> cursor(off.begin(), off.end() - 1);
    for (size_t r = 0; r < n; ++r)
      for (OffsetT e = _row_offsets[r]; e < _row_offsets[r + 1]; ++e) {
        const size_t c = static_cast<size_t>(_column_indices[e]);
        const OffsetT d = cursor[c]++;
        cols[d] = static_cast<IndexT>(r);
        if (!_nnz_values.empty()) vals[d] = _nnz_values[e];
      }
    return CSR<ValueT, IndexT, OffsetT>(std::move(off), std::move(cols), std::move(vals));
  }

private:
  std::vector<OffsetT> _row_offsets;
  std::vector<IndexT> _column_indices;
  std::vector<ValueT> _nnz_values;
};

} // namespace formats
} // namespace sygraph
