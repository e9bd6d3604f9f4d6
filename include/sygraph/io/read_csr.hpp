/*
 * io/read_csr.hpp -- CSR readers and writers (reference io/read_csr.hpp:22-411):
 *   fromMatrix  dense adjacency text ("n" then n*n values; zeros skipped)
 *   fromMM      MatrixMarket coordinate (1-based; symmetric files store the
 *               mirrored entry too), sorted by (row, col)
 *   fromCSR     text "n", n+1 offsets, then ids and values
 *   fromCOO     COO -> CSR by counting sort (in-row order = input order)
 *   toBinary / fromBinary  binary CSR: magic 0x5359475243535201, version,
 *               flags (1 directed, 2 weighted), reserved, n+1, m, arrays;
 *               a file without the magic is the headerless legacy layout.
 * The on-disk layouts are the reference's, so its datasets load unchanged.
 */
#pragma once

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <istream>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include <sygraph/formats/coo.hpp>
#include <sygraph/formats/csr.hpp>
#include <sygraph/graph/properties.hpp>
#include <sygraph/io/matrix_market.hpp>

namespace sygraph {
namespace io {
namespace csr {
namespace detail {
namespace binary {
static constexpr uint64_t magic = 0x5359475243535201ULL;
static constexpr uint8_t directed_mask = 0x1;
static constexpr uint8_t weighted_mask = 0x2;
} // namespace binary
} // namespace detail

template<typename ValueT, typename IndexT, typename OffsetT>
sygraph::formats::CSR<ValueT, IndexT, OffsetT> fromMatrix(std::istream& iss) {
  size_t n = 0;
  iss >> n;
  std::vector<OffsetT> off{0};
  std::vector<IndexT> cols;
  std::vector<ValueT> vals;
  off.reserve(n + 1);
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < n; ++j) {
      ValueT x{};
      iss >> x;
      if (x != static_cast<ValueT>(0)) {
        cols.push_back(static_cast<IndexT>(j));
        vals.push_back(x);
      }
    }
    off.push_back(static_cast<OffsetT>(cols.size()));
  }
  return sygraph::formats::CSR<ValueT, IndexT, OffsetT>(std::move(off), std::move(cols), std::move(vals));
}

template<typename ValueT, typename IndexT, typename OffsetT>
sygraph::formats::CSR<ValueT, IndexT, OffsetT> fromMM(std::istream& iss, sygraph::graph::Properties* properties = nullptr) {
  io::detail::mm::Banner banner;
  size_t rows = 0, cols = 0, nnz = 0;
  bool have_dims = false, have_banner = false;
  std::vector<std::tuple<IndexT, IndexT, ValueT>> entries;
  std::string line;
  while (std::getline(iss, line)) {
    if (line.empty()) continue;
    if (line[0] == '%') {
      if (!have_banner) {
        have_banner = true;
        banner.read(line);
        banner.template validate<ValueT, IndexT, OffsetT>();
        if (properties) {
          properties->directed = !banner.isSymmetric();
          properties->weighted = !banner.isPattern();
        }
      }
      continue;
    }
    std::istringstream ls(line);
    if (!have_dims) {
      ls >> rows >> cols >> nnz;
      have_dims = true;
      entries.reserve(banner.isSymmetric() ? 2 * nnz : nnz);
      continue;
    }
    IndexT r, c;
    ValueT x = static_cast<ValueT>(1);
    ls >> r >> c;
    if (!banner.isPattern()) ls >> x;
    entries.emplace_back(r - 1, c - 1, x);
    if (banner.isSymmetric()) entries.emplace_back(c - 1, r - 1, x);
  }
  std::sort(entries.begin(), entries.end(), [](const auto& a, const auto& b) {
    return std::get<0>(a) < std::get<0>(b) || (std::get<0>(a) == std::get<0>(b) && std::get<1>(a) < std::get<1>(b));
  });
  std::vector<OffsetT> off(rows + 1, 0);
  std::vector<IndexT> ci(entries.size());
  std::vector<ValueT> vv(entries.size());
  for (const auto& e : entries) ++off[static_cast<size_t>(std::get<0>(e)) + 1];
  for (size_t i = 0; i < rows; ++i) off[i + 1] += off[i];
  for (size_t k = 0; k < entries.size(); ++k) {
    ci[k] = std::get<1>(entries[k]);
    vv[k] = std::get<2>(entries[k]);
  }
  return sygraph::formats::CSR<ValueT, IndexT, OffsetT>(std::move(off), std::move(ci), std::move(vv));
}

template<typename ValueT, typename IndexT, typename OffsetT>
sygraph::formats::CSR<ValueT, IndexT, OffsetT> fromMM(const std::string& filename, sygraph::graph::Properties* properties = nullptr) {
  std::ifstream file(filename);
  if (!file.is_open()) throw std::runtime_error("Failed to open file: " + filename);
  return fromMM<ValueT, IndexT, OffsetT>(file, properties);
}

template<typename ValueT, typename IndexT, typename OffsetT>
sygraph::formats::CSR<ValueT, IndexT, OffsetT> fromCSR(std::istream& iss) {
  size_t n = 0;
  iss >> n;
  std::vector<OffsetT> off(n + 1);
  for (auto& o : off) iss >> o;
  const size_t m = n ? static_cast<size_t>(off.back()) : 0;
  std::vector<IndexT> cols(m);
  std::vector<ValueT> vals(m);
  for (auto& c : cols) iss >> c;
  for (auto& v : vals) iss >> v;
  return sygraph::formats::CSR<ValueT, IndexT, OffsetT>(std::move(off), std::move(cols), std::move(vals));
}

template<typename ValueT, typename IndexT, typename OffsetT>
sygraph::formats::CSR<ValueT, IndexT, OffsetT> fromCOO(const sygraph::formats::COO<ValueT, IndexT, OffsetT>& coo) {
  const auto& r = coo.getRowIndices();
  const auto& c = coo.getColumnIndices();
  const auto& w = coo.getValues();
  size_t n = 0;
  for (size_t i = 0; i < r.size(); ++i) n = std::max<size_t>(n, std::max<size_t>(r[i], c[i]) + 1);
  std::vector<OffsetT> off(n + 1, 0);
  for (auto x : r) ++off[static_cast<size_t>(x) + 1];
  for (size_t i = 0; i < n; ++i) off[i + 1] += off[i];
  std::vector<OffsetT> cursor(off.begin(), off.end() - 1);
  std::vector<IndexT> cols(r.size());
  std::vector<ValueT> vals(r.size());
  for (size_t i = 0; i < r.size(); ++i) {
    const OffsetT p = cursor[r[i]]++;
    cols[p] = c[i];
    vals[p] = w[i];
  }
  return sygraph::formats::CSR<ValueT, IndexT, OffsetT>(std::move(off), std::move(cols), std::move(vals));
}

template<typename ValueT, typename IndexT, typename OffsetT>
void toBinary(const sygraph::formats::CSR<ValueT, IndexT, OffsetT>& csr, std::ostream& oss,
              const sygraph::graph::Properties& properties = sygraph::graph::Properties()) {
  if (!oss) throw std::runtime_error("Failed to write binary CSR matrix");
  const auto& off = csr.getRowOffsets();
  const auto& cols = csr.getColumnIndices();
  const auto& vals = csr.getValues();
  const uint64_t magic = detail::binary::magic;
  const uint8_t version = 1;
  uint8_t flags = 0;
  if (properties.directed) flags |= detail::binary::directed_mask;
  if (properties.weighted) flags |= detail::binary::weighted_mask;
  const uint16_t r16 = 0;
  const uint32_t r32 = 0;
  const size_t nrows = off.size(), nnz = cols.size();
  oss.write(reinterpret_cast<const char*>(&magic), 8);
  oss.write(reinterpret_cast<const char*>(&version), 1);
  oss.write(reinterpret_cast<const char*>(&flags), 1);
  oss.write(reinterpret_cast<const char*>(&r16), 2);
  oss.write(reinterpret_cast<const char*>(&r32), 4);
  oss.write(reinterpret_cast<const char*>(&nrows), sizeof(size_t));
  oss.write(reinterpret_cast<const char*>(&nnz), sizeof(size_t));
  oss.write(reinterpret_cast<const char*>(off.data()), off.size() * sizeof(OffsetT));
  oss.write(reinterpret_cast<const char*>(cols.data()), cols.size() * sizeof(IndexT));
  oss.write(reinterpret_cast<const char*>(vals.data()), vals.size() * sizeof(ValueT));
}

template<typename ValueT, typename IndexT, typename OffsetT>
sygraph::formats::CSR<ValueT, IndexT, OffsetT> fromBinary(std::istream& iss, sygraph::graph::Properties* properties = nullptr) {
  if (!iss) throw std::runtime_error("Failed to read binary CSR matrix");
  uint64_t first = 0;
  size_t nrows = 0, nnz = 0;
  iss.read(reinterpret_cast<char*>(&first), 8);
  if (!iss) throw std::runtime_error("Failed to read binary CSR matrix");
  if (first == detail::binary::magic) {
    uint8_t version = 0, flags = 0;
    uint16_t r16 = 0;
    uint32_t r32 = 0;
    iss.read(reinterpret_cast<char*>(&version), 1);
    iss.read(reinterpret_cast<char*>(&flags), 1);
    iss.read(reinterpret_cast<char*>(&r16), 2);
    iss.read(reinterpret_cast<char*>(&r32), 4);
    iss.read(reinterpret_cast<char*>(&nrows), sizeof(size_t));
    iss.read(reinterpret_cast<char*>(&nnz), sizeof(size_t));
    if (properties) {
      properties->directed = (flags & detail::binary::directed_mask) != 0;
      properties->weighted = (flags & detail::binary::weighted_mask) != 0;
    }
  } else {
    nrows = static_cast<size_t>(first); /* legacy: no header, directed + weighted */
    iss.read(reinterpret_cast<char*>(&nnz), sizeof(size_t));
    if (properties) {
      properties->directed = true;
      properties->weighted = true;
    }
  }
  std::vector<OffsetT> off(nrows);
  std::vector<IndexT> cols(nnz);
  std::vector<ValueT> vals(nnz);
  iss.read(reinterpret_cast<char*>(off.data()), off.size() * sizeof(OffsetT));
  iss.read(reinterpret_cast<char*>(cols.data()), cols.size() * sizeof(IndexT));
  iss.read(reinterpret_cast<char*>(vals.data()), vals.size() * sizeof(ValueT));
  if (!iss) throw std::runtime_error("Failed to read binary CSR matrix");
  return sygraph::formats::CSR<ValueT, IndexT, OffsetT>(std::move(off), std::move(cols), std::move(vals));
}

} // namespace csr
} // namespace io
} // namespace sygraph
