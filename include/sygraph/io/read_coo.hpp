/*
 * io/read_coo.hpp -- COO text reader (reference io/read_coo.hpp:33-80):
 * optional '%' comment lines, a "rows cols edges" header, then "u v [w]" per
 * line; undirected input stores both directions; a third column marks the
 * graph weighted.
 */
#pragma once

#include <istream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <sygraph/formats/coo.hpp>
#include <sygraph/graph/properties.hpp>

namespace sygraph {
namespace io {
namespace coo {

template<typename ValueT, typename IndexT, typename OffsetT = types::offset_t>
sygraph::formats::COO<ValueT, IndexT, OffsetT> fromCOO(std::istream& iss, bool undirected = false,
                                                       sygraph::graph::Properties* properties = nullptr) {
  while (iss.peek() == '%') iss.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  std::string line;
  if (std::getline(iss, line).eof() && line.empty()) throw std::runtime_error("Error: could not read the first line of the file.");
  size_t rows = 0, cols = 0, edges = 0;
  std::istringstream(line) >> rows >> cols >> edges;
  std::vector<IndexT> r, c;
  std::vector<ValueT> w;
  r.reserve(edges * (undirected ? 2 : 1));
  c.reserve(r.capacity());
  w.reserve(r.capacity());
  bool weighted = false;
  while (std::getline(iss, line)) {
    std::istringstream ls(line);
    IndexT u, v;
    if (!(ls >> u >> v)) continue;
    ValueT x = static_cast<ValueT>(1);
    if (ls >> x) weighted = true;
    r.push_back(u);
    c.push_back(v);
    w.push_back(x);
    if (undirected) {
      r.push_back(v);
      c.push_back(u);
      w.push_back(x);
    }
  }
  if (properties) {
    properties->directed = !undirected;
    properties->weighted = weighted;
  }
  return sygraph::formats::COO<ValueT, IndexT, OffsetT>(std::move(r), std::move(c), std::move(w));
}

} // namespace coo
} // namespace io
} // namespace sygraph
