/*
 * io/matrices.hpp -- two small adjacency-matrix graphs used as fixtures
 * (reference io/matrices.hpp:21-41; same names and data, so tests written
 * against the reference find them).
 */
#pragma once

#include <string_view>

namespace sygraph {
namespace io {
namespace storage {
namespace matrices {

/* 0-1, 0-2, 1-2, 2-3, 2-4, 4-5 */
constexpr std::string_view symmetric_6nodes = "6\n"
                                              "0 1 1 0 0 0\n"
                                              "1 0 1 0 0 0\n"
                                              "1 1 0 1 1 0\n"
                                              "0 0 1 0 0 0\n"
                                              "0 0 1 0 0 1\n"
                                              "0 0 0 0 1 0";
/* as above without 4-5: components {0,1,2,3,4} and {5} */
constexpr std::string_view two_cc = "6\n"
                                    "0 1 1 0 0 0\n"
                                    "1 0 1 0 0 0\n"
                                    "1 1 0 1 1 0\n"
                                    "0 0 1 0 0 0\n"
                                    "0 0 1 0 0 0\n"
                                    "0 0 0 0 0 0";

} // namespace matrices
} // namespace storage
} // namespace io
} // namespace sygraph
