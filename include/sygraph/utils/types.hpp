/*
 * utils/types.hpp -- index/offset/bitmap types (reference utils/types.hpp:12-27).
 * BITMAP_SIZE defaults to 64 here: one bitmap word is one wave64 ballot.
 */
#pragma once

#include <cstddef>
#include <cstdint>

#ifndef BITMAP_SIZE
#define BITMAP_SIZE 64
#endif
#ifndef CU_SIZE
#define CU_SIZE 256
#endif

namespace sygraph {
namespace types {

typedef unsigned int index_t;
typedef size_t offset_t;
#if BITMAP_SIZE == 64
typedef uint64_t bitmap_type_t;
#else
#error "this build keeps one 64-bit bitmap word per wavefront: BITMAP_SIZE must be 64"
#endif

namespace detail {

constexpr unsigned byte_size = 8;
constexpr size_t COMPUTE_UNIT_SIZE = CU_SIZE; /* workgroup size of the operator kernels */

} // namespace detail

} // namespace types
} // namespace sygraph
