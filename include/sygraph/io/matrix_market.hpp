/*
 * io/matrix_market.hpp -- MatrixMarket banner (reference
 * io/matrix_market.hpp:21-151): "%%MatrixMarket matrix coordinate
 * <real|integer|complex|pattern> <general|symmetric|skew-symmetric|hermitian>".
 */
#pragma once

#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace sygraph {
namespace io {
namespace detail {
namespace mm {

enum class mm_object { matrix, vector };
enum class mm_format { coordinate, array };
enum class mm_field { real, integer, complex, pattern };
enum class mm_symmetry { general, symmetric, skew_symmetric, hermitian };

struct Banner {
  std::string banner;
  mm_object object = mm_object::matrix;
  mm_format format = mm_format::coordinate;
  mm_field field = mm_field::real;
  mm_symmetry symmetry = mm_symmetry::general;

  void read(const std::string& line) {
    std::istringstream iss(line);
    std::string tag, obj, fmt, fld, sym;
    iss >> tag >> obj >> fmt >> fld >> sym;
    if (tag != "%%MatrixMarket") throw std::runtime_error("Invalid MatrixMarket banner");
    if (obj == "matrix") object = mm_object::matrix;
    else if (obj == "vector") object = mm_object::vector;
    else throw std::runtime_error("Invalid object type in MatrixMarket banner");
    if (fmt == "coordinate") format = mm_format::coordinate;
    else if (fmt == "array") format = mm_format::array;
    else throw std::runtime_error("Invalid format type in MatrixMarket banner");
    if (fld == "real") field = mm_field::real;
    else if (fld == "integer") field = mm_field::integer;
    else if (fld == "complex") field = mm_field::complex;
    else if (fld == "pattern") field = mm_field::pattern;
    else throw std::runtime_error("Invalid field type in MatrixMarket banner");
    if (sym == "general") symmetry = mm_symmetry::general;
    else if (sym == "symmetric") symmetry = mm_symmetry::symmetric;
    else if (sym == "skew-symmetric") symmetry = mm_symmetry::skew_symmetric;
    else if (sym == "hermitian") symmetry = mm_symmetry::hermitian;
    else throw std::runtime_error("Invalid symmetry type in MatrixMarket banner");
    banner = tag;
  }

  bool isMatrix() const { return object == mm_object::matrix; }
  bool isVector() const { return object == mm_object::vector; }
  bool isCoordinate() const { return format == mm_format::coordinate; }
  bool isArray() const { return format == mm_format::array; }
  bool isReal() const { return field == mm_field::real; }
  bool isInteger() const { return field == mm_field::integer; }
  bool isComplex() const { return field == mm_field::complex; }
  bool isPattern() const { return field == mm_field::pattern; }
  bool isGeneral() const { return symmetry == mm_symmetry::general; }
  bool isSymmetric() const { return symmetry == mm_symmetry::symmetric; }

  template<typename ValueT, typename IndexT, typename OffsetT>
  void validate() const {
    if (object != mm_object::matrix) throw std::runtime_error("Invalid MatrixMarket object type");
    if (format != mm_format::coordinate) throw std::runtime_error("Invalid MatrixMarket format type");
    if (field == mm_field::real && !std::is_floating_point<ValueT>::value) throw std::runtime_error("Invalid MatrixMarket field type");
    if (field == mm_field::integer && !std::is_integral<ValueT>::value) throw std::runtime_error("Invalid MatrixMarket field type");
  }
};

} // namespace mm
} // namespace detail
} // namespace io
} // namespace sygraph
