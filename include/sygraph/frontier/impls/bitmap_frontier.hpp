/*
 * frontier/impls/bitmap_frontier.hpp -- the flat bitmap frontier (reference
 * frontier/impls/bitmap_frontier.hpp:23-444, deprecated there, still used by
 * its advance and intersection tests).
 *
 * It shares the storage and kernels of the two-level frontier; its device view
 * (BitmapDevice) keeps the flat-bitmap names.  computeActiveFrontier() returns
 * the number of active level-0 words, as the reference's does
 * (bitmap_frontier.hpp:396-437), and leaves their ids in getOffsets().
 */
#pragma once

#include <sygraph/frontier/impls/mlb_frontier.hpp>

namespace sygraph {
namespace frontier {
namespace detail {

template<typename T, typename B = types::bitmap_type_t>
class BitmapDevice : public MLBDevice<T, 2, B> {
public:
  using MLBDevice<T, 2, B>::MLBDevice;
  __host__ __device__ inline uint32_t getBitmapSize() const { return this->_size[0]; }
};

template<typename T>
class FrontierBitmap : public FrontierMLB<T, 2, BitmapDevice<T>> {
  using Base = FrontierMLB<T, 2, BitmapDevice<T>>;

public:
  using Base::Base;

  size_t computeActiveFrontier() const {
    int64_t count = 0;
    sg_check(sg_frontier_compact(this->handle(), 1, &count));
    return (size_t)count;
  }

  /* host copy of the active elements (ascending); `size` gets their count */
  void getActiveElements(T*& elems, size_t& size) const {
    int64_t count = 0;
    sg_check(sg_frontier_compact(this->handle(), 0, &count));
    size = (size_t)count;
    if (elems && count) {
      std::vector<int32_t> tmp(count);
      sg_check(sg_memcpy(this->getQueue().handle(), tmp.data(), this->getDeviceFrontier().getOffsets(), 4 * count));
      for (int64_t i = 0; i < count; ++i) elems[i] = static_cast<T>(tmp[i]);
    }
  }

  static void swap(FrontierBitmap& a, FrontierBitmap& b) { Base::swap(a, b); }
};

} // namespace detail
} // namespace frontier
} // namespace sygraph
