/*
 * frontier/impls/mlb_frontier.hpp -- the two-level bitmap frontier
 * (reference frontier/impls/mlb_frontier.hpp:29-449), stored by
 * libsygraph_hip (sg_frontier_t).
 *
 * Level 0 has one bit per element in 64-bit words (one word per wavefront);
 * level 1 has one bit per level-0 word that may be non-empty.  Device views
 * (MLBDevice) insert with a test-then-atomic-OR and set the level-1 bit only
 * when the level-0 word was empty; remove clears level 0 only, so level 1 stays
 * a superset (as mlb_frontier.hpp:99-103).  computeActiveFrontier is an
 * ordered set-scan-compact whose count stays on the device.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cassert>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "sygraph_hip.h"
#include <sycl/sycl.hpp>
#include <sygraph/sycl/event.hpp>
#include <sygraph/utils/types.hpp>

namespace sygraph {
namespace frontier {
namespace detail {

inline void sg_check(int st) {
  if (st != SG_OK) throw std::runtime_error(sg_last_error());
}

template<size_t Levels, typename B = types::bitmap_type_t>
class BitmapState {
public:
  std::array<size_t, Levels> size;
  std::array<std::vector<B>, Levels> data;
};

/* Device view of a frontier: plain pointers, copyable into kernels. */
template<typename T, size_t Levels = 2, typename B = types::bitmap_type_t>
class MLBDevice {
  static_assert(Levels == 2, "the device frontier has two levels");

public:
  using bitmap_type = B;

  MLBDevice() = default;
  explicit MLBDevice(const sg_frontier_view& v)
      : _num_elems(static_cast<uint32_t>(v.n)), _data{v.l0, v.l1}, _size{static_cast<uint32_t>(v.words0), static_cast<uint32_t>(v.words1)},
        _offsets(v.list), _offsets_size(v.list_size) {}

  __host__ __device__ inline uint32_t getBitmapSize() const { return _size[0]; }
  __host__ __device__ inline uint32_t getBitmapSize(unsigned level) const { return _size[level]; }
  __host__ __device__ inline uint32_t getNumElems() const { return _num_elems; }
  __host__ __device__ inline uint32_t getBitmapRange() const { return 64; }
  __host__ __device__ inline bitmap_type* getData() const { return _data[0]; }
  __host__ __device__ inline bitmap_type* getData(unsigned level) const { return _data[level]; }
  __host__ __device__ inline uint32_t getBitmapIndex(uint32_t idx) const { return idx >> 6; }
  /* compacted active elements (ascending) and their count, after computeActiveFrontier */
  __host__ __device__ inline int* getOffsets() const { return reinterpret_cast<int*>(_offsets); }
  __host__ __device__ inline uint32_t* getOffsetsSize() const { return _offsets_size; }

  __device__ inline bool check(uint32_t idx) const { return (_data[0][idx >> 6] >> (idx & 63)) & 1ull; }
  __device__ inline bool check(unsigned level, uint32_t idx) const { return (_data[level][idx >> 6] >> (idx & 63)) & 1ull; }

  __device__ inline bool insert(T idx) const {
    const uint32_t i = static_cast<uint32_t>(idx);
    const bitmap_type b = bitmap_type(1) << (i & 63);
    bitmap_type* w = &_data[0][i >> 6];
    if (*w & b) return true;
    const bitmap_type old = __hip_atomic_fetch_or(w, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) {
      const uint32_t wi = i >> 6;
      __hip_atomic_fetch_or(&_data[1][wi >> 6], bitmap_type(1) << (wi & 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
  }
  __device__ inline bool remove(uint32_t idx) const {
    __hip_atomic_fetch_and(&_data[0][idx >> 6], ~(bitmap_type(1) << (idx & 63)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  __device__ inline void set(uint32_t idx, bool val) const {
    if (val) insert(static_cast<T>(idx));
    else remove(idx);
  }

protected:
  uint32_t _num_elems = 0;
  bitmap_type* _data[Levels] = {nullptr, nullptr};
  uint32_t _size[Levels] = {0, 0};
  int32_t* _offsets = nullptr;
  uint32_t* _offsets_size = nullptr;
};

/* Host owner of a device frontier; move-only like mlb_frontier.hpp:181-206. */
template<typename T, size_t Levels = 2, typename DeviceFrontier = MLBDevice<T, Levels>>
class FrontierMLB {
public:
  using bitmap_type = typename DeviceFrontier::bitmap_type;
  using device_frontier_type = DeviceFrontier;
  using frontier_state_type = BitmapState<Levels, bitmap_type>;

  FrontierMLB(sycl::queue& q, size_t num_elems) : _queue(&q) {
    sg_check(sg_frontier_create(q.handle(), (int64_t)num_elems, &_h));
    refresh();
  }
  FrontierMLB(const FrontierMLB&) = delete;
  FrontierMLB& operator=(const FrontierMLB&) = delete;
  FrontierMLB(FrontierMLB&& o) noexcept : _queue(o._queue), _h(o._h), _view(o._view) { o._h = nullptr; }
  FrontierMLB& operator=(FrontierMLB&&) = delete;
  ~FrontierMLB() {
    if (_h) sg_frontier_destroy(_h);
  }

  size_t getBitmapSize() const { return _view.getBitmapSize(); }
  size_t getNumElems() const { return _view.getNumElems(); }
  size_t getBitmapRange() const { return _view.getBitmapRange(); }
  bool selfAllocated() const { return false; }
  sg_frontier_t handle() const { return _h; }
  sycl::queue& getQueue() const { return *_queue; }

  bool empty() const {
    int e = 0;
    sg_check(sg_frontier_empty(_h, &e));
    return e != 0;
  }
  bool check(size_t idx) const {
    int p = 0;
    sg_check(sg_frontier_check(_h, (int64_t)idx, &p));
    return p != 0;
  }
  bool insert(size_t idx) {
    sg_check(sg_frontier_insert(_h, (int64_t)idx));
    return true;
  }
  bool remove(size_t idx) {
    sg_check(sg_frontier_remove(_h, (int64_t)idx));
    return true;
  }
  size_t size() const {
    int64_t s = 0;
    sg_check(sg_frontier_size(_h, &s));
    return (size_t)s;
  }
  /* this |= other; level 1 is rebuilt (the reference leaves it stale) */
  sygraph::Event merge(const FrontierMLB& other) {
    auto e = sycl::event::begin(_queue->handle(), _queue->stream(), _queue->origin());
    sg_check(sg_frontier_merge(_h, other._h));
    e.finish();
    return e;
  }
  sygraph::Event intersect(const FrontierMLB& other) {
    auto e = sycl::event::begin(_queue->handle(), _queue->stream(), _queue->origin());
    sg_check(sg_frontier_intersect(_h, other._h));
    e.finish();
    return e;
  }

  frontier_state_type saveState() {
    frontier_state_type st;
    std::vector<bitmap_type> all((size_t)_view.getBitmapSize(0) + _view.getBitmapSize(1));
    sg_check(sg_frontier_save(_h, all.data()));
    st.size[0] = _view.getBitmapSize(0);
    st.size[1] = _view.getBitmapSize(1);
    st.data[0].assign(all.begin(), all.begin() + st.size[0]);
    st.data[1].assign(all.begin() + st.size[0], all.end());
    return st;
  }
  void loadState(const frontier_state_type& st) {
    assert(st.size[0] == _view.getBitmapSize(0) && st.size[1] == _view.getBitmapSize(1));
    std::vector<bitmap_type> all(st.data[0]);
    all.insert(all.end(), st.data[1].begin(), st.data[1].end());
    sg_check(sg_frontier_load(_h, all.data()));
  }
  void clear() { sg_check(sg_frontier_clear(_h)); }

  const DeviceFrontier& getDeviceFrontier() const { return _view; }

  /* Ordered compaction (mlb_frontier.hpp:388-442): elements of the frontier
   * (or, with invert, level-0 words that are not full) into getOffsets(). The
   * count stays on the device in getOffsetsSize(); nothing is waited for. */
  sygraph::Event computeActiveFrontier(bool invert = false) const {
    auto e = sycl::event::begin(_queue->handle(), _queue->stream(), _queue->origin());
    sg_check(sg_frontier_compact_async(_h, invert ? 2 : 0));
    e.finish();
    return e;
  }

  static void swap(FrontierMLB& a, FrontierMLB& b) {
    sg_check(sg_frontier_swap(a._h, b._h));
    a.refresh();
    b.refresh();
  }

private:
  void refresh() {
    sg_frontier_view v{};
    sg_check(sg_frontier_get_view(_h, &v));
    _view = DeviceFrontier(v);
  }

  sycl::queue* _queue;
  sg_frontier_t _h = nullptr;
  DeviceFrontier _view{};
};

} // namespace detail
} // namespace frontier
} // namespace sygraph
