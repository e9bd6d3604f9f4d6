/*
 * algorithms/tc.hpp -- triangle counting (reference algorithms/tc.hpp:25-143):
 * init(), run<EnableProfiling>(), getNumTriangles().  The count is exact and
 * 64-bit (segmented intersection on the (degree, id)-oriented DAG); the
 * reference's per-arc merge adds its running count inside the loop
 * (tc.hpp:104) and stores uint32.  getNumTriangles(v) is the number of
 * triangles whose lowest-ranked corner is v.
 */
#pragma once

#include <memory>
#include <stdexcept>
#include <vector>

#include "sygraph_hip.h"
#include <sycl/sycl.hpp>
#include <sygraph/sycl/event.hpp>
#ifdef ENABLE_PROFILING
#include <sygraph/utils/profiler.hpp>
#endif

namespace sygraph {
namespace algorithms {

template<typename GraphType>
class TC {
  using vertex_t = typename GraphType::vertex_t;

public:
  TC(GraphType& g) : _g(g) {}
  ~TC() { release(); }

  void init() {
    release();
    void* p = nullptr;
    if (sg_alloc(_g.getQueue().handle(), sizeof(uint64_t) * (_g.getVertexCount() + 1), SG_SPACE_DEVICE, &p) != SG_OK)
      throw std::runtime_error(sg_last_error());
    _per_vertex = static_cast<uint64_t*>(p);
    /* zero until run() counts, as the reference's TCInstance fills its array (tc.hpp:35-36) */
    if (sg_fill(_g.getQueue().handle(), _per_vertex, 0, 8, _g.getVertexCount() + 1) != SG_OK)
      throw std::runtime_error(sg_last_error());
    _ready = true;
  }
  void reset() {
    release();
    _ready = false;
  }

  template<bool EnableProfiling = false>
  void run() {
    if (!_ready) throw std::runtime_error("TC instance not initialized");
    sycl::queue& q = _g.getQueue();
    auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
    if (sg_tc(_g.handle(), &_count, _per_vertex) != SG_OK) throw std::runtime_error(sg_last_error());
    ev.finish();
#ifdef ENABLE_PROFILING
    sygraph::Profiler::addEvent(ev, "TC");
#endif
  }

  size_t getNumTriangles(vertex_t v) const {
    if (!_ready) throw std::runtime_error("TC instance not initialized");
    uint64_t x = 0;
    if (sg_memcpy(_g.getQueue().handle(), &x, _per_vertex + v, 8) != SG_OK) throw std::runtime_error(sg_last_error());
    return (size_t)x;
  }
  size_t getNumTriangles() const {
    if (!_ready) throw std::runtime_error("TC instance not initialized");
    return (size_t)_count;
  }

private:
  void release() {
    if (_per_vertex) sg_free(_g.getQueue().handle(), _per_vertex);
    _per_vertex = nullptr;
  }
  GraphType& _g;
  uint64_t* _per_vertex = nullptr;
  uint64_t _count = 0;
  bool _ready = false;
};

} // namespace algorithms
} // namespace sygraph
