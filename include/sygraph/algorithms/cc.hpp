/*
 * algorithms/cc.hpp -- connected components (reference algorithms/cc.hpp:34-192):
 * init(source), run<EnableProfiling>().  Labels converge to the largest vertex
 * id of each component (the fixpoint of the reference's max-label
 * propagation); libsygraph_hip reaches it by concurrent union-find on
 * undirected graphs and by atomic-max propagation on directed ones.
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
class CC {
  using vertex_t = typename GraphType::vertex_t;
  using edge_t = typename GraphType::edge_t;

public:
  CC(GraphType& g) : _g(g) {}
  ~CC() { release(); }

  void init(const vertex_t& source) {
    release();
    void* p = nullptr;
    if (sg_alloc(_g.getQueue().handle(), sizeof(uint32_t) * (_g.getVertexCount() + 1), SG_SPACE_DEVICE, &p) != SG_OK)
      throw std::runtime_error(sg_last_error());
    _labels = static_cast<uint32_t*>(p);
    _source = source;
  }
  void reset() { release(); }

  template<bool EnableProfiling = false>
  void run() {
    if (!_labels) throw std::runtime_error("CC instance not initialized");
    sycl::queue& q = _g.getQueue();
    auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
    if (sg_cc(_g.handle(), _labels, &_rounds) != SG_OK) throw std::runtime_error(sg_last_error());
    ev.finish();
#ifdef ENABLE_PROFILING
    sygraph::Profiler::addEvent(ev, "advance");
#endif
  }

  std::vector<vertex_t> getLabels() const {
    std::vector<uint32_t> h(_g.getVertexCount());
    if (sg_memcpy(_g.getQueue().handle(), h.data(), _labels, h.size() * 4) != SG_OK) throw std::runtime_error(sg_last_error());
    return std::vector<vertex_t>(h.begin(), h.end());
  }
  vertex_t getLabel(size_t v) const { return getLabels().at(v); }

  vertex_t getParents(size_t) const { throw std::runtime_error("Not implemented"); }

private:
  void release() {
    if (_labels) sg_free(_g.getQueue().handle(), _labels);
    _labels = nullptr;
  }
  GraphType& _g;
  uint32_t* _labels = nullptr;
  vertex_t _source = 0;
  int32_t _rounds = 0;
};

} // namespace algorithms
} // namespace sygraph
