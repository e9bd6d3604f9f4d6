/*
 * algorithms/bc.hpp -- betweenness centrality (reference algorithms/bc.hpp:24-224):
 * init(source), run().  Brandes dependencies of the source are accumulated
 * into bc(v) for v != source, without halving for undirected graphs, as the
 * reference does.  run() uses the batched kernel of libsygraph_hip (sg_bc):
 * FP32 sigma with per-level power-of-two scaling, FP32 delta; several sources
 * can be given to init() at once and then share one batch.
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
class BC {
  using vertex_t = typename GraphType::vertex_t;

public:
  BC(GraphType& g) : _g(g) {}
  ~BC() { release(); }

  void init(const vertex_t source) { init(std::vector<vertex_t>{source}); }
  void init(const std::vector<vertex_t>& sources) {
    release();
    void* p = nullptr;
    if (sg_alloc(_g.getQueue().handle(), sizeof(float) * (_g.getVertexCount() + 1), SG_SPACE_DEVICE, &p) != SG_OK)
      throw std::runtime_error(sg_last_error());
    _bc = static_cast<float*>(p);
    if (sg_fill(_g.getQueue().handle(), _bc, 0, 4, _g.getVertexCount()) != SG_OK) throw std::runtime_error(sg_last_error());
    _sources.assign(sources.begin(), sources.end());
  }
  void reset() { release(); }

  void run() {
    if (!_bc) throw std::runtime_error("BC instance not initialized");
    sycl::queue& q = _g.getQueue();
    std::vector<int32_t> s(_sources.begin(), _sources.end());
    auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
    if (sg_bc(_g.handle(), (int32_t)s.size(), s.data(), _bc) != SG_OK) throw std::runtime_error(sg_last_error());
    ev.finish();
#ifdef ENABLE_PROFILING
    sygraph::Profiler::addEvent(ev, "BC");
#endif
  }

  std::vector<float> getBCValues() const {
    std::vector<float> h(_g.getVertexCount());
    if (sg_memcpy(_g.getQueue().handle(), h.data(), _bc, h.size() * 4) != SG_OK) throw std::runtime_error(sg_last_error());
    return h;
  }

private:
  void release() {
    if (_bc) sg_free(_g.getQueue().handle(), _bc);
    _bc = nullptr;
  }
  GraphType& _g;
  float* _bc = nullptr;
  std::vector<vertex_t> _sources;
};

} // namespace algorithms
} // namespace sygraph
