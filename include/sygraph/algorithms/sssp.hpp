/*
 * algorithms/sssp.hpp -- SSSP (reference algorithms/sssp.hpp:26-222):
 * init(source), run<EnableProfiling>(), getDistance(v).  run() calls the
 * gfx950 frontier Bellman-Ford of libsygraph_hip (sg_sssp: atomic-min
 * relaxation over degree-class queues).  Distances have the graph's weight
 * type; unreached = n+1 (sssp.hpp:43).
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
class SSSP {
  using vertex_t = typename GraphType::vertex_t;
  using edge_t = typename GraphType::edge_t;
  using weight_t = typename GraphType::weight_t;

public:
  SSSP(GraphType& g) : _g(g) {}

  void init(const vertex_t& source) {
    if (static_cast<size_t>(source) >= _g.getVertexCount()) throw std::runtime_error("SSSP source out of range");
    release();
    void* p = nullptr;
    if (sg_alloc(_g.getQueue().handle(), sizeof(weight_t) * (_g.getVertexCount() + 1), SG_SPACE_DEVICE, &p) != SG_OK)
      throw std::runtime_error(sg_last_error());
    _dist = static_cast<weight_t*>(p);
    _source = source;
    _host.assign(_g.getVertexCount(), static_cast<weight_t>(_g.getVertexCount() + 1));
    _host[source] = 0;
    _ready = true;
  }
  void reset() {
    release();
    _ready = false;
  }
  ~SSSP() { release(); }

  template<bool EnableProfiling = false>
  void run() {
    if (!_ready) throw std::runtime_error("SSSP instance not initialized");
    sycl::queue& q = _g.getQueue();
    auto stats = std::make_unique<sg_sssp_stats>();
    auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
    if (sg_sssp(_g.handle(), (int32_t)_source, SG_LB_BUCKETING, _dist, stats.get()) != SG_OK)
      throw std::runtime_error(sg_last_error());
    ev.finish();
    if (sg_memcpy(q.handle(), _host.data(), _dist, _host.size() * sizeof(weight_t)) != SG_OK)
      throw std::runtime_error(sg_last_error());
    _rounds = stats->rounds;
#ifdef ENABLE_PROFILING
    sygraph::Profiler::addEvent(ev, "advance");
    sygraph::Profiler::addVisitedEdges((size_t)stats->reached_edges);
#endif
  }

  weight_t getDistance(size_t vertex) const { return _host.at(vertex); }
  std::vector<weight_t> getDistances() const { return _host; }
  int32_t getRounds() const { return _rounds; }

  vertex_t getParents(size_t) const { throw std::runtime_error("Not implemented"); }

private:
  void release() {
    if (_dist) sg_free(_g.getQueue().handle(), _dist);
    _dist = nullptr;
  }
  GraphType& _g;
  weight_t* _dist = nullptr;
  vertex_t _source = 0;
  bool _ready = false;
  int32_t _rounds = 0;
  std::vector<weight_t> _host;
};

} // namespace algorithms
} // namespace sygraph
