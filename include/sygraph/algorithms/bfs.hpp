/*
 * algorithms/bfs.hpp -- BFS (reference algorithms/bfs.hpp:29-351): init(source),
 * run(direction, alpha, beta) -> BFSRunDetails, getDistance(s).
 *
 * run() calls the fused gfx950 BFS of libsygraph_hip (sg_bfs): push, pull or
 * the reference's hybrid rule (push -> pull when the new frontier's degree sum
 * exceeds unexplored degree / alpha, pull -> push when |frontier| < n / beta,
 * bfs.hpp:325-350).  Unreached vertices keep distance n+1 (bfs.hpp:62-67).
 * Distances are copied to the host once per run (the reference reads device
 * memory from the host, which only works on shared-memory devices).
 */
#pragma once

#include <memory>
#include <set>
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

enum class bfs_direction { push, pull, hybrid };

struct BFSRunDetails {
  size_t iterations = 0;
  std::set<size_t> push_steps;
  std::set<size_t> pull_steps;
};

namespace detail {

inline void sg_call(int st) {
  if (st != SG_OK) throw std::runtime_error(sg_last_error());
}

template<typename GraphType>
struct BFSInstance {
  using vertex_t = typename GraphType::vertex_t;
  using edge_t = typename GraphType::edge_t;

  GraphType& G;
  vertex_t source;
  int32_t* depth = nullptr; /* device */
  std::vector<edge_t> host;  /* distances after run() */

  BFSInstance(GraphType& g, vertex_t src) : G(g), source(src) {
    if (static_cast<size_t>(src) >= g.getVertexCount()) throw std::runtime_error("BFS source out of range");
    void* p = nullptr;
    sg_call(sg_alloc(g.getQueue().handle(), sizeof(int32_t) * (g.getVertexCount() + 1), SG_SPACE_DEVICE, &p));
    depth = static_cast<int32_t*>(p);
    host.assign(g.getVertexCount(), static_cast<edge_t>(g.getVertexCount() + 1));
    host[src] = 0;
  }
  ~BFSInstance() { sg_free(G.getQueue().handle(), depth); }

  size_t getVisitedVertices() const {
    int64_t e = 0, v = 0;
    sg_call(sg_graph_reached_edges(G.handle(), depth, &e, &v));
    return (size_t)v;
  }
  size_t getVisitedEdges() const {
    int64_t e = 0, v = 0;
    sg_call(sg_graph_reached_edges(G.handle(), depth, &e, &v));
    return (size_t)e;
  }
};

} // namespace detail

template<typename GraphType>
class BFS {
  using vertex_t = typename GraphType::vertex_t;
  using edge_t = typename GraphType::edge_t;

public:
  BFS(GraphType& g) : _g(g) {}

  void init(const vertex_t& source) { _instance = std::make_unique<detail::BFSInstance<GraphType>>(_g, source); }
  void reset() { _instance.reset(); }

  BFSRunDetails run(bfs_direction direction = bfs_direction::push, float alpha = 1.0f, float beta = 1.0f) {
    if (!_instance) throw std::runtime_error("BFS instance not initialized");
    sg_bfs_options opt;
    sg_bfs_options_default(&opt);
    opt.policy = direction == bfs_direction::push ? SG_POLICY_PUSH
                                                  : (direction == bfs_direction::pull ? SG_POLICY_PULL : SG_POLICY_HYBRID);
    opt.alpha = alpha;
    opt.beta = beta;
    opt.record_levels = 1;
    auto stats = std::make_unique<sg_bfs_stats>();
    sycl::queue& q = _g.getQueue();
    auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
    detail::sg_call(sg_bfs(_g.handle(), (int32_t)_instance->source, &opt, _instance->depth, stats.get()));
    ev.finish();
    BFSRunDetails d;
    d.iterations = (size_t)stats->levels;
    for (int32_t l = 0; l < stats->levels && l < SG_MAX_LEVELS_RECORDED; ++l) {
      if (stats->level[l].direction == 0) d.push_steps.insert((size_t)l);
      else d.pull_steps.insert((size_t)l);
    }
    std::vector<int32_t> dd(_g.getVertexCount());
    detail::sg_call(sg_memcpy(q.handle(), dd.data(), _instance->depth, dd.size() * sizeof(int32_t)));
    for (size_t i = 0; i < dd.size(); ++i) _instance->host[i] = static_cast<edge_t>(dd[i]);
#ifdef ENABLE_PROFILING
    sygraph::Profiler::addEvent(ev, "advance");
    sygraph::Profiler::addVisitedEdges((size_t)stats->reached_edges);
#endif
    return d;
  }

  edge_t getDistance(size_t vertex) const { return _instance->host.at(vertex); }
  std::vector<edge_t> getDistances() const { return _instance->host; }

  /* parents are not tracked (the reference allocates them but never writes them) */
  vertex_t getParent(size_t) const { return static_cast<vertex_t>(-1); }
  std::vector<vertex_t> getParents() const { return std::vector<vertex_t>(_g.getVertexCount(), static_cast<vertex_t>(-1)); }

  /* device distances (int32, unreached = n+1) */
  const int32_t* getDeviceDistances() const { return _instance ? _instance->depth : nullptr; }

private:
  GraphType& _g;
  std::unique_ptr<detail::BFSInstance<GraphType>> _instance;
};

} // namespace algorithms
} // namespace sygraph
