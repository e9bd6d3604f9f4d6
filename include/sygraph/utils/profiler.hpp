/*
 * utils/profiler.hpp -- per-tag kernel timing and MTEPS report (reference
 * utils/profiler.hpp:18-69); same print format, so reference-style logs stay
 * comparable.  Active when ENABLE_PROFILING is defined.
 */
#pragma once

#ifdef ENABLE_PROFILING

#include <iostream>
#include <string>
#include <unordered_map>
#include <vector>

#include <sygraph/sycl/event.hpp>

namespace sygraph {

namespace detail {
inline std::unordered_map<std::string, std::vector<sygraph::Event>>& events() {
  static auto* e = new std::unordered_map<std::string, std::vector<sygraph::Event>>();
  return *e;
}
inline size_t& numVisitedEdges() {
  static auto* v = new size_t(0);
  return *v;
}
} // namespace detail

class Profiler {
public:
  static void addEvent(sygraph::Event event, std::string tag = "") {
    if (tag.empty()) tag = "default";
    detail::events()[tag].push_back(event);
  }
  static void addVisitedEdges(size_t visited_edges) { detail::numVisitedEdges() += visited_edges; }
  static void clear() {
    detail::events().clear();
    detail::numVisitedEdges() = 0;
  }
  static void print(bool detailed = false) {
    double total_ms = 0.0;
    for (auto& [tag, evs] : detail::events()) {
      double ms = 0.0;
      size_t i = 0;
      std::cout << " Kernel [" << tag << " x " << evs.size() << "] ";
      for (auto& e : evs) {
        const double t = e.getRuntime();
        if (detailed) std::cout << "(" << i++ << ") " << t << " ms ";
        ms += t;
      }
      if (detailed) std::cout << "| ";
      std::cout << "Time: " << ms << " ms" << std::endl;
      total_ms += ms;
    }
    std::cout << "Total GPU Time: " << total_ms << " ms" << std::endl;
    const double mteps = total_ms > 0.0 ? (detail::numVisitedEdges() / 1e6) / (total_ms / 1e3) : 0.0;
    std::cout << "Total Edge-Througput (MTEPS): " << mteps << " MTEPS" << std::endl;
  }
};

} // namespace sygraph

#endif
