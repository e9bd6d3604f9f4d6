/*
 * utils/profiler.hpp -- per-tag kernel timing and MTEPS report (reference
 * utils/profiler.hpp:18-69).  The printed lines keep the reference's layout
 * (" Kernel [tag x count] ... Time: t ms", "Total GPU Time: t ms",
 * "Total Edge-Througput (MTEPS): x MTEPS"), so logs of both stay comparable;
 * tags are reported in the order they were first used.  Active when
 * ENABLE_PROFILING is defined.
 */
#pragma once

#ifdef ENABLE_PROFILING

#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include <sygraph/sycl/event.hpp>

namespace sygraph {

namespace detail {
/* The process-wide log: events per tag and the TEPS numerator. */
struct ProfileLog {
  std::vector<std::pair<std::string, std::vector<sygraph::Event>>> by_tag;
  size_t visited_edges = 0;

  std::vector<sygraph::Event>& tag(const std::string& name) {
    for (auto& entry : by_tag)
      if (entry.first == name) return entry.second;
    by_tag.emplace_back(name, std::vector<sygraph::Event>{});
    return by_tag.back().second;
  }

  static ProfileLog& instance() {
    static ProfileLog log;
    return log;
  }
};
} // namespace detail

class Profiler {
public:
  static void addEvent(sygraph::Event event, std::string tag = "") {
    detail::ProfileLog::instance().tag(tag.empty() ? std::string("default") : tag).push_back(event);
  }
  static void addVisitedEdges(size_t visited_edges) { detail::ProfileLog::instance().visited_edges += visited_edges; }
  static void clear() { detail::ProfileLog::instance() = detail::ProfileLog{}; }

  static void print(bool detailed = false) {
    const detail::ProfileLog& log = detail::ProfileLog::instance();
    double total_ms = 0.0;
    for (const auto& [name, evs] : log.by_tag) {
      std::cout << " Kernel [" << name << " x " << evs.size() << "] ";
      double tag_ms = 0.0;
      for (size_t k = 0; k < evs.size(); ++k) {
        const double t = evs[k].getRuntime();
        tag_ms += t;
        if (detailed) std::cout << "(" << k << ") " << t << " ms ";
      }
      if (detailed) std::cout << "| ";
      std::cout << "Time: " << tag_ms << " ms" << std::endl;
      total_ms += tag_ms;
    }
    std::cout << "Total GPU Time: " << total_ms << " ms" << std::endl;
    const double seconds = total_ms / 1e3;
    const double mteps = total_ms > 0.0 ? static_cast<double>(log.visited_edges) / 1e6 / seconds : 0.0;
    std::cout << "Total Edge-Througput (MTEPS): " << mteps << " MTEPS" << std::endl;
  }
};

} // namespace sygraph

#endif
