/*
 * sycl/event.hpp -- sygraph::Event (reference sycl/event.hpp:19-44): a queue
 * event with wait / waitAndThrow / getRuntime in milliseconds.
 */
#pragma once

#include <sycl/sycl.hpp>

namespace sygraph {

class Event : public sycl::event {
public:
  Event() = default;
  Event(const sycl::event& e) : sycl::event(e) {}

  void wait() { sycl::event::wait(); }
  void waitAndThrow() { sycl::event::wait_and_throw(); }

  float getRuntime() const {
    auto start = sycl::event::get_profiling_info<sycl::info::event_profiling::command_start>();
    auto end = sycl::event::get_profiling_info<sycl::info::event_profiling::command_end>();
    return static_cast<float>(end - start) / 1e6f;
  }
};

} // namespace sygraph
