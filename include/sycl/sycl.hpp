/*
 * sycl/sycl.hpp -- API FACADE, not a SYCL implementation.
 *
 * The SYgraph operator API takes a `sycl::queue&` everywhere (e.g.
 * frontier/frontier.hpp:68 makeFrontier(q, G), graph/build.hpp:31 fromCSR(q,
 * ...)), and its examples construct `sycl::queue q{sycl::gpu_selector_v}`,
 * query device/platform names and allocate USM (examples/bfs/bfs.cpp:120-124,
 * examples/include/utils.hpp:127-131, tests/operators/advance.cpp:16-17).
 * So that those callers compile unchanged, this header provides exactly that
 * host-side surface, backed by libsygraph_hip: a queue is an sg_device_t (one
 * HIP device and one stream), USM maps to hipMalloc / hipMallocManaged /
 * hipHostMalloc, and events are pairs of hipEvents.  Kernel submission is
 * limited to queue::submit / handler::parallel_for over a one-dimensional
 * range (a grid-stride HIP launch of the callable) and single_task; there are
 * no buffers, accessors or nd_ranges.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sygraph_hip.h"

namespace sycl {

class exception : public std::runtime_error {
public:
  using std::runtime_error::runtime_error;
};

namespace detail {
inline void check(int st) {
  if (st != SG_OK) throw exception(sg_last_error());
}

/* One device handle per process, shared by all queues (the library's stream
 * orders everything a queue submits). */
inline std::shared_ptr<sg_device_s> shared_device(int ordinal = 0) {
  static std::weak_ptr<sg_device_s> cache;
  auto sp = cache.lock();
  if (!sp) {
    sg_device_t d = nullptr;
    check(sg_device_open(ordinal, &d));
    sp = std::shared_ptr<sg_device_s>(d, [](sg_device_t h) { sg_device_close(h); });
    cache = sp;
  }
  return sp;
}
} // namespace detail

struct gpu_selector_t {};
struct default_selector_t {};
struct cpu_selector_t {};
inline constexpr gpu_selector_t gpu_selector_v{};
inline constexpr default_selector_t default_selector_v{};
inline constexpr cpu_selector_t cpu_selector_v{};

namespace property {
namespace queue {
struct enable_profiling {};
struct in_order {};
} // namespace queue
} // namespace property

namespace info {
namespace device {
struct name {};
struct vendor {};
struct max_compute_units {};
struct sub_group_sizes {};
struct max_work_group_size {};
struct max_num_sub_groups {};
struct local_mem_size {};
struct global_mem_size {};
} // namespace device
namespace platform {
struct name {};
} // namespace platform
namespace event_profiling {
struct command_submit {};
struct command_start {};
struct command_end {};
} // namespace event_profiling
} // namespace info

class platform {
public:
  template<typename Param>
  std::string get_info() const {
    static_assert(std::is_same_v<Param, info::platform::name>, "facade: only info::platform::name");
    return "HIP";
  }
};

class device {
public:
  device() = default;
  explicit device(sg_device_t h) : _h(h) {}

  template<typename Param>
  auto get_info() const {
    sg_device_info i{};
    detail::check(sg_device_info_get(_h, &i));
    if constexpr (std::is_same_v<Param, info::device::name>) {
      return std::string(i.name[0] ? i.name : i.arch);
    } else if constexpr (std::is_same_v<Param, info::device::vendor>) {
      return std::string("AMD");
    } else if constexpr (std::is_same_v<Param, info::device::max_compute_units>) {
      return static_cast<uint32_t>(i.compute_units);
    } else if constexpr (std::is_same_v<Param, info::device::sub_group_sizes>) {
      return std::vector<size_t>{static_cast<size_t>(i.wave_size)};
    } else if constexpr (std::is_same_v<Param, info::device::max_work_group_size>) {
      return static_cast<size_t>(i.max_workgroup);
    } else if constexpr (std::is_same_v<Param, info::device::max_num_sub_groups>) {
      return static_cast<uint32_t>(i.max_waves_per_wg);
    } else if constexpr (std::is_same_v<Param, info::device::local_mem_size>) {
      return static_cast<uint64_t>(i.lds_bytes);
    } else {
      static_assert(std::is_same_v<Param, info::device::global_mem_size>, "facade: unsupported device query");
      return static_cast<uint64_t>(i.global_mem_bytes);
    }
  }
  platform get_platform() const { return platform{}; }
  bool is_gpu() const { return true; }
  sg_device_t handle() const { return _h; }

private:
  sg_device_t _h = nullptr;
};

/* A completed (or pending) piece of queue work, timed by two hipEvents. */
class event {
public:
  event() = default;

  /* record the start on `stream` now; call finish() after enqueuing the work */
  static event begin(sg_device_t dev, hipStream_t stream, hipEvent_t origin) {
    event e;
    e._s = std::make_shared<State>();
    e._s->dev = dev;
    e._s->stream = stream;
    e._s->origin = origin;
    (void)hipEventCreate(&e._s->start);
    (void)hipEventCreate(&e._s->stop);
    (void)hipEventRecord(e._s->start, stream);
    return e;
  }
  void finish() {
    if (_s) (void)hipEventRecord(_s->stop, _s->stream);
  }

  void wait() {
    if (_s && hipEventSynchronize(_s->stop) != hipSuccess) throw exception("event wait failed");
  }
  void wait_and_throw() {
    wait();
    if (_s) detail::check(sg_device_synchronize(_s->dev));
  }

  /* nanoseconds since the queue was created */
  template<typename Param>
  uint64_t get_profiling_info() const {
    if (!_s) return 0;
    hipEvent_t which = std::is_same_v<Param, info::event_profiling::command_end> ? _s->stop : _s->start;
    (void)hipEventSynchronize(_s->stop);
    float ms = 0.f;
    if (_s->origin && hipEventElapsedTime(&ms, _s->origin, which) != hipSuccess) ms = 0.f;
    return static_cast<uint64_t>(static_cast<double>(ms) * 1e6);
  }

private:
  struct State {
    sg_device_t dev = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t origin = nullptr, start = nullptr, stop = nullptr;
    ~State() {
      if (start) (void)hipEventDestroy(start);
      if (stop) (void)hipEventDestroy(stop);
    }
  };
  std::shared_ptr<State> _s;
};

/* One-dimensional index space and item (tests/formats/csr.cpp:18-22 and
 * tests/frontier/bitmap_frontier.cpp:17-20 submit such kernels). */
template<int D = 1>
class range {
  static_assert(D == 1, "facade: one-dimensional ranges only");

public:
  __host__ __device__ range(size_t n) : _n(n) {}
  __host__ __device__ size_t size() const { return _n; }
  __host__ __device__ size_t operator[](int) const { return _n; }

private:
  size_t _n;
};

template<int D = 1>
class id {
  static_assert(D == 1, "facade: one-dimensional ids only");

public:
  __host__ __device__ id(size_t i = 0) : _i(i) {}
  __host__ __device__ operator size_t() const { return _i; }
  __host__ __device__ size_t operator[](int) const { return _i; }
  __host__ __device__ size_t get(int) const { return _i; }

private:
  size_t _i;
};

namespace detail {
template<typename F>
__global__ void __launch_bounds__(256) range_kernel(size_t n, F f) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) f(id<1>(i));
}
template<typename F>
__global__ void single_task_kernel(F f) {
  f();
}
} // namespace detail

/* Command-group handler: kernels are HIP launches on the queue's stream. */
class handler {
public:
  explicit handler(hipStream_t s, unsigned max_blocks) : _s(s), _max_blocks(max_blocks) {}
  template<typename F>
  void parallel_for(range<1> r, F f) {
    if (r.size() == 0) return;
    const size_t want = (r.size() + 255) / 256;
    const unsigned blocks = (unsigned)(want < _max_blocks ? want : _max_blocks);
    detail::range_kernel<<<blocks, 256, 0, _s>>>(r.size(), f);
    if (hipGetLastError() != hipSuccess) throw exception("parallel_for: kernel launch failed");
  }
  template<typename F>
  void single_task(F f) {
    detail::single_task_kernel<<<1, 1, 0, _s>>>(f);
    if (hipGetLastError() != hipSuccess) throw exception("single_task: kernel launch failed");
  }

private:
  hipStream_t _s;
  unsigned _max_blocks;
};

class queue {
public:
  queue() : queue(gpu_selector_v) {}
  explicit queue(gpu_selector_t) { init(); }
  explicit queue(default_selector_t) { init(); }
  queue(gpu_selector_t, property::queue::enable_profiling) { init(); }
  queue(default_selector_t, property::queue::enable_profiling) { init(); }
  queue(gpu_selector_t, property::queue::in_order) { init(); }

  device get_device() const { return device(_dev.get()); }
  sg_device_t handle() const { return _dev.get(); }
  hipStream_t stream() const {
    void* s = nullptr;
    detail::check(sg_device_stream(_dev.get(), &s));
    return static_cast<hipStream_t>(s);
  }
  hipEvent_t origin() const { return _origin ? _origin->ev : nullptr; }
  unsigned max_blocks() const {
    sg_device_info i{};
    detail::check(sg_device_info_get(_dev.get(), &i));
    return (unsigned)(i.compute_units > 0 ? i.compute_units * 8 : 2048);
  }

  void wait() { detail::check(sg_device_synchronize(_dev.get())); }

  template<typename CGF>
  event submit(CGF&& cgf) {
    event e = event::begin(handle(), stream(), origin());
    handler h(stream(), max_blocks());
    cgf(h);
    e.finish();
    return e;
  }
  template<typename F>
  event parallel_for(range<1> r, F f) {
    return submit([&](handler& h) { h.parallel_for(r, f); });
  }
  void wait_and_throw() { wait(); }

  template<typename T>
  event fill(T* ptr, const T& value, size_t count) {
    static_assert(sizeof(T) == 1 || sizeof(T) == 4 || sizeof(T) == 8, "facade: fill supports 1/4/8-byte types");
    event e = event::begin(handle(), stream(), origin());
    uint64_t pattern = 0;
    std::memcpy(&pattern, &value, sizeof(T));
    detail::check(sg_fill(handle(), ptr, pattern, (int)sizeof(T), count));
    e.finish();
    return e;
  }
  template<typename T>
  event copy(const T* src, T* dst, size_t count) {
    event e = event::begin(handle(), stream(), origin());
    detail::check(sg_memcpy(handle(), dst, src, count * sizeof(T)));
    e.finish();
    return e;
  }
  event memcpy(void* dst, const void* src, size_t bytes) {
    event e = event::begin(handle(), stream(), origin());
    detail::check(sg_memcpy(handle(), dst, src, bytes));
    e.finish();
    return e;
  }
  event prefetch(const void* ptr, size_t bytes) {
    event e = event::begin(handle(), stream(), origin());
    detail::check(sg_prefetch(handle(), ptr, bytes));
    e.finish();
    return e;
  }
  event memset(void* dst, int value, size_t bytes) {
    event e = event::begin(handle(), stream(), origin());
    detail::check(sg_fill(handle(), dst, (uint64_t)(uint8_t)value, 1, bytes));
    e.finish();
    return e;
  }

private:
  struct Origin {
    hipEvent_t ev = nullptr;
    ~Origin() {
      if (ev) (void)hipEventDestroy(ev);
    }
  };
  void init() {
    _dev = detail::shared_device(0);
    _origin = std::make_shared<Origin>();
    (void)hipEventCreate(&_origin->ev);
    (void)hipEventRecord(_origin->ev, stream());
  }
  std::shared_ptr<sg_device_s> _dev;
  std::shared_ptr<Origin> _origin;
};

/* USM allocation (utils/memory.hpp:12-39 calls these). */
template<typename T>
T* malloc_device(size_t n, const queue& q) {
  void* p = nullptr;
  detail::check(sg_alloc(q.handle(), n * sizeof(T), SG_SPACE_DEVICE, &p));
  return static_cast<T*>(p);
}
template<typename T>
T* malloc_shared(size_t n, const queue& q) {
  void* p = nullptr;
  detail::check(sg_alloc(q.handle(), n * sizeof(T), SG_SPACE_SHARED, &p));
  return static_cast<T*>(p);
}
template<typename T>
T* malloc_host(size_t n, const queue& q) {
  void* p = nullptr;
  detail::check(sg_alloc(q.handle(), n * sizeof(T), SG_SPACE_HOST, &p));
  return static_cast<T*>(p);
}
inline void free(void* ptr, const queue& q) { detail::check(sg_free(q.handle(), ptr)); }

/* Reduction operators named by compute::reduce (operators/for/for.hpp:25-27). */
template<typename T = void>
struct plus {
  __host__ __device__ T operator()(const T& a, const T& b) const { return a + b; }
  __host__ __device__ static T identity() { return T(0); }
};
template<typename T = void>
struct multiplies {
  __host__ __device__ T operator()(const T& a, const T& b) const { return a * b; }
  __host__ __device__ static T identity() { return T(1); }
};
template<typename T = void>
struct minimum {
  __host__ __device__ T operator()(const T& a, const T& b) const { return b < a ? b : a; }
  __host__ __device__ static T identity() { return std::numeric_limits<T>::max(); }
};
template<typename T = void>
struct maximum {
  __host__ __device__ T operator()(const T& a, const T& b) const { return a < b ? b : a; }
  __host__ __device__ static T identity() { return std::numeric_limits<T>::lowest(); }
};

} // namespace sycl
