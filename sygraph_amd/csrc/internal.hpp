/*
 * internal.hpp -- private definitions of libsygraph_hip (not installed).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <atomic>
#include <map>
#include <string>
#include <vector>

#include "sygraph_hip.h"
#include "sygraph/detail/hip_device.hpp"

namespace sg {

void set_error(const char* fmt, ...);

/* Evaluate a HIP call; on failure record the message and return SG_ERR_HIP. */
#define SG_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      ::sg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return SG_ERR_HIP;                                                                       \
    }                                                                                          \
  } while (0)

#define SG_CHECK(cond, code, ...)  \
  do {                             \
    if (!(cond)) {                 \
      ::sg::set_error(__VA_ARGS__); \
      return (code);               \
    }                              \
  } while (0)

#define SG_TRY(expr)          \
  do {                        \
    int s_ = (expr);          \
    if (s_ != SG_OK) return s_; \
  } while (0)

} // namespace sg

struct sg_device_s {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
  int cus = 0;
  /* small pinned host scratch for stat read-backs */
  void* host_scratch = nullptr;
  size_t host_scratch_bytes = 0;
  /* bound on a grid barrier's wait inside a cooperative launch, in ticks of
   * the device's constant-rate clock (wall_clock64); past it the launch ends
   * with an error instead of waiting on */
  uint64_t wait_ticks = 0;
  int64_t wall_khz = 0;
  /* this device's copy of sg_cc_noalloc's static scratch array (cc.hip) */
  void* cc_scratch = nullptr;
  hipError_t cc_scratch_error = hipSuccess;
  /* sg_device_close and every graph, frontier and event on the device hold
   * one reference; the stream and scratch go with the last one, so handles
   * may be released in any order */
  std::atomic<int> refs{1};
};

namespace sg {
void device_release(sg_device_t dev);
void cc_scratch_init(sg_device_t dev);

/* Owning reference to a device, held by graphs, frontiers and events. */
class DevRef {
public:
  DevRef() = default;
  DevRef(const DevRef&) = delete;
  DevRef& operator=(const DevRef&) = delete;
  ~DevRef() { reset(); }
  DevRef& operator=(sg_device_t d) {
    if (d) d->refs.fetch_add(1, std::memory_order_relaxed);
    reset();
    p_ = d;
    return *this;
  }
  void reset() {
    if (p_) device_release(p_);
    p_ = nullptr;
  }
  operator sg_device_t() const { return p_; }
  sg_device_s* operator->() const { return p_; }

private:
  sg_device_t p_ = nullptr;
};
} // namespace sg

struct sg_event_s {
  sg::DevRef dev;
  hipEvent_t ev = nullptr;
};

struct sg_graph_s {
  sg::DevRef dev;
  int64_t n = 0, m = 0;
  int offset_bytes = 8;
  int flags = 0;
  void* offsets = nullptr;
  int32_t* cols = nullptr;
  void* weights = nullptr;
  void* inv_offsets = nullptr;
  int32_t* inv_cols = nullptr;
  void* inv_weights = nullptr;
  bool owns_inverse = false;
  /* lazily built helpers */
  uint64_t* isolated = nullptr; /* bit set for vertices with in-degree 0 */
  int2* pull_hint = nullptr;     /* per vertex: {in-neighbour of largest out-degree (pull probes it first),
                                    out-degree of the vertex}, or with pull_hint_offsets {that in-neighbour,
                                    low 32 bits of the start of the vertex's list} and entry n = {-1, m} */
  bool pull_hint_offsets = false;
  double mean_weight = -1.0;     /* mean arc weight, computed on first use (automatic SSSP delta) */
  /* SSSP's light-first copy of the adjacency for one bucket width: per vertex
   * the arcs lighter than the width first, then the others */
  int32_t* sssp_cols = nullptr;
  void* sssp_w = nullptr;
  int32_t* sssp_light = nullptr; /* light arcs per vertex */
  uint32_t sssp_split_bits = 0;  /* the width's bit pattern */
  bool sssp_split_valid = false;
  int64_t total_degree = 0;
  int symmetric = -1;           /* every arc (u,v) has its reverse (v,u): 1 / 0, -1 = not checked yet */
  void* bfs_ws = nullptr;       /* traversal workspaces, reused across calls */
  void* sssp_ws = nullptr;
  void* aux_queues = nullptr;   /* degree-class queue set shared by CC, multi-source BFS and batched BC */
  std::map<std::string, std::pair<void*, size_t>> buffers; /* named device buffers kept with the graph */
};

struct sg_frontier_s {
  sg::DevRef dev;
  int64_t n = 0, w0 = 0, w1 = 0;
  uint64_t* l0 = nullptr;
  uint64_t* l1 = nullptr;
  int32_t* list = nullptr;
  uint32_t* list_size = nullptr;
  uint32_t* tile_counts = nullptr; /* compaction scratch */
  int64_t ntiles = 0;
};

namespace sg {

inline int grid_for(sg_device_t dev, int64_t items, int block, int max_blocks_per_cu = 8) {
  int64_t want = (items + block - 1) / block;
  int64_t cap = static_cast<int64_t>(dev->cus) * max_blocks_per_cu;
  if (want < 1) want = 1;
  return static_cast<int>(want < cap ? want : cap);
}

int alloc_device(sg_device_t dev, void** p, size_t bytes);
int zero_async(sg_device_t dev, void* p, size_t bytes);
/* Wait until the device has stored `want` into *flag (pinned host memory),
 * spinning on it; should the stream drain without it, fail (SG_ERR_STATE,
 * "who" in the message) instead of spinning on. */
int wait_host_flag(const int32_t* flag, int32_t want, hipStream_t s, const char* who);

/* graph helpers */
int graph_ensure_isolated(sg_graph_t g);
int graph_ensure_pull_hint(sg_graph_t g);
const float* policy_device_params(sg_policy_t p); /* device parameters of a learned policy */
void bfs_workspace_release(sg_graph_t g);
void sssp_workspace_release(sg_graph_t g);
/* A device buffer of at least `bytes` kept with the graph under `name`
 * (grown, not preserved, when a call needs more); repeated calls on one graph
 * then allocate nothing.  Freed with the graph. */
int graph_buffer(sg_graph_t g, const char* name, size_t bytes, void** out);
template<typename T>
inline int graph_buffer(sg_graph_t g, const char* name, size_t bytes, T** out) {
  void* p = nullptr;
  const int st = graph_buffer(g, name, bytes, &p);
  *out = static_cast<T*>(p);
  return st;
}
void graph_buffers_release(sg_graph_t g);
void aux_queues_release(sg_graph_t g);

/* frontier compaction used by other modules */
int frontier_compact_launch(sg_frontier_t f, int mode);

} // namespace sg
