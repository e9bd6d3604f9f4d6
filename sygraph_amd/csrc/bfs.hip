/*
 * bfs.hip -- level-synchronous BFS for gfx950 (BFS::run,
 * algorithms/bfs.hpp:153-238, with the push / pull functors of :180-209 fused
 * into the advance kernels and the hybrid heuristic of :284-350).
 *
 * Frontier representation follows the direction of the level:
 *  - push levels read a vertex queue split at enqueue time into three degree
 *    classes (the thread / wavefront / workgroup bins of the bucketing load
 *    balancer, operators/advance/bucketing.hpp:100-178): a lane per small
 *    vertex, a wave per medium vertex and a workgroup per 4096-edge chunk of a
 *    large vertex.  Every queue entry carries (vertex, first edge, length), so a
 *    level reads the offsets of a frontier vertex once, when it is discovered.
 *  - pull levels read the frontier as a bitmap; one wavefront owns one 64-bit
 *    word of the visited bitmap, lane i pulls for vertex 64w+i over its
 *    in-edges and stops at the first in-neighbour in the frontier (the
 *    short-circuit of workgroup_mapped.hpp:151-175).  The new frontier word is
 *    the wave's ballot, stored without atomics.
 * The visited bitmap (2 MiB at 2^24 vertices) stays L2/MALL resident; the
 * depth array is written once per discovered vertex.  Discovery is an atomic
 * test-and-set of the visited bit, so depths are unique and bit-exact.
 *
 * Level counters live on the device in a ring of three slots: level L reads
 * the queue counts written during L-1, writes slot L%3 and clears slot
 * (L+1)%3, so no host work is needed between levels except the direction
 * decision.
 */
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

using namespace sygraph::hipdev;

namespace {

constexpr int kBlock = 256;

struct alignas(16) QEntry {
  int64_t start;
  int32_t v;
  int32_t len;
};

struct LevelCtr {
  unsigned long long discovered;
  unsigned long long disc_deg;
  unsigned long long edges;
  unsigned long long scanned;
  unsigned long long max_deg;
  uint32_t q[3]; /* entries enqueued per class: small, medium, large */
  uint32_t pad;
};

struct BfsTuning {
  int32_t t_medium; /* degree >= t_medium -> wave per vertex */
  int32_t t_large;  /* degree >= t_large -> workgroup per chunk */
  int32_t chunk;    /* edges per large-class entry */
};

template<typename OffT>
struct BfsArgs {
  const OffT* off;      /* forward CSR (push, queue entries, degrees) */
  const int32_t* cols;
  const OffT* in_off;   /* inverse CSR (pull); aliases off when undirected */
  const int32_t* in_cols;
  int32_t directed;
  int64_t n;
  int64_t w0;
  uint64_t* visited;
  uint64_t* cur_bm;
  uint64_t* next_bm;
  int32_t* depth;
  QEntry* qin[3];
  QEntry* qout[3];
  const LevelCtr* ctr_in;
  LevelCtr* ctr_out;
  LevelCtr* ctr_zero;
  int32_t level;
  BfsTuning tune;
};

struct LaneStats {
  uint64_t disc = 0, disc_deg = 0, edges = 0, scanned = 0, max_deg = 0;
};

__device__ __forceinline__ void flush_stats(LevelCtr* c, LaneStats& s) {
  uint64_t a = wave_sum_u64(s.disc);
  uint64_t b = wave_sum_u64(s.disc_deg);
  uint64_t e = wave_sum_u64(s.edges);
  uint64_t sc = wave_sum_u64(s.scanned);
  uint64_t mx = wave_max_u64(s.max_deg);
  if (lane_id() == 0) {
    if (a) atomic_add(&c->discovered, (unsigned long long)a);
    if (b) atomic_add(&c->disc_deg, (unsigned long long)b);
    if (e) atomic_add(&c->edges, (unsigned long long)e);
    if (sc) atomic_add(&c->scanned, (unsigned long long)sc);
    if (mx) atomic_max(&c->max_deg, (unsigned long long)mx);
  }
}

__device__ __forceinline__ void zero_ctr(LevelCtr* c) {
  if (blockIdx.x == 0 && threadIdx.x < sizeof(LevelCtr) / 4) reinterpret_cast<uint32_t*>(c)[threadIdx.x] = 0u;
}

/* Enqueue the vertices this wave discovered (want) into the degree classes.
 * Whole wave must call. deg/start valid where want. */
template<typename OffT>
__device__ __forceinline__ void enqueue_discovered(const BfsTuning& t, QEntry* const q[3], LevelCtr* c, bool want,
                                                   uint32_t v, int64_t start, int64_t deg) {
  const int cls = deg >= t.t_large ? 2 : (deg >= t.t_medium ? 1 : 0);
  const bool has_work = want && deg > 0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    bool w = has_work && cls == k;
    uint32_t slot = wave_append(&c->q[k], w);
    if (w) {
      QEntry qe;
      qe.start = start;
      qe.v = (int32_t)v;
      qe.len = (int32_t)deg;
      q[k][slot] = qe;
    }
  }
  if (has_work && cls == 2) {
    uint32_t nch = (uint32_t)((deg + t.chunk - 1) / t.chunk);
    uint32_t base = atomic_add(&c->q[2], nch);
    for (uint32_t i = 0; i < nch; ++i) {
      QEntry qe;
      qe.start = start + (int64_t)i * t.chunk;
      qe.v = (int32_t)v;
      int64_t rem = deg - (int64_t)i * t.chunk;
      qe.len = (int32_t)(rem < t.chunk ? rem : t.chunk);
      q[2][base + i] = qe;
    }
  }
}

/* Visit arc (u -> y) in push mode; returns true if y was discovered here. */
template<typename OffT>
__device__ __forceinline__ bool push_visit(const BfsArgs<OffT>& a, uint32_t y, int64_t* start, int64_t* deg) {
  if (!claim_bit(a.visited, y)) return false;
  a.depth[y] = a.level + 1;
  OffT b = a.off[y], e = a.off[y + 1];
  *start = (int64_t)b;
  *deg = (int64_t)(e - b);
  return true;
}

template<typename OffT>
__global__ void __launch_bounds__(kBlock) push_kernel(BfsArgs<OffT> a) {
  zero_ctr(a.ctr_zero);
  const uint32_t n_small = a.ctr_in->q[0], n_med = a.ctr_in->q[1], n_large = a.ctr_in->q[2];
  LaneStats st;
  const int lane = lane_id();
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;

  /* workgroup bin: one entry (<= chunk edges of a large vertex) per workgroup */
  for (uint32_t i = blockIdx.x; i < n_large; i += gridDim.x) {
    QEntry qe = a.qin[2][i];
    for (int64_t j0 = wave_in_block * 64; j0 < qe.len; j0 += kBlock) {
      int64_t j = j0 + lane;
      bool act = j < qe.len;
      int64_t s = 0, d = 0;
      uint32_t y = 0;
      bool disc = false;
      if (act) {
        y = (uint32_t)a.cols[qe.start + j];
        disc = push_visit(a, y, &s, &d);
      }
      if (disc) {
        st.disc++;
        st.disc_deg += d;
        st.max_deg = d > (int64_t)st.max_deg ? d : st.max_deg;
      }
      enqueue_discovered<OffT>(a.tune, a.qout, a.ctr_out, disc, y, s, d);
    }
  }
  /* wavefront bin: one medium vertex per wave, 64 consecutive arcs per step */
  for (int64_t i = gwave; i < n_med; i += nwaves) {
    QEntry qe = a.qin[1][i];
    for (int64_t j0 = 0; j0 < qe.len; j0 += 64) {
      int64_t j = j0 + lane;
      bool act = j < qe.len;
      int64_t s = 0, d = 0;
      uint32_t y = 0;
      bool disc = false;
      if (act) {
        y = (uint32_t)a.cols[qe.start + j];
        disc = push_visit(a, y, &s, &d);
      }
      if (disc) {
        st.disc++;
        st.disc_deg += d;
        st.max_deg = d > (int64_t)st.max_deg ? d : st.max_deg;
      }
      enqueue_discovered<OffT>(a.tune, a.qout, a.ctr_out, disc, y, s, d);
    }
  }
  /* thread bin: one small vertex per lane, arcs walked serially */
  for (int64_t base = gwave * 64; base < n_small; base += nwaves * 64) {
    int64_t i = base + lane;
    QEntry qe;
    qe.start = 0;
    qe.len = 0;
    qe.v = 0;
    if (i < n_small) qe = a.qin[0][i];
    for (int32_t j = 0;; ++j) {
      bool act = j < qe.len;
      if (ballot(act) == 0) break;
      int64_t s = 0, d = 0;
      uint32_t y = 0;
      bool disc = false;
      if (act) {
        y = (uint32_t)a.cols[qe.start + j];
        disc = push_visit(a, y, &s, &d);
      }
      if (disc) {
        st.disc++;
        st.disc_deg += d;
        st.max_deg = d > (int64_t)st.max_deg ? d : st.max_deg;
      }
      enqueue_discovered<OffT>(a.tune, a.qout, a.ctr_out, disc, y, s, d);
    }
  }
  flush_stats(a.ctr_out, st);
}

/* Pull level: a wave per visited-bitmap word; lanes of unvisited vertices
 * scan in-edges until the first in-neighbour in the frontier. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) pull_kernel(BfsArgs<OffT> a) {
  zero_ctr(a.ctr_zero);
  LaneStats st;
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t w = gwave; w < a.w0; w += nwaves) {
    const uint64_t vis = a.visited[w];
    if (vis == ~0ull) {
      if (lane == 0) a.next_bm[w] = 0ull;
      continue;
    }
    const int64_t v = w * 64 + lane;
    const bool active = !((vis >> lane) & 1ull) && v < a.n;
    bool found = false;
    if (active) {
      const OffT b = a.in_off[v], e = a.in_off[v + 1];
      const int64_t deg = (int64_t)(e - b);
      int64_t k = deg;
      for (OffT j = b; j < e; ++j) {
        const uint32_t u = (uint32_t)a.in_cols[j];
        if (test_bit(a.cur_bm, u)) {
          found = true;
          k = (int64_t)(j - b) + 1;
          break;
        }
      }
      if (deg > 0) st.scanned++;
      st.edges += k;
      if (found) {
        a.depth[v] = a.level + 1;
        const int64_t odeg = a.directed ? (int64_t)(a.off[v + 1] - a.off[v]) : deg;
        st.disc++;
        st.disc_deg += odeg;
        st.max_deg = odeg > (int64_t)st.max_deg ? odeg : st.max_deg;
      }
    }
    const uint64_t nb = ballot(found);
    if (lane == 0) {
      a.next_bm[w] = nb;
      if (nb) a.visited[w] = vis | nb;
    }
  }
  flush_stats(a.ctr_out, st);
}

/* Frontier queue -> bitmap (push level followed by a pull level). */
__global__ void __launch_bounds__(kBlock) queue_to_bitmap_kernel(QEntry* const* q, const LevelCtr* ctr, uint64_t* bm) {
  for (int k = 0; k < 3; ++k) {
    const uint32_t cnt = ctr->q[k];
    const QEntry* qq = q[k];
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * kBlock) {
      uint32_t v = (uint32_t)qq[i].v;
      atomic_or(&bm[v >> 6], 1ull << (v & 63));
    }
  }
}

/* Frontier bitmap -> class queues (pull level followed by a push level). */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) bitmap_to_queue_kernel(BfsArgs<OffT> a, const uint64_t* bm, LevelCtr* ctr) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t w = gwave; w < a.w0; w += nwaves) {
    const uint64_t word = bm[w];
    if (word == 0) continue;
    const bool has = (word >> lane) & 1ull;
    int64_t s = 0, d = 0;
    const uint32_t v = (uint32_t)(w * 64 + lane);
    if (has) {
      OffT b = a.off[v], e = a.off[v + 1];
      s = (int64_t)b;
      d = (int64_t)(e - b);
    }
    enqueue_discovered<OffT>(a.tune, a.qout, ctr, has, v, s, d);
  }
}

/* Level-0 setup: depth[src] = 0, visited = isolated | {src}, F0 = {src}. */
template<typename OffT>
__global__ void init_source_kernel(BfsArgs<OffT> a, int32_t src, LevelCtr* ctr_src) {
  if (blockIdx.x != 0 || threadIdx.x >= 64) return;
  const bool me = lane_id() == 0;
  int64_t s = 0, d = 0;
  if (me) {
    a.depth[src] = 0;
    a.visited[src >> 6] |= 1ull << (src & 63);
    a.cur_bm[src >> 6] |= 1ull << (src & 63);
    s = (int64_t)a.off[src];
    d = (int64_t)(a.off[src + 1] - a.off[src]);
  }
  enqueue_discovered<OffT>(a.tune, a.qout, ctr_src, me, (uint32_t)src, s, d);
}

__global__ void copy_words_kernel(uint64_t* dst, const uint64_t* src, int64_t nw) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

/* Per-graph BFS workspace, allocated on first use and reused. */
struct BfsWorkspace {
  int64_t n = 0, w0 = 0, cap_large = 0;
  uint64_t* visited = nullptr;
  uint64_t* bm[2] = {nullptr, nullptr};
  QEntry* q[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  QEntry** qptr_dev = nullptr; /* device copies of q[s][0..2] for queue_to_bitmap */
  LevelCtr* ctr = nullptr;     /* 3 slots */
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

void ws_free(BfsWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->visited);
  (void)hipFree(w->bm[0]);
  (void)hipFree(w->bm[1]);
  for (auto& s : w->q)
    for (auto* p : s) (void)hipFree(p);
  (void)hipFree(w->qptr_dev);
  (void)hipFree(w->ctr);
  for (auto e : w->ev)
    if (e) (void)hipEventDestroy(e);
  delete w;
}

BfsTuning default_tuning() { return BfsTuning{64, 4096, 4096}; }

} // namespace

namespace {
int workspace_for(sg_graph_t g, BfsWorkspace** out) {
  if (g->bfs_ws) {
    *out = static_cast<BfsWorkspace*>(g->bfs_ws);
    return SG_OK;
  }
  auto* w = new BfsWorkspace();
  w->n = g->n;
  w->w0 = (g->n + 63) / 64;
  const BfsTuning t = default_tuning();
  w->cap_large = g->m / t.chunk + g->m / t.t_large + 64;
  int st = SG_OK;
  sg_device_t dev = g->dev;
  if ((st = sg::alloc_device(dev, (void**)&w->visited, 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->bm[0], 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->bm[1], 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->ctr, sizeof(LevelCtr) * 3)) ||
      (st = sg::alloc_device(dev, (void**)&w->qptr_dev, sizeof(QEntry*) * 6))) {
    ws_free(w);
    return st;
  }
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 3; ++k) {
      int64_t cap = k == 2 ? w->cap_large : g->n + 1;
      if ((st = sg::alloc_device(dev, (void**)&w->q[s][k], sizeof(QEntry) * cap))) {
        ws_free(w);
        return st;
      }
    }
  QEntry* hp[6] = {w->q[0][0], w->q[0][1], w->q[0][2], w->q[1][0], w->q[1][1], w->q[1][2]};
  bool ok = hipMemcpy(w->qptr_dev, hp, sizeof(hp), hipMemcpyHostToDevice) == hipSuccess;
  for (auto& e : w->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  if (!ok) {
    ws_free(w);
    sg::set_error("bfs workspace: pointer upload failed");
    return SG_ERR_HIP;
  }
  g->bfs_ws = w;
  *out = w;
  return SG_OK;
}
} // namespace

namespace sg {
void bfs_workspace_release(sg_graph_t g) {
  if (g && g->bfs_ws) {
    ws_free(static_cast<BfsWorkspace*>(g->bfs_ws));
    g->bfs_ws = nullptr;
  }
}
} // namespace sg

template<typename OffT>
int run_bfs(sg_graph_t g, int32_t src, const sg_bfs_options& opt, int32_t* depth, sg_bfs_stats* stats) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  BfsWorkspace* ws = nullptr;
  SG_TRY(workspace_for(g, &ws));
  SG_TRY(sg::graph_ensure_isolated(g));
  const int64_t n = g->n, w0 = ws->w0;
  const BfsTuning tune = default_tuning();
  const int grid_main = dev->cus * 8;
  const int grid_small = std::max(1, std::min(dev->cus * 4, (int)((w0 + 255) / 256)));

  BfsArgs<OffT> a{};
  a.off = (const OffT*)g->offsets;
  a.cols = g->cols;
  a.in_off = (const OffT*)g->inv_offsets;
  a.in_cols = g->inv_cols;
  a.directed = (g->flags & SG_GRAPH_DIRECTED) ? 1 : 0;
  a.n = n;
  a.w0 = w0;
  a.visited = ws->visited;
  a.depth = depth;
  a.tune = tune;
  int cb = 0; /* current bitmap index */

  /* setup: depth = n+1, visited = isolated, cur_bm = 0, counters = 0 */
  fill_i32_kernel<<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(depth, (int32_t)(n + 1), n);
  copy_words_kernel<<<grid_small, 256, 0, s>>>(ws->visited, g->isolated, w0);
  SG_TRY(sg::zero_async(dev, ws->bm[0], 8 * w0));
  SG_TRY(sg::zero_async(dev, ws->ctr, sizeof(LevelCtr) * 3));
  /* F0 = {src}: queues of set 0, counts in ring slot 2 (read by level 0) */
  a.cur_bm = ws->bm[cb];
  for (int k = 0; k < 3; ++k) a.qout[k] = ws->q[0][k];
  init_source_kernel<OffT><<<1, 64, 0, s>>>(a, src, ws->ctr + 2);
  SG_HIP(hipGetLastError());

  const bool want_events = opt.record_levels != 0 || stats != nullptr;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  if (want_events) {
    SG_HIP(hipEventCreate(&ev_begin));
    SG_HIP(hipEventCreate(&ev_end));
    SG_HIP(hipEventRecord(ev_begin, s));
  }

  bool push = opt.policy != SG_POLICY_PULL;
  int32_t level = 0;
  uint64_t visited_deg = 0;
  LevelCtr* hctr = (LevelCtr*)dev->host_scratch;
  int32_t n_push = 0, n_pull = 0;
  {
    /* degree of the source counts as explored */
    OffT hb[2];
    SG_HIP(hipMemcpyAsync(hb, (const OffT*)g->offsets + src, 2 * sizeof(OffT), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    visited_deg = (uint64_t)(hb[1] - hb[0]);
  }
  int64_t f_vertices = 1, f_edges = (int64_t)visited_deg;
  int64_t reached_v = 1;
  for (;; ++level) {
    const int qs = level & 1;
    a.level = level;
    a.ctr_in = ws->ctr + ((level + 2) % 3);
    a.ctr_out = ws->ctr + (level % 3);
    a.ctr_zero = ws->ctr + ((level + 1) % 3);
    for (int k = 0; k < 3; ++k) {
      a.qin[k] = ws->q[qs][k];
      a.qout[k] = ws->q[qs ^ 1][k];
    }
    a.cur_bm = ws->bm[cb];
    a.next_bm = ws->bm[cb ^ 1];
    const bool timed = opt.record_levels && stats && level < SG_MAX_LEVELS_RECORDED;
    if (timed) SG_HIP(hipEventRecord(ws->ev[0], s));
    if (push) {
      push_kernel<OffT><<<grid_main, kBlock, 0, s>>>(a);
      ++n_push;
    } else {
      pull_kernel<OffT><<<grid_main, kBlock, 0, s>>>(a);
      ++n_pull;
    }
    SG_HIP(hipGetLastError());
    if (timed) SG_HIP(hipEventRecord(ws->ev[1], s));
    SG_HIP(hipMemcpyAsync(hctr, a.ctr_out, sizeof(LevelCtr), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    const LevelCtr c = *hctr;
    if (timed) {
      float ms = 0.f;
      SG_HIP(hipEventElapsedTime(&ms, ws->ev[0], ws->ev[1]));
      stats->level[level].kernel_ms = ms;
      stats->level[level].max_degree = (int64_t)c.max_deg;
    }
    if (stats && level < SG_MAX_LEVELS_RECORDED) {
      sg_level_stat& ls = stats->level[level];
      ls.direction = push ? 0 : 1;
      ls.load_balancer = opt.load_balancer;
      ls.frontier_vertices = f_vertices;
      ls.frontier_edges = f_edges;
      ls.discovered = (int64_t)c.discovered;
      ls.edges_inspected = push ? f_edges : (int64_t)c.edges;
      ls.vertices_scanned = push ? f_vertices : (int64_t)c.scanned;
    }
    visited_deg += c.disc_deg;
    reached_v += (int64_t)c.discovered;
    const uint64_t unexplored = (uint64_t)g->total_degree - visited_deg;
    if (stats && level < SG_MAX_LEVELS_RECORDED) stats->level[level].unexplored_edges = (int64_t)unexplored;
    f_vertices = (int64_t)c.discovered;
    f_edges = (int64_t)c.disc_deg;
    if (c.discovered == 0) {
      ++level;
      break;
    }
    /* direction for the next level (evaluateHeuristic, bfs.hpp:341-350) */
    bool next_push = push;
    if (opt.policy == SG_POLICY_HYBRID) {
      if (push) {
        const uint32_t fdeg = (uint32_t)c.disc_deg, udeg = (uint32_t)unexplored;
        if ((float)fdeg > ((float)udeg / opt.alpha)) next_push = false;
      } else {
        if ((float)c.discovered < ((float)n / opt.beta)) next_push = true;
      }
    }
    if (timed && push != next_push) SG_HIP(hipEventRecord(ws->ev[2], s));
    if (push && !next_push) {
      /* F_{L+1} is in queues set qs^1 with counts in slot L%3 -> bitmap */
      QEntry** qp = ws->qptr_dev + 3 * (qs ^ 1);
      SG_TRY(sg::zero_async(dev, ws->bm[cb], 8 * w0));
      queue_to_bitmap_kernel<<<grid_small, kBlock, 0, s>>>(qp, a.ctr_out, ws->bm[cb]);
      /* cb now holds F_{L+1}; the pull at L+1 reads bm[cb] */
    } else if (!push && next_push) {
      /* F_{L+1} is bm[cb^1]; build queues set qs^1 with counts in slot L%3 */
      bitmap_to_queue_kernel<OffT><<<grid_main, kBlock, 0, s>>>(a, ws->bm[cb ^ 1], a.ctr_out);
      cb ^= 1;
    } else if (!push) {
      cb ^= 1;
    }
    SG_HIP(hipGetLastError());
    if (timed && push != next_push) {
      SG_HIP(hipEventRecord(ws->ev[3], s));
      SG_HIP(hipEventSynchronize(ws->ev[3]));
      float ms = 0.f;
      SG_HIP(hipEventElapsedTime(&ms, ws->ev[2], ws->ev[3]));
      stats->level[level].convert_ms = ms;
    }
    push = next_push;
  }
  if (want_events) {
    SG_HIP(hipEventRecord(ev_end, s));
    SG_HIP(hipEventSynchronize(ev_end));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, ev_begin, ev_end));
    (void)hipEventDestroy(ev_begin);
    (void)hipEventDestroy(ev_end);
    if (stats) stats->kernel_ms = ms;
  }
  SG_HIP(hipStreamSynchronize(s));
  if (stats) {
    stats->levels = level;
    stats->n_push = n_push;
    stats->n_pull = n_pull;
    stats->reached_vertices = reached_v;
    stats->reached_edges = (int64_t)visited_deg;
  }
  return SG_OK;
}

extern "C" {

void sg_bfs_options_default(sg_bfs_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->policy = SG_POLICY_PUSH;
  opt->load_balancer = SG_LB_BUCKETING;
  opt->alpha = 15.0f;
  opt->beta = 24.0f;
}

int sg_bfs(sg_graph_t g, int32_t source, const sg_bfs_options* opt_in, int32_t* depth, sg_bfs_stats* stats) {
  SG_CHECK(g && depth, SG_ERR_INVALID, "sg_bfs: null argument");
  SG_CHECK(source >= 0 && source < g->n, SG_ERR_INVALID, "sg_bfs: source %d out of range [0,%lld)", source,
           (long long)g->n);
  sg_bfs_options opt;
  if (opt_in)
    opt = *opt_in;
  else
    sg_bfs_options_default(&opt);
  SG_CHECK(opt.policy >= SG_POLICY_PUSH && opt.policy <= SG_POLICY_HYBRID, SG_ERR_UNSUPPORTED,
           "sg_bfs: policy %d not supported", opt.policy);
  SG_CHECK(opt.policy != SG_POLICY_HYBRID || (opt.alpha > 0.f && opt.beta > 0.f), SG_ERR_INVALID,
           "sg_bfs: alpha and beta must be positive");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return run_bfs<int64_t>(g, source, opt, depth, stats);
  return run_bfs<uint32_t>(g, source, opt, depth, stats);
}

} // extern "C"
