/*
 * bfs.hip -- level-synchronous BFS for gfx950 (BFS::run,
 * algorithms/bfs.hpp:153-238, with the push / pull functors of :180-209 fused
 * into the advance kernels and the hybrid heuristic of :284-350).
 *
 * The frontier between levels is always a bitmap (one uint64 word per
 * wavefront, lane i <-> bit i).  Per direction:
 *  - push levels read vertex queues split into three degree classes (the
 *    thread / wavefront / workgroup bins of the bucketing load balancer,
 *    operators/advance/bucketing.hpp:100-178): a lane per small vertex, a wave
 *    per medium vertex, a workgroup per 4096-edge chunk of a large vertex, so
 *    even a 400K-degree hub is spread over the whole grid.  Each queue entry
 *    carries (vertex, first edge, length).  Discovery is an atomic test-and-set
 *    of the visited bit (the only atomic of the level); the new frontier is
 *    recovered afterwards as visited & ~prev.
 *  - pull levels walk the visited bitmap 64 words per wave (one coalesced
 *    load), and for each word with unvisited vertices lane i scans the
 *    in-edges of vertex 64w+i, four ids per step, until the first in-neighbour
 *    in the frontier (the short circuit of workgroup_mapped.hpp:151-175).  The
 *    new frontier word is the wave's ballot, stored without atomics.
 *  - queues for a push level are built from the frontier bitmap by an ordered
 *    set-scan-compact (count per 64-word tile, one scan, write), so no queue
 *    counter is ever contended and queue order is deterministic.
 * Level statistics (for the direction heuristic and the byte model) are
 * reduced per workgroup and added into 64 cache-line-separated slots.
 */
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

using namespace sygraph::hipdev;

#include "traverse.hpp"

using namespace sg::trav;

namespace {

constexpr int kPullHead = 16; /* ids a lane scans alone before the wave helps */

template<typename OffT>
struct BfsArgs {
  const OffT* off; /* forward CSR (push, queue entries, degrees) */
  const int32_t* cols;
  const OffT* in_off; /* inverse CSR (pull); aliases off when undirected */
  const int32_t* in_cols;
  const int32_t* hint; /* per vertex: in-neighbour of largest out-degree */
  int32_t directed;
  int64_t n;
  int64_t w0;
  uint64_t* visited;
  uint64_t* cur_bm;
  uint64_t* next_bm;
  int32_t* depth;
  QEntry* qin[3];
  const QueueHdr* qin_hdr;
  StatSlot* stats;
  StatSlot* stats_zero;
  int32_t level;
};

template<typename OffT>
__device__ __forceinline__ void push_visit(const BfsArgs<OffT>& a, uint32_t y) {
  if (claim_bit(a.visited, y)) a.depth[y] = a.level + 1;
}

/* Push level: every arc of the class queues; discovery is the test-and-set of
 * the visited bit. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) push_kernel(BfsArgs<OffT> a) {
  zero_slots(a.stats_zero);
  for_each_queued_arc<OffT>(a.qin, a.qin_hdr, a.cols,
                            [&](uint32_t, uint32_t y, int64_t) { push_visit(a, y); });
}

/* Pull level.  A wave loads 64 visited words at once and compacts their
 * unvisited vertices in registers (popcount scan + bit select), so all 64 lanes
 * pull for real vertices; new frontier bits are gathered per word in LDS. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) pull_kernel(BfsArgs<OffT> a) {
  zero_slots(a.stats_zero);
  __shared__ unsigned long long lds_new[kBlock];
  uint64_t st_disc = 0, st_deg = 0, st_edges = 0, st_scanned = 0, st_max = 0, st_miss = 0;
  const int lane = lane_id();
  unsigned long long* wnew = lds_new + (threadIdx.x & ~63);
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const int tail_bits = (int)(a.n & 63);
  for (int64_t base = gwave * 64; base < a.w0; base += nwaves * 64) {
    const int64_t myw = base + lane;
    const uint64_t myvis = myw < a.w0 ? a.visited[myw] : ~0ull;
    uint64_t myun = ~myvis;
    if (myw == a.w0 - 1 && tail_bits) myun &= (1ull << tail_bits) - 1ull;
    const uint32_t cnt = (uint32_t)__popcll(myun);
    const uint32_t incl = wave_inclusive_scan(cnt);
    const uint32_t total = shfl(incl, 63);
    wnew[lane] = 0ull;
    for (uint32_t t0 = 0; t0 < total; t0 += 64) {
      const uint32_t t = t0 + lane;
      uint32_t lo = 0, hi = 63; /* owner word: first lane with incl > t */
#pragma unroll
      for (int it = 0; it < 6; ++it) {
        const uint32_t mid = (lo + hi) >> 1;
        if (__shfl(incl, (int)mid, 64) > t) hi = mid;
        else lo = mid + 1;
      }
      const int owner = (int)lo;
      const uint32_t before = __shfl(incl - cnt, owner, 64);
      const uint64_t un = shfl_u64(myun, owner);
      const bool active = t < total;
      const uint32_t bit = active ? select_bit(un, t - before) : 0u;
      const int64_t v = (base + owner) * 64 + bit;
      bool found = false, miss = false;
      OffT b = 0, e = 0, j = 0;
      int64_t kk = 0;
      if (active) {
        b = a.in_off[v];
        e = a.in_off[v + 1];
        /* first probe: the in-neighbour of largest out-degree, read from a
         * contiguous array instead of the adjacency list's cache line */
        if (test_bit(a.cur_bm, (uint32_t)a.hint[v])) {
          found = true;
          j = e;
        } else {
          miss = true;
          j = b;
          kk = (int64_t)(e - b);
        }
        /* lane-serial head: four independent id loads then four independent
         * bitmap probes per step */
        for (int step = 0; step < kPullHead / 4 && j < e; ++step, j += 4) {
          const uint32_t u0 = (uint32_t)a.in_cols[j];
          const uint32_t u1 = j + 1 < e ? (uint32_t)a.in_cols[j + 1] : u0;
          const uint32_t u2 = j + 2 < e ? (uint32_t)a.in_cols[j + 2] : u0;
          const uint32_t u3 = j + 3 < e ? (uint32_t)a.in_cols[j + 3] : u0;
          const bool h0 = test_bit(a.cur_bm, u0), h1 = test_bit(a.cur_bm, u1);
          const bool h2 = test_bit(a.cur_bm, u2), h3 = test_bit(a.cur_bm, u3);
          if (h0 | h1 | h2 | h3) {
            const int first = h0 ? 0 : (h1 ? 1 : (h2 ? 2 : 3));
            kk = (int64_t)(j - b) + first + 1;
            found = true;
            break;
          }
        }
      }
      /* cooperative tail: the whole wave scans the rest of each long list,
       * 64 ids per step; the ballot's lowest set bit is the first hit */
      uint64_t pending = ballot(active && !found && j < e);
      while (pending) {
        const int k2 = __ffsll((unsigned long long)pending) - 1;
        pending &= pending - 1;
        const int64_t jb = shfl_i64((int64_t)j, k2);
        const int64_t je = shfl_i64((int64_t)e, k2);
        const int64_t bb = shfl_i64((int64_t)b, k2);
        int64_t hit_pos = -1;
        for (int64_t jj = jb; jj < je; jj += 64) {
          const int64_t my = jj + lane;
          const bool h = my < je && test_bit(a.cur_bm, (uint32_t)a.in_cols[my]);
          const uint64_t hm = ballot(h);
          if (hm) {
            hit_pos = jj + (__ffsll((unsigned long long)hm) - 1);
            break;
          }
        }
        if (lane == k2 && hit_pos >= 0) {
          found = true;
          kk = hit_pos - bb + 1;
        }
      }
      if (active) {
        const int64_t deg = (int64_t)(e - b);
        if (deg > 0) ++st_scanned;
        st_edges += 1 + kk; /* hint probe + list probes after a miss */
        st_miss += miss ? 1 : 0;
        if (found) {
          a.depth[v] = a.level + 1;
          atomicOr(&wnew[owner], 1ull << bit);
          const int64_t odeg = a.directed ? (int64_t)(a.off[v + 1] - a.off[v]) : deg;
          ++st_disc;
          st_deg += odeg;
          st_max = odeg > (int64_t)st_max ? odeg : st_max;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    const uint64_t mynew = wnew[lane];
    if (myw < a.w0) {
      a.next_bm[myw] = mynew;
      if (mynew) a.visited[myw] = myvis | mynew;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  }
  flush_block_stats(a.stats, st_disc, st_deg, st_edges, st_scanned, st_max, st_miss);
}

/* Level-0 setup: depth[src] = 0, visited = prev = isolated | {src},
 * cur_bm = {src}, F0 queue = {src}. */
template<typename OffT>
__global__ void init_source_kernel(BfsArgs<OffT> a, uint64_t* prev, QEntry* q0, QEntry* q1, QEntry* q2, QueueHdr* hdr,
                                   Tuning t, int32_t src) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t bit = 1ull << (src & 63);
  a.depth[src] = 0;
  a.visited[src >> 6] |= bit;
  prev[src >> 6] |= bit;
  a.cur_bm[src >> 6] |= bit;
  QEntry* const q[3] = {q0, q1, q2};
  enqueue_single(a.off, t, q, hdr, src);
}

__global__ void copy_words_kernel(uint64_t* dst, uint64_t* dst2, const uint64_t* src, int64_t nw) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t x = src[i];
    dst[i] = x;
    dst2[i] = x;
  }
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

/* Per-graph BFS workspace, allocated on first use and reused. */
struct BfsWorkspace {
  int64_t w0 = 0;
  uint64_t* visited = nullptr;
  uint64_t* prev = nullptr;
  uint64_t* bm[2] = {nullptr, nullptr};
  QueueSet qs;
  StatSlot* stats = nullptr; /* 2 areas of kSlots */
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

void ws_free(BfsWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->visited);
  (void)hipFree(w->prev);
  (void)hipFree(w->bm[0]);
  (void)hipFree(w->bm[1]);
  w->qs.release();
  (void)hipFree(w->stats);
  for (auto e : w->ev)
    if (e) (void)hipEventDestroy(e);
  delete w;
}

int workspace_for(sg_graph_t g, BfsWorkspace** out) {
  if (g->bfs_ws) {
    *out = static_cast<BfsWorkspace*>(g->bfs_ws);
    return SG_OK;
  }
  auto* w = new BfsWorkspace();
  w->w0 = (g->n + 63) / 64;
  int st = SG_OK;
  sg_device_t dev = g->dev;
  if ((st = sg::alloc_device(dev, (void**)&w->visited, 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->prev, 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->bm[0], 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->bm[1], 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->stats, sizeof(StatSlot) * kSlots * 2)) ||
      (st = w->qs.init(dev, g->n, g->m, default_tuning()))) {
    ws_free(w);
    return st;
  }
  bool ok = true;
  for (auto& e : w->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  if (!ok) {
    ws_free(w);
    sg::set_error("bfs workspace: event creation failed");
    return SG_ERR_HIP;
  }
  g->bfs_ws = w;
  *out = w;
  return SG_OK;
}

struct HostStats {
  uint64_t disc = 0, disc_deg = 0, edges = 0, scanned = 0, max_deg = 0, misses = 0;
};

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
  if (opt.policy != SG_POLICY_PUSH) SG_TRY(sg::graph_ensure_pull_hint(g));
  const int64_t n = g->n, w0 = ws->w0;
  const int grid_main = dev->cus * 8;
  const int grid_small = std::max(1, std::min(dev->cus * 4, (int)((w0 + 255) / 256)));
  const Tuning tune = default_tuning();
  QueueSet& Q = ws->qs;
  const OffT* off = (const OffT*)g->offsets;

  BfsArgs<OffT> a{};
  a.off = off;
  a.cols = g->cols;
  a.in_off = (const OffT*)g->inv_offsets;
  a.in_cols = g->inv_cols;
  a.hint = g->pull_hint;
  a.directed = (g->flags & SG_GRAPH_DIRECTED) ? 1 : 0;
  a.n = n;
  a.w0 = w0;
  a.visited = ws->visited;
  a.depth = depth;
  int cb = 0; /* bitmap holding the current frontier */
  int qs = 0; /* queue set holding the current push frontier */

  fill_i32_kernel<<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(depth, (int32_t)(n + 1), n);
  copy_words_kernel<<<grid_small, 256, 0, s>>>(ws->visited, ws->prev, g->isolated, w0);
  SG_TRY(sg::zero_async(dev, ws->bm[0], 8 * w0));
  SG_TRY(sg::zero_async(dev, ws->stats, sizeof(StatSlot) * kSlots * 2));
  a.cur_bm = ws->bm[0];
  init_source_kernel<OffT><<<1, 64, 0, s>>>(a, ws->prev, Q.q[0][0], Q.q[0][1], Q.q[0][2], Q.hdr, tune, src);
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
  int32_t n_push = 0, n_pull = 0;
  QueueHdr* hh = (QueueHdr*)dev->host_scratch;
  StatSlot* hslots = (StatSlot*)((char*)dev->host_scratch + 4096);
  SG_HIP(hipMemcpyAsync(hh, Q.hdr, sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
  SG_HIP(hipStreamSynchronize(s));
  uint64_t visited_deg = hh->deg; /* the source's arcs count as explored */
  int64_t f_vertices = 1, f_edges = (int64_t)hh->deg;
  int64_t reached_v = 1;
  for (;; ++level) {
    a.level = level;
    a.stats = ws->stats + kSlots * (level & 1);
    a.stats_zero = ws->stats + kSlots * ((level + 1) & 1);
    for (int k = 0; k < 3; ++k) a.qin[k] = Q.q[qs][k];
    a.qin_hdr = Q.hdr + qs;
    a.cur_bm = ws->bm[cb];
    a.next_bm = ws->bm[cb ^ 1];
    const bool timed = opt.record_levels && stats && level < SG_MAX_LEVELS_RECORDED;
    if (timed) SG_HIP(hipEventRecord(ws->ev[0], s));
    HostStats c;
    if (push) {
      push_kernel<OffT><<<grid_main, kBlock, 0, s>>>(a);
      if (timed) SG_HIP(hipEventRecord(ws->ev[1], s));
      /* new frontier = visited & ~prev into bm[cb^1]; per-class counts and the
       * degree statistics land in header qs^1 (queues written later if the
       * next level pushes) */
      SG_TRY((compact_to_queues<OffT, kFromDiff>(s, Q, qs ^ 1, off, ws->visited, ws->prev, a.next_bm, tune, true)));
      SG_HIP(hipMemcpyAsync(hh, Q.hdr + (qs ^ 1), sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      c.disc = hh->nvert;
      c.disc_deg = hh->deg;
      c.max_deg = hh->maxdeg;
      c.edges = (uint64_t)f_edges;
      c.scanned = (uint64_t)f_vertices;
      ++n_push;
    } else {
      pull_kernel<OffT><<<grid_main, kBlock, 0, s>>>(a);
      if (timed) SG_HIP(hipEventRecord(ws->ev[1], s));
      SG_HIP(hipGetLastError());
      SG_HIP(hipMemcpyAsync(hslots, a.stats, sizeof(StatSlot) * kSlots, hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < kSlots; ++i) {
        c.disc += hslots[i].disc;
        c.disc_deg += hslots[i].disc_deg;
        c.edges += hslots[i].edges;
        c.scanned += hslots[i].scanned;
        c.misses += hslots[i].misses;
        c.max_deg = std::max<uint64_t>(c.max_deg, hslots[i].max_deg);
      }
      ++n_pull;
    }
    if (timed) {
      float ms = 0.f;
      SG_HIP(hipEventElapsedTime(&ms, ws->ev[0], ws->ev[1]));
      sg_level_stat& ls = stats->level[level];
      ls.direction = push ? 0 : 1;
      ls.load_balancer = opt.load_balancer;
      ls.frontier_vertices = f_vertices;
      ls.frontier_edges = f_edges;
      ls.discovered = (int64_t)c.disc;
      ls.edges_inspected = (int64_t)c.edges;
      ls.vertices_scanned = (int64_t)c.scanned;
      ls.hint_misses = (int64_t)c.misses;
      ls.max_degree = (int64_t)c.max_deg;
      ls.kernel_ms = ms;
    }
    visited_deg += c.disc_deg;
    reached_v += (int64_t)c.disc;
    const uint64_t unexplored = (uint64_t)g->total_degree - visited_deg;
    if (timed) stats->level[level].unexplored_edges = (int64_t)unexplored;
    f_vertices = (int64_t)c.disc;
    f_edges = (int64_t)c.disc_deg;
    if (c.disc == 0) {
      ++level;
      break;
    }
    /* direction of the next level (evaluateHeuristic, bfs.hpp:341-350) */
    bool next_push = push;
    if (opt.policy == SG_POLICY_HYBRID) {
      if (push) {
        const uint32_t fdeg = (uint32_t)c.disc_deg, udeg = (uint32_t)unexplored;
        if ((float)fdeg > ((float)udeg / opt.alpha)) next_push = false;
      } else {
        if ((float)c.disc < ((float)n / opt.beta)) next_push = true;
      }
    }
    if (timed) SG_HIP(hipEventRecord(ws->ev[2], s));
    cb ^= 1; /* the new frontier is in bm[cb] now, after either direction */
    if (next_push) {
      if (push) {
        SG_TRY(compact_write_only<OffT>(s, Q, qs ^ 1, off, ws->bm[cb], tune));
      } else {
        SG_TRY((compact_to_queues<OffT, kFromBitmapSyncPrev>(s, Q, qs ^ 1, off, ws->visited, ws->prev, ws->bm[cb],
                                                              tune)));
      }
      qs ^= 1;
    }
    if (timed) {
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
