/*
 * msbfs.hip -- multi-source BFS, 64 sources per batch (one bit-lane each).
 *
 * Every vertex carries three 64-bit masks: visited (bit s: reached from
 * source s), frontier (bit s: at the current depth for source s) and next.
 * One pass over the graph advances all 64 traversals by a level, so each arc
 * is read once per level for the whole batch instead of once per source:
 *  - push levels: vertices with a non-empty frontier mask go through the
 *    degree-class queues (the bucketing load balancer of the single-source
 *    BFS); for each arc u -> v, next[v] |= frontier[u] & ~visited[v];
 *  - pull levels: every vertex still missing a lane ORs the frontier masks of
 *    its in-neighbours and stops as soon as all missing lanes are covered;
 *  - an update pass folds next into visited, writes depth[s][v] = level + 1
 *    for every new bit (one row per source, so the stores of a wave to one row
 *    are coalesced) and makes next the frontier.
 * A level pulls when more than 1/20 of the vertices hold frontier bits.
 * Depths are bit-exact with one BFS per source (the level at which a vertex
 * is first reached does not depend on the direction).
 */
#include <algorithm>
#include <cstring>
#include <vector>

#include "traverse.hpp"

using namespace sg::trav;

namespace {

struct MsStats {
  unsigned long long active; /* vertices with a non-empty new mask */
  unsigned long long lanes;  /* sum of new bits (vertex, source pairs reached) */
  unsigned long long pad[14];
};

template<typename OffT>
__global__ void __launch_bounds__(kBlock) ms_push_kernel(QEntry* q0, QEntry* q1, QEntry* q2, const QueueHdr* hdr,
                                                         const int32_t* cols, const uint64_t* frontier,
                                                         const uint64_t* visited, uint64_t* next) {
  QEntry* const q[3] = {q0, q1, q2};
  for_each_queued_arc<OffT>(q, hdr, cols, [&](uint32_t u, uint32_t v, int64_t) {
    const uint64_t add = frontier[u] & ~visited[v];
    if (add && (next[v] & add) != add) atomic_or(&next[v], add);
  });
}

/* one lane per vertex; a list longer than 64 arcs is ORed by the whole wave */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) ms_pull_kernel(const OffT* in_off, const int32_t* in_cols, int64_t n,
                                                         uint64_t full, const uint64_t* frontier,
                                                         const uint64_t* visited, uint64_t* next) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t base = gwave * 64; base < n; base += nwaves * 64) {
    const int64_t v = base + lane;
    uint64_t vis = full, acc = 0;
    int64_t b = 0, e = 0;
    if (v < n) {
      vis = visited[v];
      if (vis != full) {
        b = (int64_t)in_off[v];
        e = (int64_t)in_off[v + 1];
      }
    }
    const uint64_t want = full & ~vis;
    const bool coop = e - b > 64;
    if (!coop)
      for (int64_t j = b; j < e; ++j) {
        acc |= frontier[in_cols[j]];
        if ((acc & want) == want) break;
      }
    uint64_t heavy = ballot(coop);
    while (heavy) {
      const int k = __ffsll((unsigned long long)heavy) - 1;
      heavy &= heavy - 1;
      const int64_t kb = shfl_i64(b, k), ke = shfl_i64(e, k);
      const uint64_t kwant = shfl_u64(want, k);
      uint64_t part = 0;
      for (int64_t j0 = kb; j0 < ke; j0 += 64) {
        if (j0 + lane < ke) part |= frontier[in_cols[j0 + lane]];
        /* wave OR (butterfly), then stop once every wanted lane is covered */
        uint64_t w = part;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) w |= shfl_u64(w, lane ^ o);
        if ((w & kwant) == kwant) {
          part = w;
          break;
        }
        part = w;
      }
      if (lane == k) acc = part;
    }
    if (v < n) {
      const uint64_t add = acc & want;
      if (add) next[v] = add;
    }
  }
}

/* Fold next into visited, write depths, next -> frontier, bitmap of
 * non-empty frontier masks for the next push level, statistics. */
__global__ void __launch_bounds__(kBlock) ms_update_kernel(int64_t n, int64_t w0, int32_t nsrc, int32_t level,
                                                           uint64_t* visited, uint64_t* frontier, uint64_t* next,
                                                           uint64_t* active_bm, int32_t* depth, MsStats* stats) {
  const int lane = lane_id();
  unsigned long long act = 0, lanes = 0;
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t base = gwave * 64; base < w0 * 64; base += nwaves * 64) {
    const int64_t v = base + lane;
    uint64_t add = 0;
    if (v < n) {
      add = next[v];
      if (add) {
        visited[v] |= add;
        next[v] = 0;
        ++act;
        lanes += (unsigned long long)__popcll(add);
      }
      frontier[v] = add;
    }
    const uint64_t any = ballot(add != 0);
    if (lane == 0) active_bm[base >> 6] = any;
    uint64_t bits = add;
    uint64_t wave_bits = 0;
    {
      uint64_t w = bits;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) w |= shfl_u64(w, lane ^ o);
      wave_bits = w;
    }
    while (wave_bits) { /* every source with a new vertex in this wave: one coalesced row store */
      const int s = __ffsll((unsigned long long)wave_bits) - 1;
      wave_bits &= wave_bits - 1;
      if (s < nsrc && ((bits >> s) & 1ull)) depth[(int64_t)s * n + v] = level + 1;
    }
  }
  act = wave_sum_u64(act);
  lanes = wave_sum_u64(lanes);
  if (lane == 0 && (act | lanes)) {
    MsStats* st = stats + (blockIdx.x % kSlots);
    atomic_add(&st->active, act);
    atomic_add(&st->lanes, lanes);
  }
}

__global__ void ms_init_kernel(int64_t n, int64_t w0, uint64_t* visited, uint64_t* frontier, uint64_t* next,
                               uint64_t* active_bm) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    visited[i] = 0;
    frontier[i] = 0;
    next[i] = 0;
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < w0; i += (int64_t)gridDim.x * blockDim.x)
    active_bm[i] = 0;
}

/* sources of the batch: visited/frontier bit, depth 0, active bitmap */
__global__ void ms_sources_kernel(const int32_t* src, int32_t k, int64_t n, uint64_t* visited, uint64_t* frontier,
                                  uint64_t* active_bm, int32_t* depth) {
  const int i = threadIdx.x;
  if (i >= k) return;
  const int32_t v = src[i];
  atomic_or(&visited[v], 1ull << i);
  atomic_or(&frontier[v], 1ull << i);
  atomic_or(&active_bm[v >> 6], 1ull << (v & 63));
  depth[(int64_t)i * n + v] = 0;
}

__global__ void ms_fill_kernel(int32_t* p, int32_t val, int64_t cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = val;
}

struct MsBuffers {
  uint64_t *visited = nullptr, *frontier = nullptr, *next = nullptr, *active = nullptr;
  int32_t* src = nullptr;
  MsStats* stats = nullptr;
  QueueSet qs;
  ~MsBuffers() {
    void* ps[] = {visited, frontier, next, active, src, stats};
    for (void* p : ps) (void)hipFree(p);
    qs.release();
  }
};

template<typename OffT>
int run_msbfs(sg_graph_t g, int32_t nsrc, const int32_t* sources, int32_t* depth, int64_t* level_count) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n, w0 = (n + 63) / 64;
  const Tuning tune = default_tuning();
  MsBuffers B;
  SG_TRY(sg::alloc_device(dev, (void**)&B.visited, 8 * (n + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.frontier, 8 * (n + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.next, 8 * (n + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.active, 8 * (w0 + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.src, 4 * 64));
  SG_TRY(sg::alloc_device(dev, (void**)&B.stats, sizeof(MsStats) * kSlots));
  SG_TRY(B.qs.init(dev, n, g->m, tune));
  const OffT* off = (const OffT*)g->offsets;
  const OffT* in_off = (const OffT*)(g->inv_offsets ? g->inv_offsets : g->offsets);
  const int32_t* in_cols = g->inv_cols ? g->inv_cols : g->cols;
  const int grid_main = dev->cus * 8;
  const int grid_n = sg::grid_for(dev, n, kBlock, 8);
  MsStats* hst = (MsStats*)dev->host_scratch;
  int64_t levels_total = 0;

  ms_fill_kernel<<<sg::grid_for(dev, (int64_t)nsrc * n, 256, 8), 256, 0, s>>>(depth, (int32_t)(n + 1),
                                                                               (int64_t)nsrc * n);
  for (int32_t b0 = 0; b0 < nsrc; b0 += 64) {
    const int32_t k = std::min<int32_t>(64, nsrc - b0);
    const uint64_t full = k == 64 ? ~0ull : ((1ull << k) - 1ull);
    SG_HIP(hipMemcpyAsync(B.src, sources + b0, 4 * k, hipMemcpyHostToDevice, s));
    ms_init_kernel<<<grid_n, kBlock, 0, s>>>(n, w0, B.visited, B.frontier, B.next, B.active);
    int32_t* dep = depth + (int64_t)b0 * n;
    ms_sources_kernel<<<1, 64, 0, s>>>(B.src, k, n, B.visited, B.frontier, B.active, dep);
    /* the distinct sources are the first frontier */
    int64_t active = 0;
    {
      std::vector<int32_t> u(sources + b0, sources + b0 + k);
      std::sort(u.begin(), u.end());
      active = std::unique(u.begin(), u.end()) - u.begin();
    }
    for (int32_t level = 0; active > 0; ++level) {
      const bool pull = active * 20 > n;
      if (pull) {
        ms_pull_kernel<OffT><<<grid_main, kBlock, 0, s>>>(in_off, in_cols, n, full, B.frontier, B.visited, B.next);
      } else {
        SG_TRY((compact_to_queues<OffT, kFromBitmap>(s, B.qs, 0, off, nullptr, nullptr, B.active, tune)));
        ms_push_kernel<OffT><<<grid_main, kBlock, 0, s>>>(B.qs.q[0][0], B.qs.q[0][1], B.qs.q[0][2], B.qs.hdr, g->cols,
                                                          B.frontier, B.visited, B.next);
      }
      SG_TRY(sg::zero_async(dev, B.stats, sizeof(MsStats) * kSlots));
      ms_update_kernel<<<grid_n, kBlock, 0, s>>>(n, w0, k, level, B.visited, B.frontier, B.next, B.active, dep,
                                                 B.stats);
      SG_HIP(hipGetLastError());
      SG_HIP(hipMemcpyAsync(hst, B.stats, sizeof(MsStats) * kSlots, hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      active = 0;
      for (int i = 0; i < kSlots; ++i) active += (int64_t)hst[i].active;
      ++levels_total;
    }
  }
  SG_HIP(hipStreamSynchronize(s));
  if (level_count) *level_count = levels_total;
  return SG_OK;
}

} // namespace

/* Multi-source BFS: 64 sources per batch, one bit-lane each. */
extern "C" int sg_bfs_multi(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, int32_t* depth) {
  SG_CHECK(g && depth && (nsrc == 0 || sources_host), SG_ERR_INVALID, "sg_bfs_multi: null argument");
  SG_CHECK(nsrc >= 0, SG_ERR_INVALID, "sg_bfs_multi: negative source count");
  for (int32_t i = 0; i < nsrc; ++i)
    SG_CHECK(sources_host[i] >= 0 && sources_host[i] < g->n, SG_ERR_INVALID,
             "sg_bfs_multi: source %d out of range [0,%lld)", sources_host[i], (long long)g->n);
  if (nsrc == 0) return SG_OK;
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return run_msbfs<int64_t>(g, nsrc, sources_host, depth, nullptr);
  return run_msbfs<uint32_t>(g, nsrc, sources_host, depth, nullptr);
}
