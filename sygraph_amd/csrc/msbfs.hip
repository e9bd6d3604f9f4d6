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

/* buffers kept with the graph (repeated calls allocate nothing) */
struct MsBuffers {
  uint64_t *visited = nullptr, *frontier = nullptr, *next = nullptr, *active = nullptr;
  int32_t* src = nullptr;
  MsStats* stats = nullptr;
  QueueSet* qs = nullptr;
};

template<typename OffT>
int run_msbfs(sg_graph_t g, int32_t nsrc, const int32_t* sources, int32_t* depth, int64_t* level_count) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n, w0 = (n + 63) / 64;
  const Tuning tune = default_tuning();
  MsBuffers B;
  SG_TRY(sg::graph_buffer(g, "ms_visited", 8 * (n + 1), &B.visited));
  SG_TRY(sg::graph_buffer(g, "ms_frontier", 8 * (n + 1), &B.frontier));
  SG_TRY(sg::graph_buffer(g, "ms_next", 8 * (n + 1), &B.next));
  SG_TRY(sg::graph_buffer(g, "ms_active", 8 * (w0 + 1), &B.active));
  SG_TRY(sg::graph_buffer(g, "ms_src", 4 * 64, &B.src));
  SG_TRY(sg::graph_buffer(g, "ms_stats", sizeof(MsStats) * kSlots, &B.stats));
  SG_TRY(graph_queue_set(g, &B.qs));
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
        SG_TRY((compact_to_queues<OffT, kFromBitmap>(s, *B.qs, 0, off, nullptr, nullptr, B.active, tune)));
        ms_push_kernel<OffT><<<grid_main, kBlock, 0, s>>>(B.qs->q[0][0], B.qs->q[0][1], B.qs->q[0][2], B.qs->hdr, g->cols,
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

/* ------------------------------------------------------- batched BC --- */
/* Brandes' dependencies for 64 sources at once (bit-lane l <-> source l):
 * the forward pass is the multi-source BFS above, and each of its levels is
 * followed by a sigma pass over the vertices it reached; the backward pass
 * walks the levels from the deepest up.  Per level d the mask array F_d
 * holds, for every vertex, the lanes for which it lies at depth d (these are
 * the BFS's `next` arrays, kept for every level).
 *
 * Rows: T[v][l] (FP64, 64 lanes = 512 bytes per vertex).  Forward,
 *   t_l(v) = g_l(d) * sum of t_l(u) over in-arcs u -> v with l in F_d(u),
 * for l in F_{d+1}(v), where g_l(d) = 2^-e is the power of two that brings
 * the largest t_l of level d into [0.5, 1) (path counts overflow FP64 on
 * long-diameter graphs; power-of-two scaling is exact).  So t_l(v) =
 * sigma_l(v) * prod_{j<d} g_l(j), and sigma(v)/sigma(w) = t(v) g(d) / t(w)
 * for depth(w) = depth(v) + 1.  Backward, level d:
 *   delta_l(v) = t_l(v) * g_l(d) * sum of R_l(w) over out-arcs v -> w with
 *                l in F_{d+1}(w),
 *   R_l(v) = (1 + delta_l(v)) / t_l(v)   (overwrites lane l of T[v]),
 *   bc(v) += sum over l in F_d(v) of delta_l(v)   (d >= 1: sources excluded).
 * A vertex's lanes at other depths are other addresses, so a level's reads
 * and writes never overlap.
 *
 * Work mapping: a wavefront per (vertex, chunk of up to kBcChunk arcs), lane
 * = source, eight arcs in flight per step.  A vertex with one chunk is
 * finished by that wave; longer lists write per-chunk partial rows that a
 * second pass adds in chunk order -- no floating-point atomics, so results
 * are deterministic.  Work lists come from an ordered count / scan / write
 * compaction of the level's mask array. */
constexpr int kBcChunk = 256;
constexpr int kBcUnroll = 8;
constexpr int kBcSlotStride = 16; /* one 128-byte line per lane's level maximum */
constexpr int64_t kBcMaskBudget = 16ll << 30;

struct BcEntry {
  int64_t start;
  int32_t v;
  int32_t len;
  int32_t slot; /* -1: the vertex's only entry; else its partial row */
  int32_t pad;
};

struct BcVert {
  int32_t v, first, count, pad;
};

struct BcCounts {
  uint32_t entries, verts, multi, pad;
};

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

template<typename OffT>
__device__ __forceinline__ void bc_vertex_counts(const uint64_t* mask, const OffT* off, int64_t n, int64_t v,
                                                 uint32_t& ne, uint32_t& act, uint32_t& multi, int64_t& start,
                                                 int64_t& deg) {
  ne = act = multi = 0;
  start = deg = 0;
  if (v < n && mask[v]) {
    start = (int64_t)off[v];
    deg = (int64_t)(off[v + 1] - off[v]);
    ne = deg > 0 ? (uint32_t)((deg + kBcChunk - 1) / kBcChunk) : 1u;
    act = 1;
    multi = ne > 1 ? ne : 0;
  }
}

/* pass 1: per group of 64 vertices (one wave), the entries, vertices and
 * partial rows it contributes */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) bc_count_kernel(const uint64_t* mask, const OffT* off, int64_t n,
                                                          int64_t ngroups, BcCounts* groups) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t g = gwave; g < ngroups; g += nwaves) {
    uint32_t ne, act, multi;
    int64_t start, deg;
    bc_vertex_counts(mask, off, n, g * 64 + lane, ne, act, multi, start, deg);
    const uint32_t se = (uint32_t)wave_sum_u64(ne), sa = (uint32_t)wave_sum_u64(act);
    const uint32_t sm = (uint32_t)wave_sum_u64(multi);
    if (lane == 0) groups[g] = BcCounts{se, sa, sm, 0};
  }
}

/* pass 2: exclusive scan of the group counts (one 1024-thread workgroup,
 * four groups per thread per round) and the totals */
__global__ void __launch_bounds__(1024) bc_scan_kernel(BcCounts* groups, int64_t ngroups, BcCounts* total) {
  __shared__ uint32_t wt[16][3];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  uint32_t carry[3] = {0, 0, 0};
  for (int64_t base = 0; base < ngroups; base += 1024 * 4) {
    const int64_t first = base + (int64_t)threadIdx.x * 4;
    uint32_t loc[4][3], sum[3] = {0, 0, 0}, incl[3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      BcCounts c{0, 0, 0, 0};
      if (first + k < ngroups) c = groups[first + k];
      loc[k][0] = c.entries;
      loc[k][1] = c.verts;
      loc[k][2] = c.multi;
#pragma unroll
      for (int j = 0; j < 3; ++j) sum[j] += loc[k][j];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      incl[j] = wave_inclusive_scan(sum[j]);
      if (lane == 63) wt[wv][j] = incl[j];
    }
    __syncthreads();
    uint32_t run[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      uint32_t before = 0, tot = 0;
      for (int w = 0; w < 16; ++w) {
        if (w < wv) before += wt[w][j];
        tot += wt[w][j];
      }
      run[j] = carry[j] + before + incl[j] - sum[j];
      carry[j] += tot;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (first + k < ngroups) groups[first + k] = BcCounts{run[0], run[1], run[2], 0};
#pragma unroll
      for (int j = 0; j < 3; ++j) run[j] += loc[k][j];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = BcCounts{carry[0], carry[1], carry[2], 0};
}

/* pass 3: the work entries (vertex order) and the vertex list */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) bc_write_kernel(const uint64_t* mask, const OffT* off, int64_t n,
                                                          int64_t ngroups, const BcCounts* groups, BcEntry* ents,
                                                          BcVert* verts) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t g = gwave; g < ngroups; g += nwaves) {
    const int64_t v = g * 64 + lane;
    uint32_t ne, act, multi;
    int64_t start, deg;
    bc_vertex_counts(mask, off, n, v, ne, act, multi, start, deg);
    const uint32_t ie = wave_inclusive_scan(ne), ia = wave_inclusive_scan(act), im = wave_inclusive_scan(multi);
    const BcCounts b = groups[g];
    if (act) {
      const uint32_t ei = b.entries + ie - ne, vi = b.verts + ia - 1, mi = b.multi + im - multi;
      verts[vi] = BcVert{(int32_t)v, (int32_t)ei, (int32_t)ne, 0};
      for (uint32_t c = 0; c < ne; ++c) {
        BcEntry e;
        e.start = start + (int64_t)c * kBcChunk;
        e.v = (int32_t)v;
        const int64_t rem = deg - (int64_t)c * kBcChunk;
        e.len = (int32_t)(rem < kBcChunk ? rem : kBcChunk);
        e.slot = ne == 1 ? -1 : (int32_t)(mi + c);
        e.pad = 0;
        ents[ei + c] = e;
      }
    }
  }
}

/* Finish vertex v for the lanes in `own` from its arc sum `acc` (lane = source). */
template<bool kBackward>
__device__ __forceinline__ void bc_finish(int64_t v, uint64_t own, double acc, double g, double* T, double* bc64,
                                          double& mx) {
  const int lane = lane_id();
  const bool mine = (own >> lane) & 1ull;
  double* row = T + v * 64;
  if (!kBackward) {
    if (mine) {
      const double t = g * acc;
      row[lane] = t;
      mx = t > mx ? t : mx;
    }
  } else {
    double delta = 0.0;
    if (mine) {
      const double tv = row[lane];
      delta = tv * g * acc;
      row[lane] = (1.0 + delta) / tv;
    }
    const double s = wave_sum_f64(delta);
    if (lane == 0) bc64[v] += s;
  }
}

/* Per-lane maximum of the workgroup's finished t values into the level slots. */
__device__ __forceinline__ void bc_flush_max(double mx, unsigned long long* slots) {
  __shared__ double red[kBlock / 64][64];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  red[wv][lane] = mx;
  __syncthreads();
  if (threadIdx.x < 64) {
    double m = red[0][lane];
    for (int w = 1; w < kBlock / 64; ++w) m = red[w][lane] > m ? red[w][lane] : m;
    if (m > 0.0) atomic_max(&slots[lane * kBcSlotStride], (unsigned long long)__double_as_longlong(m));
  }
}

/* A wave per entry: the sum over the entry's arcs of the neighbour rows,
 * lanes restricted to own(v) & nbr(u); finishes single-entry vertices. */
template<bool kBackward>
__global__ void __launch_bounds__(kBlock) bc_task_kernel(const BcEntry* ents, const BcCounts* total,
                                                         const int32_t* cols, const uint64_t* own_mask,
                                                         const uint64_t* nbr_mask, double* T, const double* gl,
                                                         double* part, double* bc64, unsigned long long* slots) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const uint32_t ne = total->entries;
  const double g = gl[lane];
  double mx = 0.0;
  for (int64_t i = gwave; i < ne; i += nwaves) {
    const BcEntry e = ents[i];
    const uint64_t own = own_mask[e.v];
    double acc = 0.0;
    if (nbr_mask) {
      for (int32_t j = 0; j < e.len; j += kBcUnroll) {
        int32_t u[kBcUnroll];
        uint64_t mk[kBcUnroll];
#pragma unroll
        for (int k = 0; k < kBcUnroll; ++k) u[k] = j + k < e.len ? cols[e.start + j + k] : -1;
#pragma unroll
        for (int k = 0; k < kBcUnroll; ++k) mk[k] = u[k] >= 0 ? (nbr_mask[u[k]] & own) : 0ull;
        double val[kBcUnroll];
#pragma unroll
        for (int k = 0; k < kBcUnroll; ++k) val[k] = ((mk[k] >> lane) & 1ull) ? T[(int64_t)u[k] * 64 + lane] : 0.0;
#pragma unroll
        for (int k = 0; k < kBcUnroll; ++k) acc += val[k];
      }
    }
    if (e.slot < 0) bc_finish<kBackward>(e.v, own, acc, g, T, bc64, mx);
    else part[(int64_t)e.slot * 64 + lane] = acc;
  }
  if (!kBackward) bc_flush_max(mx, slots);
}

/* A wave per vertex of several entries: adds its partial rows in order. */
template<bool kBackward>
__global__ void __launch_bounds__(kBlock) bc_multi_kernel(const BcVert* verts, const BcEntry* ents,
                                                          const BcCounts* total, const uint64_t* own_mask, double* T,
                                                          const double* gl, const double* part, double* bc64,
                                                          unsigned long long* slots) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const uint32_t nv = total->verts;
  const double g = gl[lane];
  double mx = 0.0;
  for (int64_t i = gwave; i < nv; i += nwaves) {
    const BcVert vt = verts[i];
    if (vt.count <= 1) continue;
    const int64_t slot0 = ents[vt.first].slot;
    double acc = 0.0;
    for (int32_t c = 0; c < vt.count; ++c) acc += part[(slot0 + c) * 64 + lane];
    bc_finish<kBackward>(vt.v, own_mask[vt.v], acc, g, T, bc64, mx);
  }
  if (!kBackward) bc_flush_max(mx, slots);
}

/* g(d) for every lane from the level's maxima; clears the slots. */
__global__ void bc_scale_kernel(unsigned long long* slots, double* gl) {
  const int lane = threadIdx.x;
  const unsigned long long bits = slots[lane * kBcSlotStride];
  slots[lane * kBcSlotStride] = 0ull;
  const double m = __longlong_as_double((long long)bits);
  gl[lane] = m > 0.0 ? ldexp(1.0, -(ilogb(m) + 1)) : 1.0;
}

/* visited |= next, bitmap of the vertices holding new lanes, statistics */
__global__ void __launch_bounds__(kBlock) bc_update_kernel(int64_t n, int64_t w0, uint64_t* visited,
                                                           const uint64_t* next, uint64_t* active_bm, MsStats* stats) {
  const int lane = lane_id();
  unsigned long long act = 0;
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t base = gwave * 64; base < w0 * 64; base += nwaves * 64) {
    const int64_t v = base + lane;
    uint64_t add = 0;
    if (v < n) {
      add = next[v];
      if (add) {
        visited[v] |= add;
        ++act;
      }
    }
    const uint64_t any = ballot(add != 0);
    if (lane == 0) active_bm[base >> 6] = any;
  }
  act = wave_sum_u64(act);
  if (lane == 0 && act) atomic_add(&stats[blockIdx.x % kSlots].active, act);
}

/* sources of the batch: visited and level-0 mask bits, t = 1 */
__global__ void bc_sources_kernel(const int32_t* src, int32_t k, uint64_t* visited, uint64_t* mask0,
                                  uint64_t* active_bm, double* T, double* gl0) {
  const int i = threadIdx.x;
  if (i < 64) gl0[i] = 1.0;
  if (i >= k) return;
  const int32_t v = src[i];
  atomic_or(&visited[v], 1ull << i);
  atomic_or(&mask0[v], 1ull << i);
  atomic_or(&active_bm[v >> 6], 1ull << (v & 63));
  T[(int64_t)v * 64 + i] = 1.0;
}

__global__ void bc_fill_u64_kernel(uint64_t* p, int64_t cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

__global__ void bc_add_to_float_kernel(const double* acc, float* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)((double)out[i] + acc[i]);
}

/* buffers kept with the graph (repeated calls allocate nothing) */
struct BcBuffers {
  uint64_t *visited = nullptr, *active = nullptr;
  std::vector<uint64_t*> masks;
  double *T = nullptr, *part = nullptr, *bc64 = nullptr, *gtab = nullptr;
  unsigned long long* slots = nullptr;
  BcEntry* ents = nullptr;
  BcVert* verts = nullptr;
  BcCounts *groups = nullptr, *total = nullptr;
  int32_t* src = nullptr;
  MsStats* stats = nullptr;
  QueueSet* qs = nullptr;
};

/* count / scan / write of the level's work lists, then the task and
 * multi-entry passes */
template<typename OffT, bool kBackward>
int bc_level(hipStream_t s, BcBuffers& B, const OffT* off, const int32_t* cols, int64_t n, int grid,
             const uint64_t* own, const uint64_t* nbr, const double* gl) {
  const int64_t ngroups = (n + 63) / 64;
  bc_count_kernel<OffT><<<grid, kBlock, 0, s>>>(own, off, n, ngroups, B.groups);
  bc_scan_kernel<<<1, 1024, 0, s>>>(B.groups, ngroups, B.total);
  bc_write_kernel<OffT><<<grid, kBlock, 0, s>>>(own, off, n, ngroups, B.groups, B.ents, B.verts);
  bc_task_kernel<kBackward><<<grid, kBlock, 0, s>>>(B.ents, B.total, cols, own, nbr, B.T, gl, B.part, B.bc64,
                                                    B.slots);
  bc_multi_kernel<kBackward><<<grid, kBlock, 0, s>>>(B.verts, B.ents, B.total, own, B.T, gl, B.part, B.bc64,
                                                     B.slots);
  SG_HIP(hipGetLastError());
  return SG_OK;
}

template<typename OffT>
int run_bc_batched(sg_graph_t g, int32_t nsrc, const int32_t* sources, float* out) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n, m = g->m, w0 = (n + 63) / 64;
  const Tuning tune = default_tuning();
  BcBuffers B;
  const int64_t ent_cap = n + 2 * (m / kBcChunk) + 64, part_cap = 2 * (m / kBcChunk) + 64;
  SG_TRY(sg::graph_buffer(g, "ms_visited", 8 * (n + 1), &B.visited));
  SG_TRY(sg::graph_buffer(g, "ms_active", 8 * (w0 + 1), &B.active));
  SG_TRY(sg::graph_buffer(g, "bcb_rows", 8 * 64 * (size_t)(n + 1), &B.T));
  SG_TRY(sg::graph_buffer(g, "bcb_part", 8 * 64 * (size_t)part_cap, &B.part));
  SG_TRY(sg::graph_buffer(g, "bcb_bc64", 8 * (n + 1), &B.bc64));
  SG_TRY(sg::graph_buffer(g, "bcb_slots", 8 * 64 * kBcSlotStride, &B.slots));
  SG_TRY(sg::graph_buffer(g, "bcb_entries", sizeof(BcEntry) * ent_cap, &B.ents));
  SG_TRY(sg::graph_buffer(g, "bcb_verts", sizeof(BcVert) * (n + 1), &B.verts));
  SG_TRY(sg::graph_buffer(g, "bcb_groups", sizeof(BcCounts) * (w0 + 1), &B.groups));
  SG_TRY(sg::graph_buffer(g, "bcb_total", sizeof(BcCounts), &B.total));
  SG_TRY(sg::graph_buffer(g, "ms_src", 4 * 64, &B.src));
  SG_TRY(sg::graph_buffer(g, "ms_stats", sizeof(MsStats) * kSlots, &B.stats));
  SG_TRY(graph_queue_set(g, &B.qs));
  SG_TRY(sg::zero_async(dev, B.bc64, 8 * (n + 1)));
  SG_TRY(sg::zero_async(dev, B.slots, 8 * 64 * kBcSlotStride));
  const OffT* off = (const OffT*)g->offsets;
  const OffT* in_off = (const OffT*)(g->inv_offsets ? g->inv_offsets : g->offsets);
  const int32_t* in_cols = g->inv_cols ? g->inv_cols : g->cols;
  const int grid_main = dev->cus * 8;
  const int grid_n = sg::grid_for(dev, n, kBlock, 8);
  MsStats* hst = (MsStats*)dev->host_scratch;
  const int64_t max_levels = std::max<int64_t>(2, std::min<int64_t>(n + 2, kBcMaskBudget / (8 * (n + 1))));
  auto mask = [&](int32_t d, uint64_t** out) -> int { /* the level's mask array, zeroed */
    while ((int32_t)B.masks.size() <= d) {
      SG_CHECK((int64_t)B.masks.size() < max_levels, SG_ERR_UNSUPPORTED,
               "sg_bc_batched: more than %lld levels exceed the mask budget; use sg_bc", (long long)max_levels);
      uint64_t* p = nullptr;
      const std::string name = "bcb_mask_" + std::to_string(B.masks.size());
      SG_TRY(sg::graph_buffer(g, name.c_str(), 8 * (n + 1), &p));
      B.masks.push_back(p);
    }
    bc_fill_u64_kernel<<<grid_n, kBlock, 0, s>>>(B.masks[d], n + 1);
    *out = B.masks[d];
    return SG_OK;
  };
  for (int32_t b0 = 0; b0 < nsrc; b0 += 64) {
    const int32_t k = std::min<int32_t>(64, nsrc - b0);
    const uint64_t full = k == 64 ? ~0ull : ((1ull << k) - 1ull);
    /* g(d): 64 doubles per level */
    auto gl_of = [&](int32_t d) -> double* { return B.gtab + (int64_t)d * 64; };
    SG_HIP(hipMemcpyAsync(B.src, sources + b0, 4 * k, hipMemcpyHostToDevice, s));
    bc_fill_u64_kernel<<<grid_n, kBlock, 0, s>>>(B.visited, n + 1);
    bc_fill_u64_kernel<<<sg::grid_for(dev, w0 + 1, kBlock, 4), kBlock, 0, s>>>(B.active, w0 + 1);
    uint64_t* m0 = nullptr;
    SG_TRY(mask(0, &m0));
    if (!B.gtab) SG_TRY(sg::graph_buffer(g, "bcb_gtab", 8 * 64 * (size_t)(max_levels + 1), &B.gtab));
    bc_sources_kernel<<<1, 64, 0, s>>>(B.src, k, B.visited, m0, B.active, B.T, gl_of(0));
    int64_t active = 0;
    {
      std::vector<int32_t> u(sources + b0, sources + b0 + k);
      std::sort(u.begin(), u.end());
      active = std::unique(u.begin(), u.end()) - u.begin();
    }
    int32_t deepest = 0;
    for (int32_t d = 0;; ++d) {
      uint64_t* nx = nullptr;
      SG_TRY(mask(d + 1, &nx));
      uint64_t* fr = B.masks[d];
      if (active * 20 > n) {
        ms_pull_kernel<OffT><<<grid_main, kBlock, 0, s>>>(in_off, in_cols, n, full, fr, B.visited, nx);
      } else {
        SG_TRY((compact_to_queues<OffT, kFromBitmap>(s, *B.qs, 0, off, nullptr, nullptr, B.active, tune)));
        ms_push_kernel<OffT><<<grid_main, kBlock, 0, s>>>(B.qs->q[0][0], B.qs->q[0][1], B.qs->q[0][2], B.qs->hdr, g->cols,
                                                          fr, B.visited, nx);
      }
      /* sigma of the vertices reached at depth d + 1 (in-arcs from depth d) */
      SG_TRY((bc_level<OffT, false>(s, B, in_off, in_cols, n, grid_main, nx, fr, gl_of(d))));
      bc_scale_kernel<<<1, 64, 0, s>>>(B.slots, gl_of(d + 1));
      SG_TRY(sg::zero_async(dev, B.stats, sizeof(MsStats) * kSlots));
      bc_update_kernel<<<grid_n, kBlock, 0, s>>>(n, w0, B.visited, nx, B.active, B.stats);
      SG_HIP(hipGetLastError());
      SG_HIP(hipMemcpyAsync(hst, B.stats, sizeof(MsStats) * kSlots, hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      active = 0;
      for (int i = 0; i < kSlots; ++i) active += (int64_t)hst[i].active;
      if (active == 0) break;
      deepest = d + 1;
    }
    /* backward, deepest level first (level 0 holds only the sources) */
    for (int32_t d = deepest; d >= 1; --d) {
      const uint64_t* nbr = d < deepest ? B.masks[d + 1] : nullptr;
      SG_TRY((bc_level<OffT, true>(s, B, off, g->cols, n, grid_main, B.masks[d], nbr, gl_of(d))));
    }
  }
  bc_add_to_float_kernel<<<grid_n, kBlock, 0, s>>>(B.bc64, out, n);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(s));
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

/* Betweenness centrality over many sources, 64 per lane batch; the same sums
 * as sg_bc (bc.hpp:118-200 per source) up to FP64 rounding. */
extern "C" int sg_bc_batched(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, float* bc) {
  SG_CHECK(g && bc && (nsrc == 0 || sources_host), SG_ERR_INVALID, "sg_bc_batched: null argument");
  SG_CHECK(nsrc >= 0, SG_ERR_INVALID, "sg_bc_batched: negative source count");
  for (int32_t i = 0; i < nsrc; ++i)
    SG_CHECK(sources_host[i] >= 0 && sources_host[i] < g->n, SG_ERR_INVALID,
             "sg_bc_batched: source %d out of range [0,%lld)", sources_host[i], (long long)g->n);
  if (nsrc == 0) return SG_OK;
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return run_bc_batched<int64_t>(g, nsrc, sources_host, bc);
  return run_bc_batched<uint32_t>(g, nsrc, sources_host, bc);
}
