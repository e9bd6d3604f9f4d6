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

namespace {

constexpr int kBlock = 256;
constexpr int kSlots = 64;
constexpr int kPullHead = 16; /* ids a lane scans alone before the wave helps */

struct alignas(16) QEntry {
  int64_t start;
  int32_t v;
  int32_t len;
};

/* one cache line (128 B) per slot */
struct alignas(128) StatSlot {
  unsigned long long disc, disc_deg, edges, scanned, max_deg;
  unsigned long long pad[11];
};

/* per-64-word tile of the frontier compaction */
struct alignas(32) TileCount {
  uint32_t c[3]; /* queue entries per class (large class counts chunks) */
  uint32_t nvert;
  unsigned long long deg;
  unsigned long long maxdeg;
};

struct QueueHdr {
  uint32_t count[3];
  uint32_t pad;
  unsigned long long nvert, deg, maxdeg;
};

struct BfsTuning {
  int32_t t_medium; /* degree >= t_medium -> wave per vertex */
  int32_t t_large;  /* degree >= t_large -> workgroup per chunk */
  int32_t chunk;    /* edges per large-class entry */
};

template<typename OffT>
struct BfsArgs {
  const OffT* off; /* forward CSR (push, queue entries, degrees) */
  const int32_t* cols;
  const OffT* in_off; /* inverse CSR (pull); aliases off when undirected */
  const int32_t* in_cols;
  int32_t directed;
  int64_t n;
  int64_t w0;
  uint64_t* visited;
  uint64_t* prev; /* visited as of the last compaction */
  uint64_t* cur_bm;
  uint64_t* next_bm;
  int32_t* depth;
  QEntry* qin[3];
  const QueueHdr* qin_hdr;
  QEntry* qout[3];
  QueueHdr* qout_hdr;
  TileCount* tiles;
  int64_t ntiles;
  StatSlot* stats;
  StatSlot* stats_zero;
  int32_t level;
  BfsTuning tune;
};

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o, 64);
    if (lane_id() >= o) x += t;
  }
  return x;
}

/* start offset of lane `src` (src may differ per lane) */
__device__ __forceinline__ int64_t shfl_start(int64_t x, int src) {
  const uint32_t lo = __shfl((uint32_t)(x & 0xFFFFFFFFll), src, 64);
  const uint32_t hi = __shfl((uint32_t)((uint64_t)x >> 32), src, 64);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ void zero_slots(StatSlot* s) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(s);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < kSlots * 16; i += gridDim.x * kBlock) p[i] = 0ull;
}

/* Workgroup-reduce five counters and add them into slot blockIdx % kSlots.
 * All threads of the block must call. */
__device__ __forceinline__ void flush_block_stats(StatSlot* slots, uint64_t disc, uint64_t disc_deg, uint64_t edges,
                                                  uint64_t scanned, uint64_t max_deg) {
  __shared__ unsigned long long red[4][5];
  disc = wave_sum_u64(disc);
  disc_deg = wave_sum_u64(disc_deg);
  edges = wave_sum_u64(edges);
  scanned = wave_sum_u64(scanned);
  max_deg = wave_max_u64(max_deg);
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 0) {
    red[wv][0] = disc;
    red[wv][1] = disc_deg;
    red[wv][2] = edges;
    red[wv][3] = scanned;
    red[wv][4] = max_deg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
      t[0] += red[i][0];
      t[1] += red[i][1];
      t[2] += red[i][2];
      t[3] += red[i][3];
      t[4] = red[i][4] > t[4] ? red[i][4] : t[4];
    }
    StatSlot* s = slots + (blockIdx.x % kSlots);
    if (t[0]) atomic_add(&s->disc, t[0]);
    if (t[1]) atomic_add(&s->disc_deg, t[1]);
    if (t[2]) atomic_add(&s->edges, t[2]);
    if (t[3]) atomic_add(&s->scanned, t[3]);
    if (t[4]) atomic_max(&s->max_deg, t[4]);
  }
}

template<typename OffT>
__device__ __forceinline__ void push_visit(const BfsArgs<OffT>& a, uint32_t y) {
  if (claim_bit(a.visited, y)) a.depth[y] = a.level + 1;
}

/* Push level: large-class chunks per workgroup, medium vertices per wave,
 * small vertices per lane. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) push_kernel(BfsArgs<OffT> a) {
  zero_slots(a.stats_zero);
  const uint32_t n_small = a.qin_hdr->count[0], n_med = a.qin_hdr->count[1], n_large = a.qin_hdr->count[2];
  const int lane = lane_id();
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;

  for (uint32_t i = blockIdx.x; i < n_large; i += gridDim.x) {
    const QEntry qe = a.qin[2][i];
    for (int64_t j = wave_in_block * 64 + lane; j < qe.len; j += kBlock) push_visit(a, (uint32_t)a.cols[qe.start + j]);
  }
  for (int64_t i = gwave; i < n_med; i += nwaves) {
    const QEntry qe = a.qin[1][i];
    for (int64_t j = lane; j < qe.len; j += 64) push_visit(a, (uint32_t)a.cols[qe.start + j]);
  }
  /* thread bin as a wave-level merge path: the arcs of 64 small vertices are
   * laid end to end and walked 64 at a time; each lane finds the owner of its
   * arc by a binary search over the wave's inclusive length scan, read from
   * the owning lanes' registers. */
  for (int64_t base = gwave * 64; base < n_small; base += nwaves * 64) {
    const int64_t i = base + lane;
    QEntry qe;
    qe.start = 0;
    qe.v = 0;
    qe.len = 0;
    if (i < n_small) qe = a.qin[0][i];
    const uint32_t incl = wave_inclusive_scan((uint32_t)qe.len);
    const uint32_t total = shfl(incl, 63);
    for (uint32_t t0 = 0; t0 < total; t0 += 64) {
      const uint32_t t = t0 + lane;
      uint32_t lo = 0, hi = 63; /* first owner lane with incl > t */
#pragma unroll
      for (int it = 0; it < 6; ++it) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t x = __shfl(incl, (int)mid, 64);
        if (x > t) hi = mid;
        else lo = mid + 1;
      }
      const uint32_t owner = lo;
      const uint32_t before = __shfl(incl - (uint32_t)qe.len, (int)owner, 64);
      const int64_t st = shfl_start(qe.start, (int)owner);
      if (t < total) push_visit(a, (uint32_t)a.cols[st + (t - before)]);
    }
  }
}

/* Position of the r-th (0-based) set bit of x (x has more than r bits set). */
__device__ __forceinline__ uint32_t select_bit(uint64_t x, uint32_t r) {
  uint32_t pos = 0;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) {
    const uint64_t lowmask = (w == 64) ? ~0ull : ((1ull << w) - 1ull);
    const uint32_t c = (uint32_t)__popcll(x & lowmask);
    if (r >= c) {
      r -= c;
      x >>= w;
      pos += w;
    }
  }
  return pos;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) {
  const uint32_t lo = __shfl((uint32_t)(x & 0xFFFFFFFFull), src, 64);
  const uint32_t hi = __shfl((uint32_t)(x >> 32), src, 64);
  return ((uint64_t)hi << 32) | lo;
}

/* Pull level.  A wave loads 64 visited words at once and compacts their
 * unvisited vertices in registers (popcount scan + bit select), so all 64 lanes
 * pull for real vertices; new frontier bits are gathered per word in LDS. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) pull_kernel(BfsArgs<OffT> a) {
  zero_slots(a.stats_zero);
  __shared__ unsigned long long lds_new[kBlock];
  uint64_t st_disc = 0, st_deg = 0, st_edges = 0, st_scanned = 0, st_max = 0;
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
      bool found = false;
      OffT b = 0, e = 0, j = 0;
      int64_t kk = 0;
      if (active) {
        b = a.in_off[v];
        e = a.in_off[v + 1];
        j = b;
        kk = (int64_t)(e - b);
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
        const int64_t jb = shfl_start((int64_t)j, k2);
        const int64_t je = shfl_start((int64_t)e, k2);
        const int64_t bb = shfl_start((int64_t)b, k2);
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
        st_edges += kk;
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
  flush_block_stats(a.stats, st_disc, st_deg, st_edges, st_scanned, st_max);
}

template<typename OffT>
__device__ __forceinline__ int degree_class(const BfsTuning& t, int64_t d) {
  return d >= t.t_large ? 2 : (d >= t.t_medium ? 1 : 0);
}

/* Frontier compaction, pass 1: one wave per 64-word tile.
 * kDiff: frontier = visited & ~prev (after a push level), next_bm and prev updated.
 * else:  frontier = next_bm (after a pull level), prev := visited. */
template<typename OffT, bool kDiff>
__global__ void __launch_bounds__(kBlock) frontier_count_kernel(BfsArgs<OffT> a) {
  const int64_t tile = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (tile >= a.ntiles) return; /* whole wave leaves together */
  const int lane = lane_id();
  const int64_t myw = tile * 64 + lane;
  uint64_t mynew = 0;
  if (myw < a.w0) {
    const uint64_t vis = a.visited[myw];
    const uint64_t pv = a.prev[myw];
    if (kDiff) {
      mynew = vis & ~pv;
      a.next_bm[myw] = mynew;
    } else {
      mynew = a.next_bm[myw];
    }
    if (vis != pv) a.prev[myw] = vis;
  }
  uint32_t c0 = 0, c1 = 0, c2 = 0, nv = 0;
  uint64_t deg = 0, mx = 0;
  uint64_t todo = ballot(mynew != 0);
  while (todo) {
    const int k = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const uint64_t word = shfl(mynew, k);
    if ((word >> lane) & 1ull) {
      const int64_t v = (tile * 64 + k) * 64 + lane;
      const int64_t d = (int64_t)(a.off[v + 1] - a.off[v]);
      ++nv;
      deg += d;
      mx = (uint64_t)d > mx ? (uint64_t)d : mx;
      if (d > 0) {
        const int cls = degree_class<OffT>(a.tune, d);
        if (cls == 0) ++c0;
        else if (cls == 1) ++c1;
        else c2 += (uint32_t)((d + a.tune.chunk - 1) / a.tune.chunk);
      }
    }
  }
  c0 = (uint32_t)wave_sum_u64(c0);
  c1 = (uint32_t)wave_sum_u64(c1);
  c2 = (uint32_t)wave_sum_u64(c2);
  nv = (uint32_t)wave_sum_u64(nv);
  deg = wave_sum_u64(deg);
  mx = wave_max_u64(mx);
  if (lane == 0) {
    TileCount t;
    t.c[0] = c0;
    t.c[1] = c1;
    t.c[2] = c2;
    t.nvert = nv;
    t.deg = deg;
    t.maxdeg = mx;
    a.tiles[tile] = t;
  }
}

/* Pass 2: one workgroup scans the per-tile class counts (in place, exclusive)
 * and writes the totals into the queue header. */
__global__ void __launch_bounds__(kBlock) frontier_scan_kernel(TileCount* tiles, int64_t ntiles, QueueHdr* hdr) {
  __shared__ uint32_t wave_tot[4][3];
  __shared__ unsigned long long red[4][3];
  uint32_t carry[3] = {0, 0, 0};
  unsigned long long nv = 0, deg = 0, mx = 0;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (int64_t base = 0; base < ntiles; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    TileCount t{};
    if (i < ntiles) t = tiles[i];
    nv += t.nvert;
    deg += t.deg;
    mx = t.maxdeg > mx ? t.maxdeg : mx;
    uint32_t incl[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      incl[c] = wave_inclusive_scan(t.c[c]);
      if (lane == 63) wave_tot[wv][c] = incl[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint32_t before = 0, tot = 0;
      for (int w = 0; w < 4; ++w) {
        if (w < wv) before += wave_tot[w][c];
        tot += wave_tot[w][c];
      }
      if (i < ntiles) tiles[i].c[c] = carry[c] + before + incl[c] - t.c[c];
      carry[c] += tot;
    }
    __syncthreads();
  }
  nv = wave_sum_u64(nv);
  deg = wave_sum_u64(deg);
  mx = wave_max_u64(mx);
  if (lane == 0) {
    red[wv][0] = nv;
    red[wv][1] = deg;
    red[wv][2] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    QueueHdr h{};
    h.count[0] = carry[0];
    h.count[1] = carry[1];
    h.count[2] = carry[2];
    for (int w = 0; w < 4; ++w) {
      h.nvert += red[w][0];
      h.deg += red[w][1];
      h.maxdeg = red[w][2] > h.maxdeg ? red[w][2] : h.maxdeg;
    }
    *hdr = h;
  }
}

/* Pass 3: emit the class queues in vertex order at the scanned offsets. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) frontier_write_kernel(BfsArgs<OffT> a) {
  const int64_t tile = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (tile >= a.ntiles) return;
  const int lane = lane_id();
  const int64_t myw = tile * 64 + lane;
  const uint64_t mynew = myw < a.w0 ? a.next_bm[myw] : 0ull;
  const TileCount tc = a.tiles[tile];
  uint32_t pos0 = tc.c[0], pos1 = tc.c[1], pos2 = tc.c[2];
  uint64_t todo = ballot(mynew != 0);
  while (todo) {
    const int k = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const uint64_t word = shfl(mynew, k);
    const int64_t v = (tile * 64 + k) * 64 + lane;
    const bool has = (word >> lane) & 1ull;
    int64_t s = 0, d = 0;
    if (has) {
      s = (int64_t)a.off[v];
      d = (int64_t)(a.off[v + 1] - a.off[v]);
    }
    const int cls = (has && d > 0) ? degree_class<OffT>(a.tune, d) : -1;
    const uint64_t m0 = ballot(cls == 0), m1 = ballot(cls == 1);
    QEntry qe;
    qe.start = s;
    qe.v = (int32_t)v;
    qe.len = (int32_t)d;
    if (cls == 0) a.qout[0][pos0 + lanes_below(m0)] = qe;
    if (cls == 1) a.qout[1][pos1 + lanes_below(m1)] = qe;
    pos0 += (uint32_t)__popcll(m0);
    pos1 += (uint32_t)__popcll(m1);
    const uint32_t nch = cls == 2 ? (uint32_t)((d + a.tune.chunk - 1) / a.tune.chunk) : 0u;
    const uint32_t incl = wave_inclusive_scan(nch);
    if (nch) {
      const uint32_t p = pos2 + incl - nch;
      for (uint32_t c = 0; c < nch; ++c) {
        QEntry ce;
        ce.start = s + (int64_t)c * a.tune.chunk;
        ce.v = (int32_t)v;
        const int64_t rem = d - (int64_t)c * a.tune.chunk;
        ce.len = (int32_t)(rem < a.tune.chunk ? rem : a.tune.chunk);
        a.qout[2][p + c] = ce;
      }
    }
    pos2 += shfl(incl, 63);
  }
}

/* Level-0 setup: depth[src] = 0, visited = prev = isolated | {src},
 * cur_bm = {src}, F0 queue = {src}. */
template<typename OffT>
__global__ void init_source_kernel(BfsArgs<OffT> a, int32_t src) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t bit = 1ull << (src & 63);
  a.depth[src] = 0;
  a.visited[src >> 6] |= bit;
  a.prev[src >> 6] |= bit;
  a.cur_bm[src >> 6] |= bit;
  const int64_t s = (int64_t)a.off[src], d = (int64_t)(a.off[src + 1] - a.off[src]);
  QueueHdr h{};
  if (d > 0) {
    const int cls = degree_class<OffT>(a.tune, d);
    if (cls < 2) {
      QEntry qe;
      qe.start = s;
      qe.v = src;
      qe.len = (int32_t)d;
      a.qout[cls][0] = qe;
      h.count[cls] = 1;
    } else {
      const uint32_t nch = (uint32_t)((d + a.tune.chunk - 1) / a.tune.chunk);
      for (uint32_t c = 0; c < nch; ++c) {
        QEntry ce;
        ce.start = s + (int64_t)c * a.tune.chunk;
        ce.v = src;
        const int64_t rem = d - (int64_t)c * a.tune.chunk;
        ce.len = (int32_t)(rem < a.tune.chunk ? rem : a.tune.chunk);
        a.qout[2][c] = ce;
      }
      h.count[2] = nch;
    }
  }
  h.nvert = 1;
  h.deg = (unsigned long long)d;
  h.maxdeg = (unsigned long long)d;
  *a.qout_hdr = h;
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
  int64_t n = 0, w0 = 0, ntiles = 0, cap_large = 0;
  uint64_t* visited = nullptr;
  uint64_t* prev = nullptr;
  uint64_t* bm[2] = {nullptr, nullptr};
  QEntry* q[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  QueueHdr* hdr = nullptr;   /* 2 headers */
  TileCount* tiles = nullptr;
  StatSlot* stats = nullptr; /* 2 areas of kSlots */
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

void ws_free(BfsWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->visited);
  (void)hipFree(w->prev);
  (void)hipFree(w->bm[0]);
  (void)hipFree(w->bm[1]);
  for (auto& s : w->q)
    for (auto* p : s) (void)hipFree(p);
  (void)hipFree(w->hdr);
  (void)hipFree(w->tiles);
  (void)hipFree(w->stats);
  for (auto e : w->ev)
    if (e) (void)hipEventDestroy(e);
  delete w;
}

BfsTuning default_tuning() { return BfsTuning{64, 1024, 1024}; }

int workspace_for(sg_graph_t g, BfsWorkspace** out) {
  if (g->bfs_ws) {
    *out = static_cast<BfsWorkspace*>(g->bfs_ws);
    return SG_OK;
  }
  auto* w = new BfsWorkspace();
  w->n = g->n;
  w->w0 = (g->n + 63) / 64;
  w->ntiles = (w->w0 + 63) / 64;
  const BfsTuning t = default_tuning();
  w->cap_large = g->m / t.chunk + g->m / t.t_large + 64;
  int st = SG_OK;
  sg_device_t dev = g->dev;
  if ((st = sg::alloc_device(dev, (void**)&w->visited, 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->prev, 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->bm[0], 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->bm[1], 8 * (w->w0 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->hdr, sizeof(QueueHdr) * 2)) ||
      (st = sg::alloc_device(dev, (void**)&w->tiles, sizeof(TileCount) * (w->ntiles + 1))) ||
      (st = sg::alloc_device(dev, (void**)&w->stats, sizeof(StatSlot) * kSlots * 2))) {
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
  uint64_t disc = 0, disc_deg = 0, edges = 0, scanned = 0, max_deg = 0;
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
  const int64_t n = g->n, w0 = ws->w0;
  const int grid_main = dev->cus * 8;
  const int grid_small = std::max(1, std::min(dev->cus * 4, (int)((w0 + 255) / 256)));
  const int grid_tiles = (int)((ws->ntiles * 64 + kBlock - 1) / kBlock);

  BfsArgs<OffT> a{};
  a.off = (const OffT*)g->offsets;
  a.cols = g->cols;
  a.in_off = (const OffT*)g->inv_offsets;
  a.in_cols = g->inv_cols;
  a.directed = (g->flags & SG_GRAPH_DIRECTED) ? 1 : 0;
  a.n = n;
  a.w0 = w0;
  a.visited = ws->visited;
  a.prev = ws->prev;
  a.depth = depth;
  a.tiles = ws->tiles;
  a.ntiles = ws->ntiles;
  a.tune = default_tuning();
  int cb = 0; /* bitmap holding the current frontier */
  int qs = 0; /* queue set holding the current push frontier */

  fill_i32_kernel<<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(depth, (int32_t)(n + 1), n);
  copy_words_kernel<<<grid_small, 256, 0, s>>>(ws->visited, ws->prev, g->isolated, w0);
  SG_TRY(sg::zero_async(dev, ws->bm[0], 8 * w0));
  SG_TRY(sg::zero_async(dev, ws->stats, sizeof(StatSlot) * kSlots * 2));
  a.cur_bm = ws->bm[0];
  for (int k = 0; k < 3; ++k) a.qout[k] = ws->q[0][k];
  a.qout_hdr = ws->hdr + 0;
  init_source_kernel<OffT><<<1, 64, 0, s>>>(a, src);
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
  /* F0 = {src}: read its degree (counts as explored) */
  QueueHdr* hh = (QueueHdr*)dev->host_scratch;
  StatSlot* hslots = (StatSlot*)((char*)dev->host_scratch + 4096);
  SG_HIP(hipMemcpyAsync(hh, ws->hdr, sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
  SG_HIP(hipStreamSynchronize(s));
  uint64_t visited_deg = hh->deg;
  int64_t f_vertices = 1, f_edges = (int64_t)hh->deg;
  int64_t reached_v = 1;
  for (;; ++level) {
    a.level = level;
    a.stats = ws->stats + kSlots * (level & 1);
    a.stats_zero = ws->stats + kSlots * ((level + 1) & 1);
    for (int k = 0; k < 3; ++k) {
      a.qin[k] = ws->q[qs][k];
      a.qout[k] = ws->q[qs ^ 1][k];
    }
    a.qin_hdr = ws->hdr + qs;
    a.qout_hdr = ws->hdr + (qs ^ 1);
    a.cur_bm = ws->bm[cb];
    a.next_bm = ws->bm[cb ^ 1];
    const bool timed = opt.record_levels && stats && level < SG_MAX_LEVELS_RECORDED;
    if (timed) SG_HIP(hipEventRecord(ws->ev[0], s));
    HostStats c;
    if (push) {
      push_kernel<OffT><<<grid_main, kBlock, 0, s>>>(a);
      if (timed) SG_HIP(hipEventRecord(ws->ev[1], s));
      /* new frontier = visited & ~prev, counted per class */
      frontier_count_kernel<OffT, true><<<grid_tiles, kBlock, 0, s>>>(a);
      frontier_scan_kernel<<<1, kBlock, 0, s>>>(ws->tiles, ws->ntiles, a.qout_hdr);
      SG_HIP(hipGetLastError());
      SG_HIP(hipMemcpyAsync(hh, a.qout_hdr, sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
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
    /* the new frontier is bm[cb^1] in both cases */
    cb ^= 1;
    if (next_push) {
      a.next_bm = ws->bm[cb];
      if (!push) {
        frontier_count_kernel<OffT, false><<<grid_tiles, kBlock, 0, s>>>(a);
        frontier_scan_kernel<<<1, kBlock, 0, s>>>(ws->tiles, ws->ntiles, a.qout_hdr);
      }
      frontier_write_kernel<OffT><<<grid_tiles, kBlock, 0, s>>>(a);
      SG_HIP(hipGetLastError());
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
