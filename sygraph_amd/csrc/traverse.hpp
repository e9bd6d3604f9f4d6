/*
 * traverse.hpp -- pieces shared by the traversal kernels (BFS, SSSP, CC,
 * batched BC): degree-class frontier queues, the ordered bitmap -> queue
 * compaction, spread statistics slots and wave64 helpers.
 *
 * A frontier moves between levels as a bitmap (one uint64 word per wave,
 * lane i <-> bit i).  When a level needs it as work items, an ordered
 * set-scan-compact turns it into three queues by degree class -- the
 * thread / wavefront / workgroup bins of the bucketing load balancer
 * (operators/advance/bucketing.hpp:100-178):
 *   class 0  degree <  t_medium : one lane per vertex (wave merge path)
 *   class 1  degree <  t_large  : one wave per vertex
 *   class 2  larger             : one workgroup per `chunk` arcs
 * Each entry carries (vertex, first arc, arc count).  Compaction is three
 * passes -- per-64-word tile counts, one scan, write -- so no counter is ever
 * contended and queue order is deterministic (ascending vertex id).
 */
#pragma once

#include "internal.hpp"

namespace sg::trav {

using namespace sygraph::hipdev;

constexpr int kBlock = 256;
constexpr int kSlots = 64;

struct alignas(16) QEntry {
  int64_t start;
  int32_t v;
  int32_t len;
};

/* one 128-byte cache line per slot */
struct alignas(128) StatSlot {
  unsigned long long disc, disc_deg, edges, scanned, max_deg, misses;
  unsigned long long pad[10];
};

/* per-64-word tile of the compaction */
struct alignas(32) TileCount {
  uint32_t c[3]; /* queue entries per class (class 2 counts chunks) */
  uint32_t nvert;
  unsigned long long deg;
  unsigned long long maxdeg;
};

/* Agent-coherent stores and loads: relaxed atomics at agent scope, which
 * write through to / read from the point of coherence that every XCD's L2
 * shares, so one workgroup can hand data to another within a launch without
 * an L2 write-back or invalidate. */
__device__ __forceinline__ void store_coherent(TileCount* p, const TileCount& t) {
  uint64_t w[4];
  __builtin_memcpy(w, &t, sizeof(w));
  uint64_t* q = reinterpret_cast<uint64_t*>(p);
#pragma unroll
  for (int k = 0; k < 4; ++k) __hip_atomic_store(q + k, w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ TileCount load_coherent(const TileCount* p) {
  uint64_t w[4];
  const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  TileCount t;
  __builtin_memcpy(&t, w, sizeof(w));
  return t;
}

/* Wait until this wave's stores have completed (s_waitcnt vmcnt(0)); the
 * signal fences keep the compiler from moving memory operations across. */
__device__ __forceinline__ void wait_stores_complete() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70); /* gfx9 encoding: vmcnt 0; expcnt 7 and lgkmcnt 15 are no-waits */
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

/* True in exactly one workgroup of the grid, the last to get here; it also
 * resets the counter for the next launch.  All threads call.  What the last
 * workgroup may then read of the grid's earlier writes:
 *  - kCoherent: what was written with agent-coherent stores (every wave
 *    waits for its stores to complete before the arrival is counted), read
 *    with agent-coherent loads; nothing is written back or invalidated;
 *  - otherwise: the writes of threads that pass `publish` (they release at
 *    agent scope, an L2 write-back, so only threads with data for the last
 *    workgroup should), read with ordinary loads after the last workgroup's
 *    acquire (an L2 invalidate). */
constexpr int kArriveGroups = 8;  /* workgroups are dealt to the 8 XCDs round robin */
constexpr int kArriveStride = 16; /* counters 64 B apart */
constexpr int kArriveWords = (kArriveGroups + 1) * kArriveStride;

/* Arrivals are counted in two stages, per group of workgroups (blockIdx.x %
 * kArriveGroups: one counter per XCD) and then per group, so no single
 * counter takes more than 1/8 of the grid's atomics.  `ctr` holds
 * kArriveWords zeroed words. */
template<bool kCoherent>
__device__ __forceinline__ bool last_to_arrive(uint32_t* ctr, bool publish) {
  __shared__ uint32_t s_last;
  if (kCoherent) wait_stores_complete();
  else if (publish) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t grid = gridDim.x, g = blockIdx.x % kArriveGroups;
    const uint32_t ngroups = grid < (uint32_t)kArriveGroups ? grid : (uint32_t)kArriveGroups;
    const uint32_t in_group = (grid - g + kArriveGroups - 1) / kArriveGroups;
    uint32_t* gc = ctr + (1 + g) * kArriveStride;
    uint32_t last = 0;
    if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == in_group - 1) {
      __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ngroups - 1 ? 1u : 0u;
      if (last) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_last = last;
  }
  __syncthreads();
  const bool last = s_last != 0;
  if (!kCoherent && last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return last;
}

struct QueueHdr {
  uint32_t count[3];
  uint32_t pad;
  unsigned long long nvert, deg, maxdeg;
};

struct Tuning {
  int32_t t_medium; /* degree >= t_medium -> wave per vertex */
  int32_t t_large;  /* degree >= t_large -> workgroup per chunk */
  int32_t chunk;    /* arcs per class-2 entry */
};

inline Tuning default_tuning() { return Tuning{64, 1024, 1024}; }

/* ------------------------------------------------------------ helpers --- */
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o, 64);
    if (lane_id() >= o) x += t;
  }
  return x;
}

/* 64-bit value of lane `src` (src may differ per lane; all lanes active) */
__device__ __forceinline__ int64_t shfl_i64(int64_t x, int src) {
  const uint32_t lo = __shfl((uint32_t)(x & 0xFFFFFFFFll), src, 64);
  const uint32_t hi = __shfl((uint32_t)((uint64_t)x >> 32), src, 64);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) { return (uint64_t)shfl_i64((int64_t)x, src); }

/* Position of the r-th (0-based) set bit of x; x has more than r bits set. */
__device__ __forceinline__ uint32_t select_bit(uint64_t x, uint32_t r) {
  uint32_t pos = 0;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) {
    const uint32_t c = (uint32_t)__popcll(x & ((1ull << w) - 1ull));
    if (r >= c) {
      r -= c;
      x >>= w;
      pos += w;
    }
  }
  return pos;
}

/* owner lane of flattened item t: first lane whose inclusive scan exceeds t
 * (whole wave must call; t may exceed the total, then the result is 63) */
__device__ __forceinline__ int scan_owner(uint32_t incl, uint32_t t) {
  uint32_t lo = 0, hi = 63;
#pragma unroll
  for (int it = 0; it < 6; ++it) {
    const uint32_t mid = (lo + hi) >> 1;
    if (__shfl(incl, (int)mid, 64) > t) hi = mid;
    else lo = mid + 1;
  }
  return (int)lo;
}

__device__ __forceinline__ void zero_slots(StatSlot* s) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(s);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < kSlots * 16; i += gridDim.x * kBlock) p[i] = 0ull;
}

/* Workgroup-reduce six counters and add them into slot blockIdx % kSlots.
 * Every thread of the (kBlock-sized) workgroup must call. */
__device__ __forceinline__ void flush_block_stats(StatSlot* slots, uint64_t disc, uint64_t disc_deg, uint64_t edges,
                                                  uint64_t scanned, uint64_t max_deg, uint64_t misses = 0) {
  __shared__ unsigned long long red[kBlock / 64][6];
  disc = wave_sum_u64(disc);
  disc_deg = wave_sum_u64(disc_deg);
  edges = wave_sum_u64(edges);
  scanned = wave_sum_u64(scanned);
  max_deg = wave_max_u64(max_deg);
  misses = wave_sum_u64(misses);
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 0) {
    red[wv][0] = disc;
    red[wv][1] = disc_deg;
    red[wv][2] = edges;
    red[wv][3] = scanned;
    red[wv][4] = max_deg;
    red[wv][5] = misses;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < kBlock / 64; ++i) {
      t[0] += red[i][0];
      t[1] += red[i][1];
      t[2] += red[i][2];
      t[3] += red[i][3];
      t[4] = red[i][4] > t[4] ? red[i][4] : t[4];
      t[5] += red[i][5];
    }
    StatSlot* s = slots + (blockIdx.x % kSlots);
    if (t[0]) atomic_add(&s->disc, t[0]);
    if (t[1]) atomic_add(&s->disc_deg, t[1]);
    if (t[2]) atomic_add(&s->edges, t[2]);
    if (t[3]) atomic_add(&s->scanned, t[3]);
    if (t[4]) atomic_max(&s->max_deg, t[4]);
    if (t[5]) atomic_add(&s->misses, t[5]);
  }
}

__device__ __forceinline__ int degree_class(const Tuning& t, int64_t d) {
  return d >= t.t_large ? 2 : (d >= t.t_medium ? 1 : 0);
}

/* A wave's walk over the set bits of its 64 words (lane i holds word i), in
 * vertex order, one vertex per lane and step.  Word by word (a step per
 * non-empty word, lane = bit) suits dense words; flattened (a step per 64 set
 * bits, the owner word found by a search of the popcount scan) suits sparse
 * ones; the walk takes the one with fewer steps.  All lanes call. */
struct TileWalk {
  uint64_t words;  /* ballot of the non-empty words (word by word) */
  uint32_t incl;   /* inclusive scan of the popcounts (flattened) */
  uint32_t pc;     /* this lane's popcount */
  uint32_t total;  /* set bits of the tile */
  uint32_t steps;
  bool flat;
};

__device__ __forceinline__ TileWalk tile_walk(uint64_t mine) {
  TileWalk w;
  w.words = ballot(mine != 0);
  w.pc = (uint32_t)__popcll(mine);
  w.incl = wave_inclusive_scan(w.pc);
  w.total = __shfl(w.incl, 63, 64);
  const uint32_t flat_steps = (w.total + 63) / 64, word_steps = (uint32_t)__popcll(w.words);
  w.flat = flat_steps < word_steps;
  w.steps = w.flat ? flat_steps : word_steps;
  return w;
}

/* Vertex of this lane at `step` (false: none); steps in increasing order. */
__device__ __forceinline__ bool tile_walk_item(TileWalk& w, uint64_t mine, int64_t tile, uint32_t step, int64_t& v) {
  const int lane = lane_id();
  if (w.flat) {
    const uint32_t t = step * 64 + lane;
    const int owner = scan_owner(w.incl, t);
    const uint32_t before = __shfl(w.incl - w.pc, owner, 64);
    const uint64_t word = shfl_u64(mine, owner);
    if (t >= w.total) return false;
    v = (tile * 64 + owner) * 64 + select_bit(word, t - before);
    return true;
  }
  const int k = __ffsll((unsigned long long)w.words) - 1; /* the step-th non-empty word */
  w.words &= w.words - 1;
  const uint64_t word = shfl_u64(mine, k);
  v = (tile * 64 + k) * 64 + lane;
  return (word >> lane) & 1ull;
}

/* ------------------------------------------------------- compaction --- */
enum CompactMode {
  kFromDiff = 0, /* frontier = visited & ~prev; writes frontier bitmap and prev */
  kFromBitmapSyncPrev = 1, /* frontier = bitmap; prev := visited */
  kFromBitmap = 2 /* frontier = bitmap */
};

template<typename OffT>
struct CompactArgs {
  const OffT* off;
  int64_t w0, ntiles;
  const uint64_t* visited; /* kFromDiff / kFromBitmapSyncPrev */
  uint64_t* prev;          /* kFromDiff / kFromBitmapSyncPrev */
  uint64_t* frontier;      /* bitmap: written (kFromDiff) or read */
  TileCount* tiles;        /* per wave (64-word tile): class counts */
  TileCount* blocks;       /* per workgroup (4 tiles): sums, then exclusive offsets after the scan */
  QEntry* qout[3];
  QueueHdr* hdr;
  Tuning tune;
  const int32_t* split; /* optional: per vertex, the length of the first part of its list (SSSP's light
                           arcs); the entries then cover the first part (phase 0) or the rest (phase 1) */
  int32_t* depth;       /* optional (kFromDiff): depth of the new frontier's vertices, written here in
                           vertex order rather than scattered by the push that discovered them */
  int32_t split_k;      /* without `split`: a first part of min(split_k, degree) arcs (0: none) */
};

/* The arc range [s, s + d) of vertex v's queue entries. */
template<typename OffT>
__device__ __forceinline__ void entry_range(const CompactArgs<OffT>& a, int64_t v, int phase, int64_t& s, int64_t& d) {
  s = (int64_t)a.off[v];
  d = (int64_t)(a.off[v + 1] - a.off[v]);
  if (a.split || a.split_k > 0) {
    const int64_t first = a.split ? (int64_t)a.split[v] : (d < a.split_k ? d : (int64_t)a.split_k);
    if (phase) {
      s += first;
      d -= first;
    } else {
      d = first;
    }
  }
}

/* pass 1: one wave per 64-word tile; the workgroup also sums its four tiles
 * (all threads of the workgroup must call) */
template<typename OffT, int kMode>
__device__ __forceinline__ void compact_count_body(const CompactArgs<OffT>& a, int phase = 0, int32_t depth_val = 0) {
  __shared__ TileCount s_wave[kBlock / 64];
  const int64_t tile = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6; /* tiles >= ntiles see only words >= w0 */
  const int lane = lane_id();
  const int64_t myw = tile * 64 + lane;
  uint64_t mynew = 0;
  if (myw < a.w0) {
    if (kMode == kFromBitmap) {
      mynew = a.frontier[myw];
    } else {
      const uint64_t vis = a.visited[myw];
      const uint64_t pv = a.prev[myw];
      if (kMode == kFromDiff) {
        mynew = vis & ~pv;
        a.frontier[myw] = mynew;
      } else {
        mynew = a.frontier[myw];
      }
      if (vis != pv) a.prev[myw] = vis;
    }
  }
  uint32_t c0 = 0, c1 = 0, c2 = 0, nv = 0;
  uint64_t deg = 0, mx = 0;
  /* the tile's set bits in vertex order, walked word by word (lane = bit) or,
   * when that takes fewer steps, flattened (a scan of the words' popcounts:
   * 64 vertices per step however sparse the words are) */
  TileWalk tw = tile_walk(mynew);
  for (uint32_t step = 0; step < tw.steps; ++step) {
    int64_t v;
    if (tile_walk_item(tw, mynew, tile, step, v)) {
      if (kMode == kFromDiff && a.depth) a.depth[v] = depth_val;
      int64_t s0, d;
      entry_range(a, v, phase, s0, d);
      ++nv;
      deg += d;
      mx = (uint64_t)d > mx ? (uint64_t)d : mx;
      if (d > 0) {
        const int cls = degree_class(a.tune, d);
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
    if (tile < a.ntiles) a.tiles[tile] = t;
    s_wave[threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    TileCount b{};
    for (int w = 0; w < kBlock / 64; ++w) {
      for (int c = 0; c < 3; ++c) b.c[c] += s_wave[w].c[c];
      b.nvert += s_wave[w].nvert;
      b.deg += s_wave[w].deg;
      b.maxdeg = s_wave[w].maxdeg > b.maxdeg ? s_wave[w].maxdeg : b.maxdeg;
    }
    store_coherent(&a.blocks[blockIdx.x], b); /* may be scanned by the last workgroup of this launch */
  }
}

template<typename OffT, int kMode>
__global__ void __launch_bounds__(kBlock) compact_count_kernel(CompactArgs<OffT> a, int phase) {
  compact_count_body<OffT, kMode>(a, phase);
}


/* pass 2: one 1024-thread workgroup scans the per-workgroup class counts of
 * pass 1 (in place, exclusive) and writes the totals into the queue header.
 * Each thread owns kScanPer consecutive entries, so one workgroup scan covers
 * 2,048 entries (8,192 tiles, 33.5 M vertices); larger graphs loop with a
 * carry. */
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 2;

/* Must be called by all kThreads threads of the workgroup; thread 0 gets the
 * totals in *out (also stored to *hdr).  kThreads * kPer entries per pass. */
template<int kThreads, int kPer, bool kCoherent = false>
__device__ __forceinline__ void compact_scan_body_t(TileCount* tiles, int64_t ntiles, QueueHdr* hdr, QueueHdr* out) {
  __shared__ uint32_t wave_tot[kThreads / 64][3];
  __shared__ unsigned long long red[kThreads / 64][3];
  uint32_t carry[3] = {0, 0, 0};
  unsigned long long nv = 0, deg = 0, mx = 0;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (int64_t base = 0; base < ntiles; base += (int64_t)kThreads * kPer) {
    const int64_t first = base + (int64_t)threadIdx.x * kPer;
    uint32_t loc[kPer][3];
    uint32_t sum[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      TileCount t{};
      if (first + k < ntiles) t = kCoherent ? load_coherent(&tiles[first + k]) : tiles[first + k];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        loc[k][c] = t.c[c];
        sum[c] += t.c[c];
      }
      nv += t.nvert;
      deg += t.deg;
      mx = t.maxdeg > mx ? t.maxdeg : mx;
    }
    uint32_t incl[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      incl[c] = wave_inclusive_scan(sum[c]);
      if (lane == 63) wave_tot[wv][c] = incl[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint32_t before = 0, tot = 0;
      for (int w = 0; w < kThreads / 64; ++w) {
        if (w < wv) before += wave_tot[w][c];
        tot += wave_tot[w][c];
      }
      uint32_t run = carry[c] + before + incl[c] - sum[c];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        if (first + k < ntiles) tiles[first + k].c[c] = run;
        run += loc[k][c];
      }
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
    for (int w = 0; w < kThreads / 64; ++w) {
      h.nvert += red[w][0];
      h.deg += red[w][1];
      h.maxdeg = red[w][2] > h.maxdeg ? red[w][2] : h.maxdeg;
    }
    *hdr = h;
    if (out) *out = h;
  }
}

__device__ __forceinline__ void compact_scan_body(TileCount* tiles, int64_t ntiles, QueueHdr* hdr, QueueHdr* out) {
  compact_scan_body_t<kScanThreads, kScanPer>(tiles, ntiles, hdr, out);
}

static __global__ void __launch_bounds__(kScanThreads) compact_scan_kernel(TileCount* tiles, int64_t ntiles,
                                                                           QueueHdr* hdr) {
  compact_scan_body(tiles, ntiles, hdr, nullptr);
}

/* pass 3: emit the class queues in vertex order at the scanned offsets.
 * `blk` and `tid` are the workgroup and thread of the count pass's layout
 * (kBlock threads per workgroup), so a wider workgroup can act as several.
 * Returns whether the wave wrote any entry. */
template<typename OffT>
__device__ __forceinline__ bool compact_write_body_at(const CompactArgs<OffT>& a, int phase, int64_t blk, int tid) {
  const int64_t tile = (blk * kBlock + tid) >> 6;
  if (tile >= a.ntiles) return false;
  const int lane = lane_id();
  const int64_t myw = tile * 64 + lane;
  const uint64_t mynew = myw < a.w0 ? a.frontier[myw] : 0ull;
  /* this tile's offsets: its workgroup's exclusive offsets plus the counts of
   * the workgroup's earlier tiles */
  const TileCount bc = a.blocks[blk];
  uint32_t pos0 = bc.c[0], pos1 = bc.c[1], pos2 = bc.c[2];
  for (int64_t k = blk * (kBlock / 64); k < tile; ++k) {
    const TileCount tc = a.tiles[k];
    pos0 += tc.c[0];
    pos1 += tc.c[1];
    pos2 += tc.c[2];
  }
  /* the same walk, so the same vertex order, as the count pass */
  TileWalk tw = tile_walk(mynew);
  const bool wrote = tw.total != 0;
  for (uint32_t step = 0; step < tw.steps; ++step) {
    int64_t v = 0;
    const bool has = tile_walk_item(tw, mynew, tile, step, v);
    int64_t s = 0, d = 0;
    if (has) entry_range(a, v, phase, s, d);
    const int cls = (has && d > 0) ? degree_class(a.tune, d) : -1;
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
    pos2 += __shfl(incl, 63, 64);
  }
  return wrote;
}

template<typename OffT>
__device__ __forceinline__ bool compact_write_body(const CompactArgs<OffT>& a, int phase = 0) {
  return compact_write_body_at(a, phase, (int64_t)blockIdx.x, (int)threadIdx.x);
}

template<typename OffT>
__global__ void __launch_bounds__(kBlock) compact_write_kernel(CompactArgs<OffT> a, int phase) {
  compact_write_body<OffT>(a, phase);
}


/* Single-vertex queue (level-0 source) written by one thread. */
template<typename OffT>
__device__ __forceinline__ void enqueue_single(const OffT* off, const Tuning& t, QEntry* const q[3], QueueHdr* hdr,
                                               int32_t v) {
  const int64_t s = (int64_t)off[v], d = (int64_t)(off[v + 1] - off[v]);
  QueueHdr h{};
  if (d > 0) {
    const int cls = degree_class(t, d);
    if (cls < 2) {
      QEntry qe;
      qe.start = s;
      qe.v = v;
      qe.len = (int32_t)d;
      q[cls][0] = qe;
      h.count[cls] = 1;
    } else {
      const uint32_t nch = (uint32_t)((d + t.chunk - 1) / t.chunk);
      for (uint32_t c = 0; c < nch; ++c) {
        QEntry ce;
        ce.start = s + (int64_t)c * t.chunk;
        ce.v = v;
        const int64_t rem = d - (int64_t)c * t.chunk;
        ce.len = (int32_t)(rem < t.chunk ? rem : t.chunk);
        q[2][c] = ce;
      }
      h.count[2] = nch;
    }
  }
  h.nvert = 1;
  h.deg = (unsigned long long)d;
  h.maxdeg = (unsigned long long)d;
  *hdr = h;
}

/* Visit every queued arc with `visit(src, dst, arc_index)`.  All threads of
 * the workgroup must call.  `variant` picks the work mapping (the result is
 * the same set of visits):
 *   0  class 2 chunks per workgroup, class 1 a wave per vertex, class 0 as a
 *      wave-level merge path (the arcs of 64 small vertices laid end to end,
 *      64 per step) -- the bucketing balancer;
 *   1  classes 0 and 1 both through the merge path (edge-balanced);
 *   2  class 2 chunks per wave instead of per workgroup. */
/* for_each_queued_arc with four arcs per lane per call: visit4(v, e, ok)
 * receives four arc heads, their arc indices and validity, so the visitor can
 * issue the four arcs' memory accesses together (memory-level parallelism
 * for latency-bound levels).  Same work mappings as for_each_queued_arc. */
template<typename OffT, typename Visit4>
__device__ __forceinline__ void for_each_queued_arc4(QEntry* const q[3], const QueueHdr* hdr, const int32_t* cols,
                                                     Visit4&& visit4, int variant = 0) {
  const uint32_t n_small = hdr->count[0], n_med = hdr->count[1], n_large = hdr->count[2];
  const int lane = lane_id();
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  uint32_t v[4];
  int64_t e[4];
  bool ok[4];
  /* one list [start, start + len) walked by `width` threads from position `first` */
  auto strided = [&](const QEntry& qe, int64_t first, int64_t width) {
    for (int64_t j0 = first; j0 < qe.len; j0 += 4 * width) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t j = j0 + k * width;
        ok[k] = j < qe.len;
        e[k] = qe.start + j;
        v[k] = ok[k] ? (uint32_t)cols[e[k]] : 0u;
      }
      visit4(v, e, ok);
    }
  };
  if (variant == 2) {
    for (int64_t i = gwave; i < n_large; i += nwaves) strided(q[2][i], lane, 64);
  } else {
    for (uint32_t i = blockIdx.x; i < n_large; i += gridDim.x) strided(q[2][i], wave_in_block * 64 + lane, kBlock);
  }
  if (variant != 1) {
    for (int64_t i = gwave; i < n_med; i += nwaves) strided(q[1][i], lane, 64);
  }
  for (int cls = 0; cls < (variant == 1 ? 2 : 1); ++cls) {
    const QEntry* qq = q[cls];
    const uint32_t cnt = cls == 0 ? n_small : n_med;
    for (int64_t base = gwave * 64; base < cnt; base += nwaves * 64) {
      const int64_t i = base + lane;
      QEntry qe;
      qe.start = 0;
      qe.v = 0;
      qe.len = 0;
      if (i < cnt) qe = qq[i];
      const uint32_t incl = wave_inclusive_scan((uint32_t)qe.len);
      const uint32_t total = __shfl(incl, 63, 64);
      const uint32_t excl = incl - (uint32_t)qe.len;
      for (uint32_t t0 = 0; t0 < total; t0 += 4 * 64) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t t = t0 + 64 * k + lane;
          const int owner = scan_owner(incl, t);
          const uint32_t before = __shfl(excl, owner, 64);
          const int64_t st = shfl_i64(qe.start, owner);
          ok[k] = t < total;
          e[k] = st + (t - before);
          v[k] = ok[k] ? (uint32_t)cols[e[k]] : 0u;
        }
        visit4(v, e, ok);
      }
    }
  }
}

template<typename OffT, typename Visit>
__device__ __forceinline__ void for_each_queued_arc(QEntry* const q[3], const QueueHdr* hdr, const int32_t* cols,
                                                    Visit&& visit, int variant = 0) {
  const uint32_t n_small = hdr->count[0], n_med = hdr->count[1], n_large = hdr->count[2];
  const int lane = lane_id();
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  if (variant == 2) {
    for (int64_t i = gwave; i < n_large; i += nwaves) {
      const QEntry qe = q[2][i];
      for (int64_t j = lane; j < qe.len; j += 64) {
        const int64_t e = qe.start + j;
        visit((uint32_t)qe.v, (uint32_t)cols[e], e);
      }
    }
  } else {
    for (uint32_t i = blockIdx.x; i < n_large; i += gridDim.x) {
      const QEntry qe = q[2][i];
      for (int64_t j = wave_in_block * 64 + lane; j < qe.len; j += kBlock) {
        const int64_t e = qe.start + j;
        visit((uint32_t)qe.v, (uint32_t)cols[e], e);
      }
    }
  }
  if (variant != 1) {
    for (int64_t i = gwave; i < n_med; i += nwaves) {
      const QEntry qe = q[1][i];
      for (int64_t j = lane; j < qe.len; j += 64) {
        const int64_t e = qe.start + j;
        visit((uint32_t)qe.v, (uint32_t)cols[e], e);
      }
    }
  }
  /* merge path over class 0 (and class 1 for variant 1) */
  for (int cls = 0; cls < (variant == 1 ? 2 : 1); ++cls) {
    const QEntry* qq = q[cls];
    const uint32_t cnt = cls == 0 ? n_small : n_med;
    for (int64_t base = gwave * 64; base < cnt; base += nwaves * 64) {
      const int64_t i = base + lane;
      QEntry qe;
      qe.start = 0;
      qe.v = 0;
      qe.len = 0;
      if (i < cnt) qe = qq[i];
      const uint32_t incl = wave_inclusive_scan((uint32_t)qe.len);
      const uint32_t total = __shfl(incl, 63, 64);
      for (uint32_t t0 = 0; t0 < total; t0 += 64) {
        const uint32_t t = t0 + lane;
        const int owner = scan_owner(incl, t);
        const uint32_t before = __shfl(incl - (uint32_t)qe.len, owner, 64);
        const int64_t st = shfl_i64(qe.start, owner);
        const int32_t src = __shfl(qe.v, owner, 64);
        if (t < total) {
          const int64_t e = st + (t - before);
          visit((uint32_t)src, (uint32_t)cols[e], e);
        }
      }
    }
  }
}

/* --------------------------------------------------------- host side --- */
/* Two sets of class queues (current / next) plus compaction scratch. */
struct QueueSet {
  int64_t w0 = 0, ntiles = 0, cap_large = 0;
  QEntry* q[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  QueueHdr* hdr = nullptr; /* 2 headers */
  TileCount* tiles = nullptr;
  TileCount* blocks = nullptr;
  int64_t nblocks = 0;

  /* largest number of class-2 chunks any frontier can produce under t */
  static int64_t large_capacity(int64_t m, const Tuning& t) { return m / t.chunk + m / t.t_large + 64; }
  bool fits(int64_t m, const Tuning& t) const { return large_capacity(m, t) <= cap_large; }

  int init(sg_device_t dev, int64_t n, int64_t m, const Tuning& t) {
    w0 = (n + 63) / 64;
    ntiles = (w0 + 63) / 64;
    nblocks = (ntiles + kBlock / 64 - 1) / (kBlock / 64);
    cap_large = large_capacity(m, t);
    int st = SG_OK;
    if ((st = alloc_device(dev, (void**)&hdr, sizeof(QueueHdr) * 2)) ||
        (st = alloc_device(dev, (void**)&tiles, sizeof(TileCount) * (ntiles + 1))) ||
        (st = alloc_device(dev, (void**)&blocks, sizeof(TileCount) * (nblocks + 1))))
      return st;
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < 3; ++k)
        if ((st = alloc_device(dev, (void**)&q[s][k], sizeof(QEntry) * (k == 2 ? cap_large : n + 1)))) return st;
    return SG_OK;
  }
  void release() {
    for (auto& s : q)
      for (auto*& p : s) {
        (void)hipFree(p);
        p = nullptr;
      }
    (void)hipFree(hdr);
    (void)hipFree(tiles);
    (void)hipFree(blocks);
    hdr = nullptr;
    tiles = nullptr;
    blocks = nullptr;
  }
  int grid_tiles() const { return (int)nblocks; }
};

/* The queue set kept with the graph for CC, multi-source BFS and batched BC
 * (default tuning), created on first use and freed with the graph. */
inline int graph_queue_set(sg_graph_t g, QueueSet** out) {
  if (!g->aux_queues) {
    auto* qs = new QueueSet();
    const int st = qs->init(g->dev, g->n, g->m, default_tuning());
    if (st != SG_OK) {
      qs->release();
      delete qs;
      return st;
    }
    g->aux_queues = qs;
  }
  *out = static_cast<QueueSet*>(g->aux_queues);
  return SG_OK;
}

template<typename OffT>
inline CompactArgs<OffT> make_compact_args(QueueSet& qs, int dst, const OffT* off, const uint64_t* visited,
                                           uint64_t* prev, uint64_t* frontier, const Tuning& t) {
  CompactArgs<OffT> a{};
  a.off = off;
  a.w0 = qs.w0;
  a.ntiles = qs.ntiles;
  a.visited = visited;
  a.prev = prev;
  a.frontier = frontier;
  a.tiles = qs.tiles;
  a.blocks = qs.blocks;
  for (int k = 0; k < 3; ++k) a.qout[k] = qs.q[dst][k];
  a.hdr = qs.hdr + dst;
  a.tune = t;
  return a;
}

/* Launch the three compaction passes into queue set `dst`.  count_only skips
 * the write pass (totals still land in hdr[dst]).  split_k > 0: the entries
 * cover each list after its first split_k arcs. */
template<typename OffT, int kMode>
inline int compact_to_queues(hipStream_t s, QueueSet& qs, int dst, const OffT* off, const uint64_t* visited,
                             uint64_t* prev, uint64_t* frontier, const Tuning& t, bool count_only = false,
                             int32_t split_k = 0) {
  CompactArgs<OffT> a{};
  a.off = off;
  a.w0 = qs.w0;
  a.ntiles = qs.ntiles;
  a.visited = visited;
  a.prev = prev;
  a.frontier = frontier;
  a.tiles = qs.tiles;
  a.blocks = qs.blocks;
  for (int k = 0; k < 3; ++k) a.qout[k] = qs.q[dst][k];
  a.hdr = qs.hdr + dst;
  a.tune = t;
  a.split_k = split_k;
  const int phase = split_k > 0 ? 1 : 0;
  compact_count_kernel<OffT, kMode><<<qs.grid_tiles(), kBlock, 0, s>>>(a, phase);
  compact_scan_kernel<<<1, kScanThreads, 0, s>>>(qs.blocks, qs.nblocks, qs.hdr + dst);
  if (!count_only) compact_write_kernel<OffT><<<qs.grid_tiles(), kBlock, 0, s>>>(a, phase);
  SG_HIP(hipGetLastError());
  return SG_OK;
}

/* Only the write pass (counts and scan already done by compact_to_queues
 * with count_only = true on the same frontier). */
template<typename OffT>
inline int compact_write_only(hipStream_t s, QueueSet& qs, int dst, const OffT* off, uint64_t* frontier,
                              const Tuning& t) {
  CompactArgs<OffT> a{};
  a.off = off;
  a.w0 = qs.w0;
  a.ntiles = qs.ntiles;
  a.frontier = frontier;
  a.tiles = qs.tiles;
  a.blocks = qs.blocks;
  for (int k = 0; k < 3; ++k) a.qout[k] = qs.q[dst][k];
  a.hdr = qs.hdr + dst;
  a.tune = t;
  compact_write_kernel<OffT><<<qs.grid_tiles(), kBlock, 0, s>>>(a, 0);
  SG_HIP(hipGetLastError());
  return SG_OK;
}

/* ------------------------------------------------- grid barrier --- */
/* Barrier over all workgroups of a cooperative launch (every workgroup
 * co-resident).  Arrival counters only grow -- barrier t completes at t
 * arrivals per member -- so nothing is reset between barriers; the block is
 * zeroed before each launch.  Workgroups are grouped by blockIdx % 8 (the
 * XCD they are dealt to): a group's last arrival bumps the top counter, whose
 * last arrival publishes generation t.  The wait is bounded: past wait_ticks
 * of the constant-rate clock kBarTimeout is raised and every workgroup
 * leaves.  Returns false when the launch must stop (a flag of stop_mask is
 * set); the caller's stores before the barrier are visible to every
 * workgroup after it (agent-scope release / acquire). */
/* Launch a kernel whose workgroups meet at grid barriers.  The caller sizes
 * the grid from the kernel's occupancy so that every workgroup is resident at
 * once on the device.  The launch is cooperative where the runtime grants it
 * (the runtime then also checks that the grid fits); where it refuses for
 * another reason than the grid's size -- a process whose device another
 * process is already running cooperative launches on gets "the operation
 * cannot be performed in the present state" -- the same grid is launched as
 * an ordinary kernel, and the barriers' bounded wait still ends it should its
 * workgroups not all become resident.  The refusal is remembered for the
 * process. */
inline hipError_t launch_resident(const void* fn, int grid, int block, void** args, hipStream_t s) {
  static std::atomic<int> coop_refused{0};
  if (!coop_refused.load(std::memory_order_relaxed)) {
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, 0, s);
    if (e == hipSuccess || e == hipErrorCooperativeLaunchTooLarge) return e;
    (void)hipGetLastError();
    coop_refused.store(1, std::memory_order_relaxed);
  }
  return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s);
}

constexpr int kBarGroups = 8;
constexpr int kBarStride = 32; /* counters 128 bytes apart */
constexpr uint32_t kBarTimeout = 1u << 31;
constexpr int kBarFlat = 64; /* grids up to this size use one arrival counter */

struct GridBar {
  uint32_t grp[kBarGroups * kBarStride];
  uint32_t top[kBarStride];
  uint32_t gen[kBarStride];
  uint32_t flags;
  uint32_t pad[31];
};

__device__ __forceinline__ void bar_raise(GridBar* b, uint32_t f) { atomic_or(&b->flags, f); }

/* kFences: the caller exchanges data through ordinary loads and stores, so
 * the barrier releases (an L2 write-back of every workgroup's XCD) and
 * acquires (an L2 invalidate) at agent scope.  Without: everything the
 * workgroups exchange is written with agent-coherent stores and read with
 * agent-coherent loads (relaxed atomics at agent scope, which go to the
 * point of coherence every XCD shares), and each wave only waits for its own
 * stores to complete before its workgroup arrives -- no L2 is written back or
 * invalidated, which costs tens of microseconds per barrier once hundreds of
 * workgroups each do it. */
template<bool kFences = true>
__device__ __noinline__ bool grid_barrier(GridBar* c, uint32_t& t, uint64_t wait_ticks, uint32_t stop_mask) {
  __shared__ uint32_t s_stop;
  if (kFences) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  else wait_stores_complete();
  __syncthreads();
  ++t;
  if (threadIdx.x == 0) {
    const uint32_t grid = gridDim.x, g = blockIdx.x % kBarGroups;
    const uint32_t ngroups = grid < (uint32_t)kBarGroups ? grid : (uint32_t)kBarGroups;
    const uint32_t in_group = (grid - g + kBarGroups - 1) / kBarGroups;
    /* up to kBarFlat workgroups arrive on the top counter directly: one
     * atomic round trip fewer, and few enough not to queue long on it */
    const bool flat = grid <= (uint32_t)kBarFlat;
    bool released = false;
    if (flat || atomic_add(&c->grp[g * kBarStride], 1u) == t * in_group - 1) {
      if (kFences) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if (atomic_add(&c->top[0], 1u) == t * (flat ? grid : ngroups) - 1) {
        if (kFences) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
          __hip_atomic_store(&c->gen[0], t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          atomic_store(&c->gen[0], t);
        }
        released = true;
      }
    }
    if (!released) {
      const uint64_t t0 = wall_clock64();
      while ((kFences ? __hip_atomic_load(&c->gen[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                      : atomic_load(&c->gen[0])) < t) {
        if (atomic_load(&c->flags) & (stop_mask | kBarTimeout)) break;
        if ((uint64_t)(wall_clock64() - t0) > wait_ticks) {
          bar_raise(c, kBarTimeout);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    s_stop = (atomic_load(&c->flags) & (stop_mask | kBarTimeout)) ? 1u : 0u;
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();
  if (kFences) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return s_stop == 0;
}

/* A queue entry written with agent-coherent stores / read with
 * agent-coherent loads (entries one workgroup hands to another within a
 * launch whose barriers do not fence). */
__device__ __forceinline__ void store_entry_coherent(QEntry* p, const QEntry& e) {
  uint64_t w[2];
  __builtin_memcpy(w, &e, sizeof(w));
  uint64_t* q = reinterpret_cast<uint64_t*>(p);
  __hip_atomic_store(q, w[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, w[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ QEntry load_entry_coherent(const QEntry* p) {
  uint64_t w[2];
  const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
  w[0] = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  w[1] = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  QEntry e;
  __builtin_memcpy(&e, w, sizeof(w));
  return e;
}

} // namespace sg::trav
