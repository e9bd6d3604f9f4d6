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

#include "policy.hpp"
#include "traverse.hpp"

using namespace sg::trav;

namespace {

constexpr int kPullHead = 16; /* steps of four ids in the longest lane-serial pull head */

template<typename OffT>
struct BfsArgs {
  const OffT* off; /* forward CSR (push, queue entries, degrees) */
  const int32_t* cols;
  const OffT* in_off; /* inverse CSR (pull); aliases off when undirected */
  const int32_t* in_cols;
  const int2* hint; /* per vertex: {in-neighbour of largest out-degree, out-degree}, or with hint_offsets
                       {that in-neighbour, low 32 bits of the vertex's list start}, entry n = {-1, m} */
  int32_t hint_offsets;
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
};

/* Push level: every arc of the class queues; discovery is the test-and-set of
 * the visited bit. */
template<typename OffT>
__device__ __forceinline__ void push_level(const BfsArgs<OffT>& a, int variant, int32_t level) {
  zero_slots(a.stats_zero);
  /* four arcs per lane: four visited-word loads, then the atomic ORs of the
   * unset ones (no result needed: the count pass finds the new vertices as
   * visited & ~prev and writes their depths in vertex order) */
  for_each_queued_arc4<OffT>(
      a.qin, a.qin_hdr, a.cols,
      [&](const uint32_t (&y)[4], const int64_t (&)[4], const bool (&ok)[4]) {
        uint64_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = ok[k] ? a.visited[y[k] >> 6] : ~0ull;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint64_t bit = 1ull << (y[k] & 63);
          if (!(w[k] & bit)) atomic_or(&a.visited[y[k] >> 6], bit);
        }
      },
      variant);
}

/* Pull level.  A wave takes a chunk of up to 64 visited words (fewer when
 * the bitmap has fewer words than 64 per wave, so every wave has a chunk) and
 * compacts the chunk's unvisited vertices in registers (popcount scan + bit
 * select).  Each lane then probes the hints of kPullBatch of them at once --
 * their hint entries loaded together, then their bitmap probes -- so a lane
 * keeps kPullBatch dependent chains in flight instead of one.  A hint that
 * hits settles its vertex; the misses go to a per-wave list in LDS and are
 * pulled one per lane from their adjacency lists.  New frontier bits are
 * gathered per word in LDS. */
constexpr int kPullBatch = 8;

template<typename OffT>
__device__ __forceinline__ void pull_level(const BfsArgs<OffT>& a, int variant, int32_t level) {
  zero_slots(a.stats_zero);
  /* arcs a lane scans alone before the wave helps: 16 / 4 / 64 */
  const int head_steps = variant == 1 ? 1 : (variant == 2 ? kPullHead : kPullHead / 4);
  __shared__ unsigned long long lds_new[kBlock];
  /* a miss: its vertex and the hint's two words (list start / end, or out-degree) */
  __shared__ uint32_t lds_miss[3][kBlock / 64][64 * kPullBatch];
  uint64_t st_disc = 0, st_deg = 0, st_edges = 0, st_scanned = 0, st_max = 0, st_miss = 0;
  const int lane = lane_id();
  unsigned long long* wnew = lds_new + (threadIdx.x & ~63);
  uint32_t* wmiss = lds_miss[0][threadIdx.x >> 6];
  uint32_t* wm_y = lds_miss[1][threadIdx.x >> 6];
  uint32_t* wm_e = lds_miss[2][threadIdx.x >> 6];
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const int tail_bits = (int)(a.n & 63);
  /* words per chunk: 64, halved while the chunks would leave waves idle (at least 8) */
  int chunk = 64;
  while (chunk > 8 && (a.w0 + chunk - 1) / chunk < nwaves) chunk >>= 1;
  const uint64_t below = (1ull << lane) - 1ull;
  for (int64_t base = gwave * chunk; base < a.w0; base += nwaves * chunk) {
    const int64_t myw = base + lane;
    const bool mine = lane < chunk && myw < a.w0;
    const uint64_t myvis = mine ? a.visited[myw] : ~0ull;
    uint64_t myun = ~myvis;
    if (myw == a.w0 - 1 && tail_bits) myun &= (1ull << tail_bits) - 1ull;
    const uint32_t cnt = (uint32_t)__popcll(myun);
    const uint32_t incl = wave_inclusive_scan(cnt);
    const uint32_t total = shfl(incl, 63);
    wnew[lane] = 0ull;
    for (uint32_t t0 = 0; t0 < total; t0 += 64 * kPullBatch) {
      uint32_t vv[kPullBatch];
      bool act[kPullBatch];
#pragma unroll
      for (int k = 0; k < kPullBatch; ++k) {
        const uint32_t t = t0 + 64 * k + lane;
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
        act[k] = t < total;
        vv[k] = (uint32_t)((base + owner) * 64 + (act[k] ? select_bit(un, t - before) : 0u));
      }
      /* first probes: the in-neighbour of largest out-degree, read with the
       * vertex's out-degree (or its list start) from one 8-byte entry */
      int2 hd[kPullBatch];
      uint32_t rend[kPullBatch];
#pragma unroll
      for (int k = 0; k < kPullBatch; ++k) {
        hd[k] = act[k] ? a.hint[vv[k]] : int2{0, 0};
        rend[k] = (act[k] && a.hint_offsets) ? (uint32_t)a.hint[vv[k] + 1].y : 0u;
      }
      uint32_t nmiss = 0;
#pragma unroll
      for (int k = 0; k < kPullBatch; ++k) {
        const bool hit = act[k] && test_bit(a.cur_bm, (uint32_t)hd[k].x);
        if (hit) {
          const int32_t odeg = a.hint_offsets ? (int32_t)(rend[k] - (uint32_t)hd[k].y) : hd[k].y;
          a.depth[vv[k]] = level + 1;
          atomicOr(&wnew[(int64_t)(vv[k] >> 6) - base], 1ull << (vv[k] & 63));
          ++st_scanned; /* active vertices have in-arcs: in-degree 0 is pre-marked visited */
          ++st_edges;   /* the hint probe */
          ++st_disc;
          st_deg += odeg;
          st_max = odeg > (int64_t)st_max ? odeg : st_max;
        }
        const uint64_t mm = ballot(act[k] && !hit);
        if (act[k] && !hit) {
          const uint32_t slot = nmiss + (uint32_t)__popcll(mm & below);
          wmiss[slot] = vv[k];
          wm_y[slot] = (uint32_t)hd[k].y;
          wm_e[slot] = rend[k];
        }
        nmiss += (uint32_t)__popcll(mm);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
      /* the misses, one per lane: the adjacency list after the hint */
      for (uint32_t m0 = 0; m0 < nmiss; m0 += 64) {
        const bool active = m0 + lane < nmiss;
        const uint32_t v = active ? wmiss[m0 + lane] : 0u;
        const uint32_t hy = active ? wm_y[m0 + lane] : 0u, he = active ? wm_e[m0 + lane] : 0u;
        bool found = false;
        OffT b = 0, e = 0, j = 0;
        int64_t kk = 0;
        int32_t odeg = 0;
        if (active) {
          if (a.hint_offsets) { /* the list's bounds came with the hints */
            b = (OffT)hy;
            e = (OffT)he;
            odeg = (int32_t)(he - hy);
          } else {
            odeg = (int32_t)hy;
            b = a.in_off[v];
            e = a.in_off[v + 1];
          }
          j = b;
          kk = (int64_t)(e - b);
          /* a single in-arc is the hint itself, which missed: no parent and
           * no list to read */
          if (e - b == 1) {
            j = e;
            kk = 0;
          }
          /* lane-serial head: four independent id loads then four
           * independent bitmap probes per step */
          for (int step = 0; step < head_steps && j < e; ++step, j += 4) {
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
          ++st_scanned;
          st_edges += 1 + kk; /* hint probe + list probes after the miss */
          ++st_miss;
          if (found) {
            a.depth[v] = level + 1;
            atomicOr(&wnew[(int64_t)(v >> 6) - base], 1ull << (v & 63));
            ++st_disc;
            st_deg += odeg;
            st_max = odeg > (int64_t)st_max ? odeg : st_max;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    const uint64_t mynew = wnew[lane];
    if (mine) {
      a.next_bm[myw] = mynew;
      if (mynew) a.visited[myw] = myvis | mynew;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  }
  flush_block_stats(a.stats, st_disc, st_deg, st_edges, st_scanned, st_max, st_miss);
}

/* Device-resident control block of the level loop.  Every kernel of a step
 * reads it (kernel boundaries, and the last-workgroup hand-offs inside the
 * count and write passes, order the accesses); count_dyn_kernel (finish_level)
 * and the bursts write it. */
struct BfsCtl {
  int32_t level;    /* the level about to run */
  int32_t dir;      /* its direction: 0 push, 1 pull */
  int32_t done;     /* a level discovered nothing: later steps are no-ops */
  int32_t write;    /* queues for the next (push) level must be written */
  int32_t lb;       /* work-mapping variant of the level about to run */
  int32_t burst;    /* the next level is small enough for a burst */
  int32_t skip;     /* a burst ran this level: the next step only counts and decides it */
  int32_t n_push, n_pull;
  int64_t f_vertices, f_edges; /* frontier entering the level */
  unsigned long long visited_deg, reached;
};

/* One level's record, written into pinned host memory (levels below
 * SG_MAX_LEVELS_RECORDED) by finish_level or, inside a burst, by
 * burst_body. */
struct alignas(64) LevelLog {
  int32_t level, dir, done, next_dir, action, burst;
  int64_t f_vertices, f_edges, disc, disc_deg, edges, scanned, misses, max_deg, unexplored;
};

/* One step's record (a step = the launches the host enqueues together: one
 * level, plus the levels of a burst), written into a ring in pinned host
 * memory by the step's last working kernel; the host reads it after the
 * step's completion event. */
struct alignas(64) StepLog {
  int32_t step, last_level, done, n_push, n_pull;
  int32_t next_dir; /* direction of the level after last_level */
  int32_t fail;     /* a grid barrier timed out */
  int64_t reached_v, reached_e;
  int64_t next_f_edges; /* its frontier's arcs */
};

constexpr int kRing = 8; /* steps in the log / event rings (> lookahead + 1) */

/* A step's record, published once, by the step's last writer: every field,
 * then (after a system-scope fence) the step number the host polls for, so
 * a host that sees the number reads a whole record. */
__device__ __forceinline__ void publish_step(StepLog* slot, StepLog sl) {
  const int32_t stp = sl.step;
  sl.step = -1;
  *slot = sl;
  __threadfence_system();
  __hip_atomic_store(&slot->step, stp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __threadfence_system();
}

struct DecideParams {
  int32_t policy;
  float alpha, beta;
  int64_t n;
  unsigned long long total_degree;
  int32_t lb_fixed;     /* variant under the push / pull / hybrid policies */
  const float* mlp;     /* learned policy parameters (SG_POLICY_LEARNED) */
  float f6, f7;         /* the graph's constant features (average degree, vertex count) */
  int64_t burst_arcs;   /* frontier arcs up to which push levels run in a burst (0: never) */
  unsigned long long learned_gate; /* learned policy: a level whose frontier has fewer arcs than the
                                      graph's arcs / learned_gate pushes (variant 0) without asking the
                                      network; 0: the network decides every level */
};

/* Default of sg_bfs_options.learned_gate.  A frontier with fewer arcs than
 * 1/1024 of the graph's arcs is cheaper to push than any pull, which scans
 * every word of the visited bitmap and the lists of the vertices not yet
 * reached (the reference's rule too pushes such frontiers until the arcs left
 * unexplored fall below 15 times the frontier's); on long traversals such
 * levels are nearly all the levels, and inferring for them costs ~1.3 us
 * each, 5-10 % of a grid level. */
constexpr int64_t kLearnedGateDefault = 1024;

/* evaluateHeuristic (algorithms/bfs.hpp:341-350): push -> pull when the new
 * frontier's arcs exceed the unexplored arcs / alpha (both cast to uint32 and
 * compared as float); pull -> push when the frontier is smaller than n / beta. */
__device__ __forceinline__ int32_t next_direction(const DecideParams& p, int32_t dir, unsigned long long disc,
                                                  unsigned long long disc_deg, unsigned long long unexplored) {
  if (p.policy != SG_POLICY_HYBRID) return dir;
  if (dir == 0) {
    const uint32_t fdeg = (uint32_t)disc_deg, udeg = (uint32_t)unexplored;
    return ((float)fdeg > ((float)udeg / p.alpha)) ? 1 : 0;
  }
  return ((float)disc < ((float)p.n / p.beta)) ? 0 : 1;
}

/* Action (direction * 3 + variant) of the next level; every lane of one
 * wavefront calls with the same arguments.  Under the learned policy `pol`
 * holds the parameters, staged by the workgroup (stage_policy). */
__device__ __forceinline__ int32_t decide_action(const DecideParams& p, const sg::policy::Staged* pol, int32_t dir,
                                                 unsigned long long disc, unsigned long long disc_deg,
                                                 unsigned long long unexplored, int32_t next_level,
                                                 unsigned long long max_deg) {
  if (p.policy == SG_POLICY_LEARNED) {
    if (p.learned_gate > 0 && disc_deg * p.learned_gate < p.total_degree) return 0;
    return sg::policy::infer_staged(pol, (double)disc, (double)disc_deg, (double)unexplored, next_level,
                                    (double)max_deg, p.f6, p.f7);
  }
  return next_direction(p, dir, disc, disc_deg, unexplored) * 3 + p.lb_fixed;
}

/* All threads of the workgroup copy the learned parameters into LDS; the
 * caller's next workgroup barrier publishes them. */
__device__ __forceinline__ void stage_policy(const DecideParams& p, sg::policy::Staged* pol) {
  if (p.policy == SG_POLICY_LEARNED) sg::policy::stage(pol, p.mlp, (int)threadIdx.x, (int)blockDim.x);
}

/* A traversal's start in one launch (each kernel boundary costs a few
 * microseconds, a fifth of a small level): depth = n + 1 except depth[src] =
 * 0, visited = prev = isolated | {src}, cur_bm = {src}, the stat slots
 * cleared, and (thread 0) the F0 queue = {src} and the control block for
 * level 0. */
template<typename OffT>
__global__ void bfs_start_kernel(BfsArgs<OffT> a, uint64_t* prev, const uint64_t* isolated,
                                 unsigned long long* stat_words, int64_t n_stat_words, QEntry* q0, QEntry* q1,
                                 QEntry* q2, QueueHdr* hdr, Tuning t, int32_t src, BfsCtl* ctl, int32_t first_dir,
                                 int32_t first_lb, int64_t burst_arcs) {
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t sw = src >> 6;
  const uint64_t bit = 1ull << (src & 63);
  for (int64_t i = tid; i < a.n; i += stride) a.depth[i] = i == src ? 0 : (int32_t)(a.n + 1);
  for (int64_t i = tid; i < a.w0; i += stride) {
    const uint64_t x = isolated[i] | (i == sw ? bit : 0ull);
    a.visited[i] = x;
    prev[i] = x;
    a.cur_bm[i] = i == sw ? bit : 0ull;
  }
  for (int64_t i = tid; i < n_stat_words; i += stride) stat_words[i] = 0ull;
  if (tid != 0) return;
  QEntry* const q[3] = {q0, q1, q2};
  enqueue_single(a.off, t, q, hdr, src);
  const unsigned long long d = (unsigned long long)(a.off[src + 1] - a.off[src]);
  BfsCtl c{};
  c.dir = first_dir;
  c.lb = first_lb;
  c.f_vertices = 1;
  c.f_edges = (int64_t)d;
  c.visited_deg = d; /* the source's arcs count as explored */
  c.reached = 1;
  /* a pushing start with a small source list begins in a burst (launched
   * ahead of the first step) */
  c.burst = (first_dir == 0 && burst_arcs > 0 && (int64_t)d <= burst_arcs) ? 1 : 0;
  *ctl = c;
}

/* The traversal of one level, push or pull as the control block says.  Four
 * waves per SIMD, what four workgroups of four waves per CU place: the pull's
 * eight hint chains per lane take 106 registers (four chains: 80 registers,
 * 3 % slower; sixteen: 162 registers and three workgroups per CU, 11 %
 * slower). */
constexpr int kLevelWavesPerSimd = 4;
template<typename OffT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kLevelWavesPerSimd))) level_kernel(BfsArgs<OffT> a, const BfsCtl* ctl) {
  if (ctl->done) return;
  if (ctl->skip) { /* the level was traversed by the burst that ended the previous step */
    zero_slots(a.stats_zero);
    return;
  }
  const int32_t level = ctl->level;
  if (ctl->dir == 0) push_level(a, ctl->lb, level);
  else pull_level(a, ctl->lb, level);
}

/* Totals of a pull level's 64 stat slots and the decision taken on them. */
struct PullTotals {
  unsigned long long disc, disc_deg, edges, scanned, misses, max_deg;
  int32_t act;
};

/* Where the last workgroup of a step's count pass writes its results. */
struct FinishArgs {
  TileCount* tiles; /* the count pass's per-workgroup sums, scanned in place */
  int64_t ntiles;
  QueueHdr* hdr_out;
  int32_t step;
  LevelLog* levlog;
  StepLog* steplog;
  uint32_t* arrive; /* zero between launches */
};

/* Scan of the count pass, the push-level decision (wave 0, so the learned
 * policy can run on it), the level record, and the control block for the
 * next level.  Unless a burst takes over the next levels, this ends the
 * step's work and writes the step record.  Run by the last workgroup of the
 * count pass; all kBlock threads call.  `mode` is the count pass's mode (0:
 * nothing counted, the next level pulls or the traversal ends). */
__device__ __forceinline__ void finish_level(const FinishArgs& f, BfsCtl* ctl, const DecideParams& p,
                                             const sg::policy::Staged* s_pol, int32_t mode, const PullTotals* pt) {
  __shared__ QueueHdr tot;
  /* the workgroup sums were written with agent-coherent stores during this launch */
  if (mode != 0) compact_scan_body_t<kBlock, 8, true>(f.tiles, f.ntiles, f.hdr_out, &tot);
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int32_t dir = ctl->dir, level = ctl->level;
  unsigned long long disc, disc_deg, edges, scanned, misses, max_deg;
  int32_t act;
  const unsigned long long vdeg_before = ctl->visited_deg;
  if (dir == 0) {
    disc = tot.nvert;
    disc_deg = tot.deg;
    max_deg = tot.maxdeg;
    edges = (unsigned long long)ctl->f_edges;
    scanned = (unsigned long long)ctl->f_vertices;
    misses = 0;
    act = decide_action(p, s_pol, 0, disc, disc_deg, p.total_degree - (vdeg_before + disc_deg), level + 1,
                        max_deg);
  } else {
    disc = pt->disc;
    disc_deg = pt->disc_deg;
    max_deg = pt->max_deg;
    edges = pt->edges;
    scanned = pt->scanned;
    misses = pt->misses;
    act = pt->act;
  }
  if (threadIdx.x != 0) return;
  const int32_t next = act / 3;
  const unsigned long long vdeg = vdeg_before + disc_deg;
  const int32_t done = disc == 0 ? 1 : 0;
  if (level < SG_MAX_LEVELS_RECORDED) {
    LevelLog e;
    e.level = level;
    e.dir = dir;
    e.done = done;
    e.next_dir = next;
    e.action = dir * 3 + ctl->lb;
    e.burst = ctl->skip; /* traversed by the burst that ended the previous step */
    e.f_vertices = ctl->f_vertices;
    e.f_edges = ctl->f_edges;
    e.disc = (int64_t)disc;
    e.disc_deg = (int64_t)disc_deg;
    e.edges = (int64_t)edges;
    e.scanned = (int64_t)scanned;
    e.misses = (int64_t)misses;
    e.max_deg = (int64_t)max_deg;
    e.unexplored = (int64_t)(p.total_degree - vdeg);
    f.levlog[level] = e;
  }
  const int32_t n_push = ctl->n_push + (dir == 0 ? 1 : 0), n_pull = ctl->n_pull + (dir == 0 ? 0 : 1);
  const unsigned long long reached = ctl->reached + disc;
  const int32_t burst = (!done && next == 0 && p.burst_arcs > 0 && (int64_t)disc_deg <= p.burst_arcs) ? 1 : 0;
  ctl->visited_deg = vdeg;
  ctl->reached = reached;
  ctl->f_vertices = (int64_t)disc;
  ctl->f_edges = (int64_t)disc_deg;
  ctl->write = (!done && next == 0) ? 1 : 0;
  ctl->done = done;
  ctl->dir = next;
  ctl->lb = act % 3;
  ctl->level = level + 1;
  ctl->n_push = n_push;
  ctl->n_pull = n_pull;
  ctl->burst = burst;
  ctl->skip = 0;
  if (!burst) {
    StepLog sl{};
    sl.step = f.step;
    sl.last_level = level;
    sl.done = done;
    sl.n_push = n_push;
    sl.n_pull = n_pull;
    sl.next_dir = next;
    sl.reached_v = (int64_t)reached;
    sl.reached_e = (int64_t)vdeg;
    sl.next_f_edges = (int64_t)disc_deg;
    publish_step(&f.steplog[f.step % kRing], sl);
  }
  __threadfence_system();
}

/* Count pass of the level.  After a push it counts visited & ~prev.  After a
 * pull the next action must be known first (queues are only built when the
 * next level pushes): wave 0 of every workgroup reduces the pull's 64 stat
 * slots and decides -- the same inputs, so the same answer everywhere, with no
 * extra launch.  The last workgroup to finish then scans, decides (after a
 * push) and updates the control block (finish_level). */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) count_dyn_kernel(CompactArgs<OffT> a, BfsCtl* ctl, const StatSlot* slots,
                                                           DecideParams p, FinishArgs f) {
  __shared__ int32_t s_mode;
  __shared__ sg::policy::Staged s_pol;
  __shared__ PullTotals s_pt;
  if (ctl->done) return;
  const int32_t dir = ctl->dir, level = ctl->level;
  if (dir == 0) {
    compact_count_body<OffT, kFromDiff>(a, 0, level + 1); /* the push left the depths to this pass */
    if (!last_to_arrive<true>(f.arrive, false)) return;
    stage_policy(p, &s_pol);
    __syncthreads();
    finish_level(f, ctl, p, &s_pol, 1, nullptr);
    return;
  }
  stage_policy(p, &s_pol);
  __syncthreads();
  if (threadIdx.x < 64) {
    const StatSlot& x = slots[threadIdx.x];
    const unsigned long long disc = wave_sum_u64(x.disc), disc_deg = wave_sum_u64(x.disc_deg);
    const unsigned long long max_deg = wave_max_u64(x.max_deg);
    const unsigned long long unexplored = p.total_degree - (ctl->visited_deg + disc_deg);
    const int32_t act = decide_action(p, &s_pol, 1, disc, disc_deg, unexplored, level + 1, max_deg);
    const int32_t mode = (disc > 0 && act / 3 == 0) ? 2 : 0;
    const unsigned long long edges = wave_sum_u64(x.edges), scanned = wave_sum_u64(x.scanned);
    const unsigned long long misses = wave_sum_u64(x.misses);
    if (threadIdx.x == 0) {
      s_mode = mode;
      s_pt = PullTotals{disc, disc_deg, edges, scanned, misses, max_deg, act};
    }
  }
  __syncthreads();
  const int32_t mode = s_mode;
  if (mode == 2) compact_count_body<OffT, kFromBitmapSyncPrev>(a);
  if (!last_to_arrive<true>(f.arrive, false)) return;
  finish_level(f, ctl, p, &s_pol, mode, &s_pt);
}

/* ------------------------------------------------------------ bursts --- */
/* Small push levels are bound by launch and dispatch cost, not bandwidth: a
 * level with a few thousand frontier arcs takes three grid-wide launches
 * (about 25 us) for microseconds of memory work.  A burst, one 512-thread
 * workgroup (the last of the write pass, or burst_kernel before the first
 * step), runs such levels back to back with
 * workgroup barriers between them, for as long as the policy keeps pushing and
 * the frontier keeps at most burst_arcs arcs.  The frontier lives in two
 * lists of (first arc, vertex, degree) entries; each wave flattens 64 entries
 * at a time into arcs (merge path) and claims discoveries four arcs per lane
 * per round, so four claims are in flight per lane.
 *
 * Hand-back.  A burst that ends because a level discovered nothing ends the
 * traversal.  Otherwise the level at which it stops (its frontier grew past
 * burst_arcs, or the policy chose pull next) has been traversed but is
 * treated as a push level that level_kernel ran: its discoveries are in
 * visited and depth but not in prev, the control block describes it, and the
 * next step skips the traversal (`skip`) and only counts (visited & ~prev:
 * the exact next frontier and its queues), scans and decides it.  Earlier
 * burst levels set the prev bits of their discoveries when the following
 * level starts, so the count sees that level's discoveries only. */
constexpr int kBurstThreads = 512;
constexpr int kBurstWaves = kBurstThreads / 64;
/* 2,048: with grid bursts spreading their waves over the compute units, a
 * grid level of 2-4 K arcs is faster there than in one workgroup (4096^2
 * grid 90.6 -> 87.0 ms; RMAT scale 24 within 0.5 % over 1,024-8,192) */
constexpr int kBurstDefaultArcs = 2048;
constexpr int kBurstMaxArcs = 65536; /* capacity of each burst list (entries) */

template<typename OffT>
struct BurstArgs {
  const OffT* off;
  const int32_t* cols;
  uint64_t* visited;
  uint64_t* prev;
  int32_t* depth;
  QEntry* q[3];    /* queues of the next step (Q[nxt]) */
  QueueHdr* hdr;
  QEntry* list[2][2]; /* burst frontier lists [parity][small / big], kBurstMaxArcs entries each */
  Tuning tune;
};

struct BurstAcc {
  unsigned long long disc, deg, mx, nsmall, nbig;
};

constexpr int kBurstBig = 64;     /* lists this long go to the big list, walked in pieces by every wave */
constexpr int kBurstPiece = 256;  /* arcs per piece of a big list: four per lane */

/* Claim the (up to) four arcs y[k] of each lane: atomic OR on the visited
 * word issued together with the two offset loads a discovery needs, one LDS
 * counter update per wave and list for all four, then the stores.  A
 * discovery gets depth level + 1 and goes to the small or the big list of the
 * next level by its degree. */
template<typename OffT>
__device__ __forceinline__ void burst_claim4(const BurstArgs<OffT>& a, const uint32_t (&y)[4], const bool (&ok)[4],
                                             int32_t level, QEntry* out_small, QEntry* out_big, uint32_t* cnt2,
                                             BurstAcc& acc) {
  const int lane = lane_id();
  uint64_t old[4];
  int64_t vs[4], ve[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    old[k] = ~0ull;
    vs[k] = ve[k] = 0;
    if (ok[k]) {
      old[k] = atomic_or(&a.visited[y[k] >> 6], 1ull << (y[k] & 63));
      vs[k] = (int64_t)a.off[y[k]];
      ve[k] = (int64_t)a.off[y[k] + 1];
    }
  }
  bool nw[4], big[4];
  uint64_t ms[4], mb[4];
  uint32_t cs = 0, cb = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nw[k] = ((old[k] >> (y[k] & 63)) & 1ull) == 0ull;
    big[k] = ve[k] - vs[k] >= kBurstBig;
    ms[k] = ballot(nw[k] && !big[k]);
    mb[k] = ballot(nw[k] && big[k]);
    cs += (uint32_t)__popcll(ms[k]);
    cb += (uint32_t)__popcll(mb[k]);
  }
  if ((cs | cb) == 0) return;
  uint32_t ps = 0, pb = 0;
  if (lane == 0) {
    if (cs) ps = atomicAdd(&cnt2[0], cs);
    if (cb) pb = atomicAdd(&cnt2[1], cb);
  }
  ps = __shfl(ps, 0, 64);
  pb = __shfl(pb, 0, 64);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (nw[k]) {
      const uint32_t v = y[k];
      a.depth[v] = level + 1;
      QEntry e;
      e.start = vs[k];
      e.v = (int32_t)v;
      e.len = (int32_t)(ve[k] - vs[k]);
      if (big[k]) out_big[pb + lanes_below(mb[k])] = e;
      else out_small[ps + lanes_below(ms[k])] = e;
      ++acc.disc;
      acc.deg += (unsigned long long)e.len;
      acc.mx = (unsigned long long)e.len > acc.mx ? (unsigned long long)e.len : acc.mx;
    }
    ps += (uint32_t)__popcll(ms[k]);
    pb += (uint32_t)__popcll(mb[k]);
  }
  acc.nsmall += lane == 0 ? cs : 0;
  acc.nbig += lane == 0 ? cb : 0;
}

/* Short lists: a wave lays 64 entries end to end (merge path) and takes four
 * arcs per lane per round; waves take groups of 64 entries in turn. */
template<typename OffT>
__device__ __forceinline__ void burst_visit_small(const BurstArgs<OffT>& a, const QEntry* ents, uint32_t cnt,
                                                  int32_t level, QEntry* out_small, QEntry* out_big, uint32_t* cnt2,
                                                  BurstAcc& acc) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (uint32_t base = (uint32_t)wv * 64; base < cnt; base += kBurstThreads) {
    QEntry qe;
    qe.start = 0;
    qe.v = 0;
    qe.len = 0;
    if (base + lane < cnt) qe = ents[base + lane];
    const uint32_t incl = wave_inclusive_scan((uint32_t)qe.len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - (uint32_t)qe.len;
    for (uint32_t t0 = 0; t0 < total; t0 += 4 * 64) {
      uint32_t y[4];
      bool ok[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t t = t0 + 64 * k + lane;
        const int owner = scan_owner(incl, t);
        const uint32_t before = __shfl(excl, owner, 64);
        const int64_t st = shfl_i64(qe.start, owner);
        ok[k] = t < total;
        y[k] = ok[k] ? (uint32_t)a.cols[st + (t - before)] : 0u;
      }
      burst_claim4(a, y, ok, level, out_small, out_big, cnt2, acc);
    }
  }
}

/* Long lists: every list is cut into pieces of kBurstPiece arcs, and the
 * pieces of the level go to the waves in turn, so one hub keeps every wave
 * busy. */
template<typename OffT>
__device__ __forceinline__ void burst_visit_big(const BurstArgs<OffT>& a, const QEntry* ents, uint32_t cnt,
                                                int32_t level, QEntry* out_small, QEntry* out_big, uint32_t* cnt2,
                                                BurstAcc& acc) {
  const int lane = lane_id();
  const uint32_t wv = threadIdx.x >> 6;
  uint32_t pbase = 0; /* index of the entry's first piece among the level's pieces */
  for (uint32_t i = 0; i < cnt; ++i) {
    const QEntry qe = ents[i];
    const uint32_t np = (uint32_t)((qe.len + kBurstPiece - 1) / kBurstPiece);
    const uint32_t first = pbase + (wv + kBurstWaves - pbase % kBurstWaves) % kBurstWaves;
    for (uint32_t pc = first; pc < pbase + np; pc += kBurstWaves) {
      const int64_t o0 = (int64_t)(pc - pbase) * kBurstPiece;
      uint32_t y[4];
      bool ok[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t o = o0 + 64 * k + lane;
        ok[k] = o < qe.len;
        y[k] = ok[k] ? (uint32_t)a.cols[qe.start + o] : 0u;
      }
      burst_claim4(a, y, ok, level, out_small, out_big, cnt2, acc);
    }
    pbase += np;
  }
}

/* One barrier per level: the counter and the partial sums are double
 * buffered by level parity, so a wave may start the next level while others
 * still read this level's sums; every wave evaluates the (deterministic)
 * decision itself. */
template<typename OffT>
__device__ __forceinline__ void burst_body(const BurstArgs<OffT>& a, BfsCtl* ctl, const DecideParams& p,
                                           int32_t step, LevelLog* levlog, StepLog* steplog, bool publish) {
  __shared__ sg::policy::Staged s_pol;
  __shared__ uint32_t s_count[2][2];
  __shared__ unsigned long long s_red[2][kBurstWaves][5];
  stage_policy(p, &s_pol);
  if (threadIdx.x == 0) s_count[0][0] = s_count[0][1] = 0;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  int32_t level = ctl->level, lb = ctl->lb;
  int64_t f_vertices = ctl->f_vertices, f_edges = ctl->f_edges;
  unsigned long long vdeg = ctl->visited_deg, reached = ctl->reached;
  int32_t n_push = ctl->n_push, done = 0, act = 0;
  const QueueHdr h0 = *a.hdr;
  int cur = 0;
  uint32_t n_small = 0, n_big = 0; /* entries of the lists of the level about to run (after the first) */
  unsigned long long disc = 0, disc_deg = 0, max_deg = 0;
  __syncthreads();
  for (int it = 0;; ++it) {
    const int par = it & 1;
    BurstAcc acc{0, 0, 0, 0, 0};
    QEntry* out_small = a.list[cur ^ 1][0];
    QEntry* out_big = a.list[cur ^ 1][1];
    if (it == 0) { /* the queues: class 0 lists are short, classes 1 and 2 long */
      burst_visit_small(a, a.q[0], h0.count[0], level, out_small, out_big, s_count[par], acc);
      burst_visit_big(a, a.q[1], h0.count[1], level, out_small, out_big, s_count[par], acc);
      burst_visit_big(a, a.q[2], h0.count[2], level, out_small, out_big, s_count[par], acc);
    } else {
      /* the previous burst level's discoveries join prev (fire and forget:
       * nothing in the burst reads prev) */
      for (uint32_t i = threadIdx.x; i < n_small + n_big; i += kBurstThreads) {
        const uint32_t v = (uint32_t)(i < n_small ? a.list[cur][0][i] : a.list[cur][1][i - n_small]).v;
        atomic_or(&a.prev[v >> 6], 1ull << (v & 63));
      }
      burst_visit_small(a, a.list[cur][0], n_small, level, out_small, out_big, s_count[par], acc);
      burst_visit_big(a, a.list[cur][1], n_big, level, out_small, out_big, s_count[par], acc);
    }
    acc.disc = wave_sum_u64(acc.disc);
    acc.deg = wave_sum_u64(acc.deg);
    acc.mx = wave_max_u64(acc.mx);
    acc.nsmall = wave_sum_u64(acc.nsmall);
    acc.nbig = wave_sum_u64(acc.nbig);
    if (lane == 0) {
      s_red[par][wv][0] = acc.disc;
      s_red[par][wv][1] = acc.deg;
      s_red[par][wv][2] = acc.mx;
      s_red[par][wv][3] = acc.nsmall;
      s_red[par][wv][4] = acc.nbig;
    }
    if (threadIdx.x == 0) s_count[par ^ 1][0] = s_count[par ^ 1][1] = 0; /* last used two levels ago */
    __syncthreads(); /* discoveries, the lists and the partial sums are complete */
    disc = 0;
    disc_deg = 0;
    max_deg = 0;
    n_small = 0;
    n_big = 0;
    for (int w = 0; w < kBurstWaves; ++w) {
      disc += s_red[par][w][0];
      disc_deg += s_red[par][w][1];
      max_deg = s_red[par][w][2] > max_deg ? s_red[par][w][2] : max_deg;
      n_small += (uint32_t)s_red[par][w][3];
      n_big += (uint32_t)s_red[par][w][4];
    }
    const unsigned long long unexplored = p.total_degree - (vdeg + disc_deg);
    act = decide_action(p, &s_pol, 0, disc, disc_deg, unexplored, level + 1, max_deg);
    done = disc == 0 ? 1 : 0;
    const bool more = !done && act / 3 == 0 && (int64_t)disc_deg <= p.burst_arcs;
    if (!more && !done) break; /* the next step counts and decides this level (hand-back) */
    if (threadIdx.x == 0 && level < SG_MAX_LEVELS_RECORDED) {
      LevelLog e;
      e.level = level;
      e.dir = 0;
      e.done = done;
      e.next_dir = act / 3;
      e.action = lb;
      e.burst = 1;
      e.f_vertices = f_vertices;
      e.f_edges = f_edges;
      e.disc = (int64_t)disc;
      e.disc_deg = (int64_t)disc_deg;
      e.edges = f_edges;
      e.scanned = f_vertices;
      e.misses = 0;
      e.max_deg = (int64_t)max_deg;
      e.unexplored = (int64_t)unexplored;
      levlog[level] = e;
    }
    vdeg += disc_deg;
    reached += disc;
    ++n_push;
    f_vertices = (int64_t)disc;
    f_edges = (int64_t)disc_deg;
    lb = act % 3;
    ++level;
    cur ^= 1;
    if (!more) break; /* done */
  }
  if (threadIdx.x == 0) {
    /* done: the traversal ended inside the burst.  Otherwise `level` is the
     * level the burst traversed last, described as a push level that the
     * next step only counts and decides. */
    ctl->level = level;
    ctl->dir = 0;
    ctl->lb = lb;
    ctl->done = done;
    ctl->write = 0;
    ctl->burst = 0;
    ctl->skip = done ? 0 : 1;
    ctl->n_push = n_push;
    ctl->visited_deg = vdeg;
    ctl->reached = reached;
    ctl->f_vertices = f_vertices;
    ctl->f_edges = f_edges;
    StepLog sl{};
    sl.step = step;
    sl.last_level = level - 1;
    sl.done = done;
    sl.n_push = n_push;
    sl.n_pull = ctl->n_pull;
    sl.reached_v = (int64_t)reached;
    sl.reached_e = (int64_t)vdeg;
    sl.next_f_edges = f_edges; /* the level handed back, a push level */
    /* the step's record, unless the step's own count pass writes it later
     * (a burst ahead of the step that hands a level back) */
    if (publish || done) publish_step(&steplog[step % kRing], sl);
    __threadfence_system();
  }
}

/* A burst from the source, before the first step (all kBurstThreads threads
 * of one workgroup); returns at once unless the control block asks for one. */
template<typename OffT>
__global__ void __launch_bounds__(kBurstThreads) burst_kernel(BurstArgs<OffT> a, BfsCtl* ctl, DecideParams p,
                                                              int32_t step, LevelLog* levlog, StepLog* steplog) {
  if (ctl->done || !ctl->burst) return;
  burst_body(a, ctl, p, step, levlog, steplog, false);
}

/* Queue write of a level that the next one pushes.  A workgroup of
 * kBurstThreads threads writes the tiles of two count-pass workgroups; when
 * the control block asks for a burst, the last workgroup to finish runs it on
 * the queues the grid has just written. */
template<typename OffT>
__global__ void __launch_bounds__(kBurstThreads) write_dyn_kernel(CompactArgs<OffT> a, BurstArgs<OffT> ba,
                                                                  BfsCtl* ctl, DecideParams p, int32_t step,
                                                                  LevelLog* levlog, StepLog* steplog,
                                                                  uint32_t* arrive) {
  static_assert(kBurstThreads == 2 * kBlock, "a write workgroup covers two count-pass workgroups");
  if (!ctl->write) return;
  const bool burst = ctl->burst != 0;
  const bool wrote =
      compact_write_body_at<OffT>(a, 0, (int64_t)blockIdx.x * 2 + (threadIdx.x >> 8), (int)(threadIdx.x & (kBlock - 1)));
  /* a burst level has at most burst_arcs arcs, so few waves write entries and publish them */
  if (!burst || !last_to_arrive<false>(arrive, wrote)) return;
  burst_body(ba, ctl, p, step, levlog, steplog, true);
}

/* ------------------------------------------------------- grid bursts --- */
/* Push levels too large for one workgroup but far too small to fill the chip
 * (a 4096 x 4096 grid has ~8,000 levels of a few thousand arcs each) still
 * pay the ~25 us of a step's three launches.  A grid burst runs such levels
 * in one cooperative launch, with one grid barrier per level instead: every
 * workgroup of the launch is co-resident, the frontier lives in two global
 * lists, and the totals that decide the next level are summed with atomics
 * and read by every workgroup after the barrier, so all of them take the
 * same decision.  It starts from the queues the previous step wrote (like
 * level_kernel), keeps going while the policy pushes and the frontier has at
 * most gb_arcs arcs, and hands the level at which it stops back exactly like
 * a burst: traversed, described by the control block as a push level, its
 * discoveries in visited and depth but not in prev, so that the same step's
 * level_kernel skips it and its count pass counts and decides it.
 *
 * Lists: short adjacency lists (< kBurstBig arcs) as one entry each, laid
 * end to end 64 at a time by a wave (merge path); longer ones as pieces of
 * kGbPiece arcs, one wave per piece, so a hub is spread over many waves. */
constexpr int kGbThreads = 512;
constexpr int kGbWaves = kGbThreads / 64;
constexpr int kGbPiece = 256;
constexpr int64_t kGbDefaultArcs = 1 << 18;
constexpr int64_t kGbMaxArcs = 1 << 24;

/* per level: entries of the next short list and pieces of the next long
 * list (three rotating copies, so the copy a level adds into was cleared two
 * levels earlier and read one level earlier) */
struct alignas(128) GbSums {
  unsigned long long nsmall, npieces;
  unsigned long long pad[14];
};

/* per workgroup and level: its discoveries, their arcs and their largest
 * degree, written by the workgroup and read by every workgroup after the
 * level's barrier (two rotating copies: a workgroup writes a copy only after
 * the barrier that follows every read of its previous contents).  One slot
 * per workgroup, not a shared atomic counter: hundreds of workgroups adding
 * into one line serialize for tens of microseconds per level. */
constexpr int kGbMaxWgs = 1024;
/* Default workgroups of a grid burst: the levels it takes have at most 2^18
 * arcs, sixteen per lane of 32 workgroups of 512; a level's time is its
 * busiest waves' memory round trips plus the barrier, and fewer workgroups
 * make the barrier cheaper (8 to 512 swept on the 4096^2 grid: 8 to 64
 * alike, 6 % slower at 256, 12 % at 512). */
constexpr int kGbDefaultWgs = 32;
struct alignas(32) GbSlot {
  unsigned long long disc, deg, mx, pad;
};

template<typename OffT>
struct GbArgs {
  const OffT* off;
  const int32_t* cols;
  uint64_t* visited;
  uint64_t* prev;
  int32_t* depth;
  QEntry* q[3]; /* the queues of the level about to run */
  const QueueHdr* hdr;
  QEntry* small[2];
  QEntry* pieces[2];
  int64_t cap_small, cap_pieces;
  GbSums* sums;  /* [3] */
  GbSlot* slots; /* [2][kGbMaxWgs] */
  GridBar* bar;
  int64_t gb_arcs;
  uint64_t wait_ticks;
};

/* Claim the (up to) four arcs y[k] of each lane; new vertices get depth
 * level + 1 and join the short list or, cut into pieces, the long list. */
template<typename OffT>
__device__ __forceinline__ void gb_claim4(const GbArgs<OffT>& a, const uint32_t (&y)[4], const bool (&ok)[4],
                                          int32_t level, QEntry* out_small, QEntry* out_pieces, GbSums* sm,
                                          BurstAcc& acc) {
  const int lane = lane_id();
  uint64_t old[4];
  int64_t vs[4], ve[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    old[k] = ~0ull;
    vs[k] = ve[k] = 0;
    if (ok[k]) {
      const uint64_t bit = 1ull << (y[k] & 63);
      const uint64_t w = a.visited[y[k] >> 6];
      if (!(w & bit)) {
        old[k] = atomic_or(&a.visited[y[k] >> 6], bit);
        vs[k] = (int64_t)a.off[y[k]];
        ve[k] = (int64_t)a.off[y[k] + 1];
      }
    }
  }
  bool nw[4];
  uint32_t ns = 0, np = 0, npc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nw[k] = ((old[k] >> (y[k] & 63)) & 1ull) == 0ull;
    const int64_t len = ve[k] - vs[k];
    npc[k] = (nw[k] && len >= kBurstBig) ? (uint32_t)((len + kGbPiece - 1) / kGbPiece) : 0u;
    ns += (nw[k] && len < kBurstBig) ? 1u : 0u;
    np += npc[k];
  }
  const uint32_t is = wave_inclusive_scan(ns), ip = wave_inclusive_scan(np);
  const uint32_t ts = __shfl(is, 63, 64), tp = __shfl(ip, 63, 64);
  if ((ts | tp) == 0) return;
  unsigned long long bs = 0, bp = 0;
  if (lane == 63) {
    if (ts) bs = atomic_add(&sm->nsmall, (unsigned long long)ts);
    if (tp) bp = atomic_add(&sm->npieces, (unsigned long long)tp);
  }
  bs = (unsigned long long)shfl_i64((int64_t)bs, 63) + (is - ns);
  bp = (unsigned long long)shfl_i64((int64_t)bp, 63) + (ip - np);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!nw[k]) continue;
    const uint32_t v = y[k];
    const int64_t len = ve[k] - vs[k];
    a.depth[v] = level + 1;
    ++acc.disc;
    acc.deg += (unsigned long long)len;
    acc.mx = (unsigned long long)len > acc.mx ? (unsigned long long)len : acc.mx;
    if (npc[k] == 0) {
      if ((int64_t)bs < a.cap_small) store_entry_coherent(out_small + bs, QEntry{vs[k], (int32_t)v, (int32_t)len});
      ++bs;
    } else {
      for (uint32_t c = 0; c < npc[k]; ++c, ++bp) {
        const int64_t o = (int64_t)c * kGbPiece;
        const int64_t l = len - o < kGbPiece ? len - o : kGbPiece;
        if ((int64_t)bp < a.cap_pieces) store_entry_coherent(out_pieces + bp, QEntry{vs[k] + o, (int32_t)v, (int32_t)l});
      }
    }
  }
}

/* short lists, 64 entries per wave laid end to end, four arcs per lane per round */
template<typename OffT>
__device__ __forceinline__ void gb_visit_small(const GbArgs<OffT>& a, const QEntry* ents, int64_t cnt, int32_t level,
                                               QEntry* out_small, QEntry* out_pieces, GbSums* sm, BurstAcc& acc,
                                               int64_t gwave, int64_t nwaves, bool coherent) {
  const int lane = lane_id();
  for (int64_t base = gwave * 64; base < cnt; base += nwaves * 64) {
    QEntry qe{0, 0, 0};
    if (base + lane < cnt) {
      qe = coherent ? load_entry_coherent(ents + base + lane) : ents[base + lane];
      /* a list entry of an earlier level of the burst: its vertex joins prev
       * (fire and forget; nothing in the burst reads prev) */
      if (coherent) atomic_or(&a.prev[(uint32_t)qe.v >> 6], 1ull << (qe.v & 63));
    }
    const uint32_t incl = wave_inclusive_scan((uint32_t)qe.len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - (uint32_t)qe.len;
    for (uint32_t t0 = 0; t0 < total; t0 += 4 * 64) {
      uint32_t y[4];
      bool ok[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t t = t0 + 64 * k + lane;
        const int owner = scan_owner(incl, t);
        const uint32_t before = __shfl(excl, owner, 64);
        const int64_t st = shfl_i64(qe.start, owner);
        ok[k] = t < total;
        y[k] = ok[k] ? (uint32_t)a.cols[st + (t - before)] : 0u;
      }
      gb_claim4(a, y, ok, level, out_small, out_pieces, sm, acc);
    }
  }
}

/* long lists: a wave per entry (a piece, or a class 1 / class 2 queue entry),
 * 256 arcs per round */
template<typename OffT>
__device__ __forceinline__ void gb_visit_pieces(const GbArgs<OffT>& a, const QEntry* ents, int64_t cnt, int32_t level,
                                                QEntry* out_small, QEntry* out_pieces, GbSums* sm, BurstAcc& acc,
                                                int64_t gwave, int64_t nwaves, bool coherent) {
  const int lane = lane_id();
  for (int64_t i = gwave; i < cnt; i += nwaves) {
    const QEntry qe = coherent ? load_entry_coherent(ents + i) : ents[i];
    if (coherent && lane == 0 && qe.start == (int64_t)a.off[qe.v]) /* the vertex's first piece */
      atomic_or(&a.prev[(uint32_t)qe.v >> 6], 1ull << (qe.v & 63));
    for (int64_t o0 = 0; o0 < qe.len; o0 += 4 * 64) {
      uint32_t y[4];
      bool ok[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t o = o0 + 64 * k + lane;
        ok[k] = o < qe.len;
        y[k] = ok[k] ? (uint32_t)a.cols[qe.start + o] : 0u;
      }
      gb_claim4(a, y, ok, level, out_small, out_pieces, sm, acc);
    }
  }
}

template<typename OffT>
__global__ void __launch_bounds__(kGbThreads) gburst_kernel(GbArgs<OffT> a, BfsCtl* ctl, DecideParams p, int32_t step,
                                                            LevelLog* levlog, StepLog* steplog) {
  __shared__ sg::policy::Staged s_pol;
  __shared__ unsigned long long s_red[kGbWaves][3];
  __shared__ unsigned long long s_tot[5];
  if (ctl->done || ctl->skip || ctl->dir != 0 || ctl->f_edges > a.gb_arcs) return; /* the same for every workgroup */
  stage_policy(p, &s_pol);
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  /* the waves' ranks run across the workgroups first (rank = wave * grid +
   * workgroup): a level of a few thousand entries needs a few dozen waves,
   * and ranked within a workgroup first they all sat on the first few
   * compute units, queueing their memory requests there */
  const int64_t gwave = (int64_t)(threadIdx.x >> 6) * gridDim.x + blockIdx.x;
  const int64_t nwaves = ((int64_t)gridDim.x * kGbThreads) >> 6;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  int32_t level = ctl->level, lb = ctl->lb;
  int64_t f_vertices = ctl->f_vertices, f_edges = ctl->f_edges;
  unsigned long long vdeg = ctl->visited_deg, reached = ctl->reached;
  int32_t n_push = ctl->n_push, done = 0, act = 0;
  const int32_t n_pull = ctl->n_pull;
  const QueueHdr h0 = *a.hdr;
  uint32_t bt = 0;
  int64_t n_small = 0, n_pieces = 0;
  __syncthreads();
  for (int it = 0;; ++it) {
    const int cur = it & 1;
    GbSums* sm = a.sums + (it % 3);
    if (lead) { /* the copy of the level after next: last read before the previous barrier */
      unsigned long long* z = reinterpret_cast<unsigned long long*>(a.sums + ((it + 1) % 3));
      for (int k = 0; k < 2; ++k) atomic_store(z + k, 0ull);
    }
    BurstAcc acc{0, 0, 0, 0, 0};
    QEntry* out_small = a.small[cur ^ 1];
    QEntry* out_pieces = a.pieces[cur ^ 1];
    if (it == 0) {
      gb_visit_small(a, a.q[0], h0.count[0], level, out_small, out_pieces, sm, acc, gwave, nwaves, false);
      gb_visit_pieces(a, a.q[1], h0.count[1], level, out_small, out_pieces, sm, acc, gwave, nwaves, false);
      gb_visit_pieces(a, a.q[2], h0.count[2], level, out_small, out_pieces, sm, acc, gwave, nwaves, false);
    } else {
      /* the previous level's discoveries join prev as their entries are read */
      gb_visit_small(a, a.small[cur], n_small, level, out_small, out_pieces, sm, acc, gwave, nwaves, true);
      gb_visit_pieces(a, a.pieces[cur], n_pieces, level, out_small, out_pieces, sm, acc, gwave, nwaves, true);
    }
    acc.disc = wave_sum_u64(acc.disc);
    acc.deg = wave_sum_u64(acc.deg);
    acc.mx = wave_max_u64(acc.mx);
    if (lane == 0) {
      s_red[wv][0] = acc.disc;
      s_red[wv][1] = acc.deg;
      s_red[wv][2] = acc.mx;
    }
    __syncthreads();
    GbSlot* slots = a.slots + (int64_t)(it & 1) * kGbMaxWgs;
    if (threadIdx.x == 0) {
      unsigned long long d = 0, g = 0, x = 0;
      for (int w = 0; w < kGbWaves; ++w) {
        d += s_red[w][0];
        g += s_red[w][1];
        x = s_red[w][2] > x ? s_red[w][2] : x;
      }
      atomic_store(&slots[blockIdx.x].disc, d);
      atomic_store(&slots[blockIdx.x].deg, g);
      atomic_store(&slots[blockIdx.x].mx, x);
    }
    /* no fences: the lists, sums and slots are written and read coherently */
    if (!grid_barrier<false>(a.bar, bt, a.wait_ticks, 0u)) break; /* timed out: reported through the bar flags */
    if (threadIdx.x < 64) { /* wave 0 totals the slots of every workgroup and reads the list sizes */
      unsigned long long d = 0, g = 0, x = 0, lists = 0;
      if (threadIdx.x < 2) lists = atomic_load(threadIdx.x == 0 ? &sm->nsmall : &sm->npieces);
      for (uint32_t i = threadIdx.x; i < gridDim.x; i += 64) {
        d += atomic_load(&slots[i].disc);
        g += atomic_load(&slots[i].deg);
        const unsigned long long y = atomic_load(&slots[i].mx);
        x = y > x ? y : x;
      }
      d = wave_sum_u64(d);
      g = wave_sum_u64(g);
      x = wave_max_u64(x);
      if (threadIdx.x == 0) {
        s_tot[0] = d;
        s_tot[1] = g;
        s_tot[2] = x;
        s_tot[3] = lists;
      }
      if (threadIdx.x == 1) s_tot[4] = lists;
    }
    __syncthreads();
    const unsigned long long disc = s_tot[0], disc_deg = s_tot[1], max_deg = s_tot[2];
    const int64_t ns = (int64_t)s_tot[3], np = (int64_t)s_tot[4];
    const unsigned long long unexplored = p.total_degree - (vdeg + disc_deg);
    act = decide_action(p, &s_pol, 0, disc, disc_deg, unexplored, level + 1, max_deg);
    done = disc == 0 ? 1 : 0;
    const bool more = !done && act / 3 == 0 && (int64_t)disc_deg <= a.gb_arcs && ns <= a.cap_small &&
                      np <= a.cap_pieces;
    if (!more && !done) break; /* the count pass of this step counts and decides this level */
    if (lead && level < SG_MAX_LEVELS_RECORDED) {
      LevelLog e;
      e.level = level;
      e.dir = 0;
      e.done = done;
      e.next_dir = act / 3;
      e.action = lb;
      e.burst = 2;
      e.f_vertices = f_vertices;
      e.f_edges = f_edges;
      e.disc = (int64_t)disc;
      e.disc_deg = (int64_t)disc_deg;
      e.edges = f_edges;
      e.scanned = f_vertices;
      e.misses = 0;
      e.max_deg = (int64_t)max_deg;
      e.unexplored = (int64_t)unexplored;
      levlog[level] = e;
    }
    vdeg += disc_deg;
    reached += disc;
    ++n_push;
    f_vertices = (int64_t)disc;
    f_edges = (int64_t)disc_deg;
    lb = act % 3;
    ++level;
    n_small = ns;
    n_pieces = np;
    if (!more) break; /* done */
  }
  if (atomic_load(&a.bar->flags) & kBarTimeout) {
    /* a barrier timed out: stop the traversal and report it in the step record */
    if (threadIdx.x == 0) {
      ctl->done = 1;
      StepLog sl{};
      sl.step = step;
      sl.last_level = level;
      sl.done = 1;
      sl.fail = 1;
      publish_step(&steplog[step % kRing], sl);
    }
    return;
  }
  if (!lead) return;
  ctl->level = level;
  ctl->dir = 0;
  ctl->lb = lb;
  ctl->done = done;
  ctl->write = 0;
  ctl->burst = 0;
  ctl->skip = done ? 0 : 1;
  ctl->n_push = n_push;
  ctl->visited_deg = vdeg;
  ctl->reached = reached;
  ctl->f_vertices = f_vertices;
  ctl->f_edges = f_edges;
  StepLog sl{};
  sl.step = step;
  sl.last_level = level - 1;
  sl.done = done;
  sl.n_push = n_push;
  sl.n_pull = n_pull;
  sl.reached_v = (int64_t)reached;
  sl.reached_e = (int64_t)vdeg;
  sl.next_f_edges = f_edges;
  /* a level handed back is counted and decided by this step's count pass,
   * which writes the step's record */
  if (done) publish_step(&steplog[step % kRing], sl);
  __threadfence_system();
}

__global__ void gb_reset_kernel(GridBar* bar, GbSums* sums) {
  uint32_t* b = reinterpret_cast<uint32_t*>(bar);
  for (int i = threadIdx.x; i < (int)(sizeof(GridBar) / 4); i += blockDim.x) b[i] = 0u;
  unsigned long long* z = reinterpret_cast<unsigned long long*>(sums);
  for (int i = threadIdx.x; i < (int)(3 * sizeof(GbSums) / 8); i += blockDim.x) z[i] = 0ull;
}

/* Per-graph BFS workspace, allocated on first use and reused. */
struct BfsWorkspace {
  int64_t w0 = 0;
  uint64_t* visited = nullptr;
  uint64_t* prev = nullptr;
  uint64_t* bm[2] = {nullptr, nullptr};
  QueueSet qs;
  StatSlot* stats = nullptr; /* 2 areas of kSlots */
  BfsCtl* ctl = nullptr;
  QEntry* burst_list[4] = {nullptr, nullptr, nullptr, nullptr}; /* kBurstMaxArcs entries each */
  uint32_t* arrive = nullptr; /* arrival counters of the count and write passes (kArriveWords each), zero between launches */
  /* grid bursts, allocated on first use for gb_arcs arcs */
  int64_t gb_arcs = 0;
  GridBar* gb_bar = nullptr;
  GbSums* gb_sums = nullptr;
  GbSlot* gb_slots = nullptr;
  QEntry* gb_small[2] = {nullptr, nullptr};
  QEntry* gb_pieces[2] = {nullptr, nullptr};
  LevelLog* levlog = nullptr; /* SG_MAX_LEVELS_RECORDED entries, pinned coherent host memory */
  StepLog* steplog = nullptr; /* kRing entries, pinned coherent host memory */
  /* per step: after its traversal kernel (record_levels only) and at its end */
  hipEvent_t ev_mid[kRing] = {}, ev_done[kRing] = {};
  hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_first = nullptr; /* ev_first: before step 0 */
};

void ws_free(BfsWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->visited);
  (void)hipFree(w->prev);
  (void)hipFree(w->bm[0]);
  (void)hipFree(w->bm[1]);
  w->qs.release();
  (void)hipFree(w->stats);
  (void)hipFree(w->ctl);
  for (QEntry* l : w->burst_list) (void)hipFree(l);
  (void)hipFree(w->arrive);
  (void)hipFree(w->gb_bar);
  (void)hipFree(w->gb_sums);
  (void)hipFree(w->gb_slots);
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(w->gb_small[k]);
    (void)hipFree(w->gb_pieces[k]);
  }
  if (w->levlog) (void)hipHostFree(w->levlog);
  if (w->steplog) (void)hipHostFree(w->steplog);
  for (int r = 0; r < kRing; ++r)
    for (hipEvent_t e : {w->ev_mid[r], w->ev_done[r]})
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {w->ev_begin, w->ev_end, w->ev_first})
    if (e) (void)hipEventDestroy(e);
  delete w;
}

/* list capacities of a grid burst of at most gb_arcs frontier arcs: a level
 * discovers at most its frontier's arcs, and while the discoveries' arcs stay
 * within gb_arcs their pieces number at most gb_arcs / 256 + gb_arcs / 64
 * (a larger count ends the burst, whose lists are then not needed) */
inline int64_t gb_cap_small(int64_t arcs) { return arcs + 64; }
inline int64_t gb_cap_pieces(int64_t arcs) { return arcs / 32 + 64; }

int gb_buffers(sg_graph_t g, BfsWorkspace* w, int64_t arcs) {
  if (w->gb_arcs >= arcs) return SG_OK;
  sg_device_t dev = g->dev;
  SG_HIP(hipStreamSynchronize(dev->stream));
  (void)hipFree(w->gb_bar);
  (void)hipFree(w->gb_sums);
  (void)hipFree(w->gb_slots);
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(w->gb_small[k]);
    (void)hipFree(w->gb_pieces[k]);
    w->gb_small[k] = w->gb_pieces[k] = nullptr;
  }
  w->gb_bar = nullptr;
  w->gb_sums = nullptr;
  w->gb_slots = nullptr;
  w->gb_arcs = 0;
  SG_TRY(sg::alloc_device(dev, (void**)&w->gb_bar, sizeof(GridBar)));
  SG_TRY(sg::alloc_device(dev, (void**)&w->gb_sums, 3 * sizeof(GbSums)));
  SG_TRY(sg::alloc_device(dev, (void**)&w->gb_slots, 2 * kGbMaxWgs * sizeof(GbSlot)));
  for (int k = 0; k < 2; ++k) {
    SG_TRY(sg::alloc_device(dev, (void**)&w->gb_small[k], sizeof(QEntry) * gb_cap_small(arcs)));
    SG_TRY(sg::alloc_device(dev, (void**)&w->gb_pieces[k], sizeof(QEntry) * gb_cap_pieces(arcs)));
  }
  w->gb_arcs = arcs;
  return SG_OK;
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
      (st = sg::alloc_device(dev, (void**)&w->ctl, sizeof(BfsCtl))) ||
      (st = sg::alloc_device(dev, (void**)&w->burst_list[0], sizeof(QEntry) * kBurstMaxArcs)) ||
      (st = sg::alloc_device(dev, (void**)&w->burst_list[1], sizeof(QEntry) * kBurstMaxArcs)) ||
      (st = sg::alloc_device(dev, (void**)&w->burst_list[2], sizeof(QEntry) * kBurstMaxArcs)) ||
      (st = sg::alloc_device(dev, (void**)&w->burst_list[3], sizeof(QEntry) * kBurstMaxArcs)) ||
      (st = sg::alloc_device(dev, (void**)&w->arrive, 8 * kArriveWords)) ||
      (st = sg::zero_async(dev, w->arrive, 8 * kArriveWords)) ||
      (st = w->qs.init(dev, g->n, g->m, default_tuning()))) {
    ws_free(w);
    return st;
  }
  bool ok = hipHostMalloc((void**)&w->levlog, sizeof(LevelLog) * SG_MAX_LEVELS_RECORDED, hipHostMallocCoherent) ==
                hipSuccess &&
            hipHostMalloc((void**)&w->steplog, sizeof(StepLog) * kRing, hipHostMallocCoherent) == hipSuccess;
  for (int r = 0; r < kRing && ok; ++r)
    ok = hipEventCreate(&w->ev_mid[r]) == hipSuccess && hipEventCreate(&w->ev_done[r]) == hipSuccess;
  ok = ok && hipEventCreate(&w->ev_begin) == hipSuccess && hipEventCreate(&w->ev_end) == hipSuccess &&
       hipEventCreate(&w->ev_first) == hipSuccess;
  if (!ok) {
    ws_free(w);
    sg::set_error("bfs workspace: event or pinned log allocation failed");
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

/* Host side of the level loop.  Each step is a fixed sequence of three
 * launches (traverse; decide + count, then scan + finish in the last
 * workgroup; write, then a burst in the last workgroup) whose
 * behaviour the device-resident control block selects, so the host never
 * has to decide anything between levels: it keeps `lookahead` steps enqueued
 * beyond the one it waits for and only reads each step's record (pinned host
 * memory) after that step's completion event, to stop after the first empty
 * level.  A step runs one level, plus the levels of a burst when the next
 * frontiers are small.  Steps enqueued beyond the last level find `done` set
 * and return at once. */
template<typename OffT>
int run_bfs(sg_graph_t g, int32_t src, const sg_bfs_options& opt, int32_t* depth, sg_bfs_stats* stats,
            bool sync_end = true, int64_t* reached = nullptr) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  BfsWorkspace* ws = nullptr;
  SG_TRY(workspace_for(g, &ws));
  SG_TRY(sg::graph_ensure_isolated(g));
  if (opt.policy != SG_POLICY_PUSH) SG_TRY(sg::graph_ensure_pull_hint(g));
  const int64_t n = g->n, w0 = ws->w0;
  /* the level kernel's workgroups, all resident at once (a pull gives each wave a chunk of the bitmap);
   * four per compute unit measured best on RMAT scale 24 (2 to 8 swept: 14 % less pull time than 8) */
  const int grid_level = dev->cus * (opt.blocks_per_cu > 0 ? opt.blocks_per_cu : 4);
  Tuning tune = default_tuning();
  if (opt.t_medium > 0) tune.t_medium = opt.t_medium;
  if (opt.t_large > 0) tune.t_large = opt.t_large;
  if (opt.chunk > 0) tune.chunk = opt.chunk;
  if (!ws->qs.fits(g->m, tune)) { /* queues sized for a coarser tuning: grow them */
    SG_HIP(hipStreamSynchronize(s));
    ws->qs.release();
    SG_TRY(ws->qs.init(dev, n, g->m, tune));
  }
  QueueSet& Q = ws->qs;
  const OffT* off = (const OffT*)g->offsets;

  BfsArgs<OffT> a{};
  a.off = off;
  a.cols = g->cols;
  a.in_off = (const OffT*)g->inv_offsets;
  a.in_cols = g->inv_cols;
  a.hint = g->pull_hint;
  a.hint_offsets = g->pull_hint_offsets ? 1 : 0;
  a.directed = (g->flags & SG_GRAPH_DIRECTED) ? 1 : 0;
  a.n = n;
  a.w0 = w0;
  a.visited = ws->visited;
  a.depth = depth;
  a.cur_bm = ws->bm[0];

  /* work-mapping variant under the fixed policies (see sg_bfs_options) */
  const int32_t lb_fixed = opt.load_balancer == SG_LB_WORKGROUP ? 1 : (opt.load_balancer == SG_LB_SUBGROUP ? 2 : 0);
  const int64_t burst_arcs = opt.burst_arcs < 0 ? 0 : (opt.burst_arcs == 0 ? kBurstDefaultArcs : opt.burst_arcs);
  const int64_t gb_arcs = opt.grid_burst_arcs < 0 ? 0 : (opt.grid_burst_arcs == 0 ? kGbDefaultArcs : opt.grid_burst_arcs);
  int gb_grid = 0;
  if (gb_arcs > 0 && opt.policy != SG_POLICY_PULL) {
    SG_TRY(gb_buffers(g, ws, gb_arcs));
    static int occ = 0; /* per offset width (one instantiation each) */
    if (occ == 0) {
      int nb = 0;
      SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gburst_kernel<OffT>, kGbThreads, 0));
      occ = std::max(1, nb);
    }
    /* at most what is resident at once (the barrier needs every workgroup) */
    gb_grid = std::min(opt.grid_burst_workgroups > 0 ? opt.grid_burst_workgroups : kGbDefaultWgs, dev->cus * occ);
    gb_grid = std::min(gb_grid, kGbMaxWgs);
  }
  bfs_start_kernel<OffT><<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(
      a, ws->prev, g->isolated, reinterpret_cast<unsigned long long*>(ws->stats),
      (int64_t)(sizeof(StatSlot) * kSlots * 2 / 8), Q.q[0][0], Q.q[0][1], Q.q[0][2], Q.hdr, tune, src, ws->ctl,
      opt.policy == SG_POLICY_PULL ? 1 : 0, opt.policy == SG_POLICY_LEARNED ? 0 : lb_fixed, burst_arcs);
  SG_HIP(hipGetLastError());

  const bool timed = opt.record_levels && stats;
  const bool want_total = stats != nullptr;
  const int lookahead = opt.sync_per_level ? 0 : 1;
  DecideParams P;
  P.policy = opt.policy;
  P.alpha = opt.alpha;
  P.beta = opt.beta;
  P.n = n;
  P.total_degree = (unsigned long long)g->total_degree;
  P.lb_fixed = lb_fixed;
  P.mlp = opt.policy == SG_POLICY_LEARNED ? sg::policy_device_params(opt.learned) : nullptr;
  P.f6 = sg::policy::feature(6, 0, 0, 0, 0, 0, (double)n, (double)g->m);
  P.f7 = sg::policy::feature(7, 0, 0, 0, 0, 0, (double)n, (double)g->m);
  P.burst_arcs = burst_arcs;
  P.learned_gate = opt.learned_gate < 0 ? 0ull
                                        : (unsigned long long)(opt.learned_gate == 0 ? kLearnedGateDefault : opt.learned_gate);
  if (want_total) SG_HIP(hipEventRecord(ws->ev_begin, s));

  /* the burst of a step reads the queue set its count pass wrote */
  auto burst_args = [&](int qset) {
    BurstArgs<OffT> ba;
    ba.off = off;
    ba.cols = g->cols;
    ba.visited = ws->visited;
    ba.prev = ws->prev;
    ba.depth = depth;
    for (int k = 0; k < 3; ++k) ba.q[k] = Q.q[qset][k];
    ba.hdr = Q.hdr + qset;
    for (int k = 0; k < 4; ++k) ba.list[k >> 1][k & 1] = ws->burst_list[k];
    ba.tune = tune;
    return ba;
  };
  auto enqueue = [&](int32_t step, bool gb) -> int {
    BfsArgs<OffT> la = a;
    const int cur = step & 1, nxt = (step + 1) & 1, r = step % kRing;
    la.stats = ws->stats + kSlots * cur;
    la.stats_zero = ws->stats + kSlots * nxt;
    for (int k = 0; k < 3; ++k) la.qin[k] = Q.q[cur][k];
    la.qin_hdr = Q.hdr + cur;
    la.cur_bm = ws->bm[cur];
    la.next_bm = ws->bm[nxt];
    CompactArgs<OffT> ca = make_compact_args<OffT>(Q, nxt, off, ws->visited, ws->prev, ws->bm[nxt], tune);
    ca.depth = depth; /* depths of the vertices a push level discovered */
    const BurstArgs<OffT> ba = burst_args(nxt);
    /* timing: the traversal kernel runs from the previous step's end (or
     * ev_first) to ev_mid; each event is a stream barrier, so only these */
    if (timed && step == 0) SG_HIP(hipEventRecord(ws->ev_first, s));
    if (gb) { /* a grid burst ahead of the traversal kernel (returns at once unless the level qualifies) */
      GbArgs<OffT> gba;
      gba.off = off;
      gba.cols = g->cols;
      gba.visited = ws->visited;
      gba.prev = ws->prev;
      gba.depth = depth;
      for (int k = 0; k < 3; ++k) gba.q[k] = Q.q[cur][k];
      gba.hdr = Q.hdr + cur;
      for (int k = 0; k < 2; ++k) {
        gba.small[k] = ws->gb_small[k];
        gba.pieces[k] = ws->gb_pieces[k];
      }
      gba.cap_small = gb_cap_small(gb_arcs);
      gba.cap_pieces = gb_cap_pieces(gb_arcs);
      gba.sums = ws->gb_sums;
      gba.slots = ws->gb_slots;
      gba.bar = ws->gb_bar;
      gba.gb_arcs = gb_arcs;
      gba.wait_ticks = dev->wait_ticks;
      gb_reset_kernel<<<1, 256, 0, s>>>(ws->gb_bar, ws->gb_sums);
      int32_t stp = step;
      void* args[] = {&gba, &ws->ctl, &P, &stp, &ws->levlog, &ws->steplog};
      SG_HIP(launch_resident(reinterpret_cast<const void*>(gburst_kernel<OffT>), gb_grid, kGbThreads, args, s));
    }
    level_kernel<OffT><<<grid_level, kBlock, 0, s>>>(la, ws->ctl);
    if (timed) SG_HIP(hipEventRecord(ws->ev_mid[r], s));
    FinishArgs fa{Q.blocks, Q.nblocks, Q.hdr + nxt, step, ws->levlog, ws->steplog, ws->arrive};
    count_dyn_kernel<OffT><<<Q.grid_tiles(), kBlock, 0, s>>>(ca, ws->ctl, la.stats, P, fa);
    write_dyn_kernel<OffT><<<(Q.grid_tiles() + 1) / 2, kBurstThreads, 0, s>>>(ca, ba, ws->ctl, P, step, ws->levlog,
                                                                             ws->steplog, ws->arrive + kArriveWords);
    SG_HIP(hipGetLastError());
    /* the host waits for an untimed step by polling its record: an event per
     * step is a stream barrier, a few microseconds of every level */
    if (timed) SG_HIP(hipEventRecord(ws->ev_done[r], s));
    return SG_OK;
  };

  /* no record of an earlier traversal may pass for one of this one (the
   * stream is idle: the previous call ended with a synchronize) */
  for (int k = 0; k < kRing; ++k) __atomic_store_n(&ws->steplog[k].step, -1, __ATOMIC_RELEASE);
  /* a start in a burst: the levels from the source while they stay small,
   * handed back to step 0 (which then only counts and decides) or ending the
   * traversal; the kernel returns at once unless init set ctl->burst */
  if (burst_arcs > 0 && opt.policy != SG_POLICY_PULL)
    burst_kernel<OffT><<<1, kBurstThreads, 0, s>>>(burst_args(0), ws->ctl, P, 0, ws->levlog, ws->steplog);
  SG_TRY(enqueue(0, false));
  for (int k = 1; k <= lookahead; ++k) SG_TRY(enqueue(k, false));
  StepLog last{};
  int32_t first_level = 0;
  /* steps in a row that left a push level of at most gb_arcs arcs (and more
   * than a single-workgroup burst takes) for the next step; from two on, the
   * steps enqueued start with a grid burst */
  int32_t small_push = 0;
  for (int32_t step = 0;; ++step) {
    const int r = step % kRing;
    if (timed) SG_HIP(hipEventSynchronize(ws->ev_done[r]));
    else SG_TRY(sg::wait_host_flag(&ws->steplog[r].step, step, s, "sg_bfs"));
    std::memcpy(&last, (const void*)&ws->steplog[r], sizeof(last));
    if (last.fail) {
      SG_HIP(hipStreamSynchronize(s));
      sg::set_error("sg_bfs: a grid barrier waited longer than the device's wait limit (step %d)", step);
      return SG_ERR_HIP;
    }
    if (last.step != step || last.last_level < first_level) {
      SG_HIP(hipStreamSynchronize(s));
      sg::set_error("sg_bfs: step record %d out of order (found step %d, level %d)", step, last.step,
                    last.last_level);
      return SG_ERR_STATE;
    }
    if (stats) {
      float k_ms = 0.f, c_ms = 0.f;
      if (timed) {
        const hipEvent_t start = step == 0 ? ws->ev_first : ws->ev_done[(step + kRing - 1) % kRing];
        SG_HIP(hipEventElapsedTime(&k_ms, start, ws->ev_mid[r]));
        SG_HIP(hipEventElapsedTime(&c_ms, ws->ev_mid[r], ws->ev_done[r]));
      }
      for (int32_t L = first_level; L <= last.last_level && L < SG_MAX_LEVELS_RECORDED; ++L) {
        LevelLog e;
        std::memcpy(&e, (const void*)&ws->levlog[L], sizeof(e));
        sg_level_stat& ls = stats->level[L];
        std::memset(&ls, 0, sizeof(ls));
        ls.direction = e.dir;
        ls.load_balancer = opt.load_balancer;
        ls.frontier_vertices = e.f_vertices;
        ls.frontier_edges = e.f_edges;
        ls.discovered = e.disc;
        ls.edges_inspected = e.edges;
        ls.vertices_scanned = e.scanned;
        ls.hint_misses = e.misses;
        ls.max_degree = e.max_deg;
        ls.unexplored_edges = e.unexplored;
        ls.action = e.action;
        ls.burst = e.burst;
        if (L == first_level) { /* the step's later levels ran in a burst, timed with the count and write */
          ls.kernel_ms = k_ms;
          ls.convert_ms = c_ms;
        }
      }
    }
    if (last.done) break;
    if ((int64_t)last.last_level + 1 >= n + 1) { /* more levels than vertices: the control block is corrupt */
      SG_HIP(hipStreamSynchronize(s));
      sg::set_error("sg_bfs: no termination after %lld levels", (long long)(last.last_level + 1));
      return SG_ERR_STATE;
    }
    first_level = last.last_level + 1;
    const bool qualifies = gb_grid > 0 && last.next_dir == 0 && last.next_f_edges > burst_arcs &&
                           last.next_f_edges <= gb_arcs;
    small_push = qualifies ? small_push + 1 : 0;
    SG_TRY(enqueue(step + 1 + lookahead, small_push >= 2));
  }
  if (reached) { /* reached vertices and their arcs */
    reached[0] = last.reached_v;
    reached[1] = last.reached_e;
  }
  /* a traversal of a batch leaves the stream running: the next one's start
   * is ordered after this one's trailing steps (sg_bfs_many) */
  if (!sync_end) return SG_OK;
  if (want_total) SG_HIP(hipEventRecord(ws->ev_end, s));
  SG_HIP(hipStreamSynchronize(s));
  if (stats) {
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, ws->ev_begin, ws->ev_end));
    stats->kernel_ms = ms;
    stats->levels = last.last_level + 1;
    stats->n_push = last.n_push;
    stats->n_pull = last.n_pull;
    stats->reached_vertices = last.reached_v;
    stats->reached_edges = last.reached_e;
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

} // extern "C"

namespace {
/* The options sg_bfs and sg_bfs_many accept (opt_in null: the defaults). */
int bfs_options_checked(const sg_bfs_options* opt_in, sg_bfs_options& opt) {
  if (opt_in)
    opt = *opt_in;
  else
    sg_bfs_options_default(&opt);
  SG_CHECK(opt.policy >= SG_POLICY_PUSH && opt.policy <= SG_POLICY_LEARNED, SG_ERR_UNSUPPORTED,
           "sg_bfs: policy %d not supported", opt.policy);
  SG_CHECK(opt.policy != SG_POLICY_LEARNED || opt.learned, SG_ERR_INVALID,
           "sg_bfs: the learned policy needs a trained sg_policy_t in options.learned");
  SG_CHECK(opt.load_balancer >= SG_LB_WORKITEM && opt.load_balancer <= SG_LB_BUCKETING, SG_ERR_INVALID,
           "sg_bfs: load balancer %d out of range", opt.load_balancer);
  SG_CHECK(opt.t_medium >= 0 && opt.t_large >= 0 && opt.chunk >= 0 && opt.blocks_per_cu >= 0 &&
               opt.blocks_per_cu <= 64 && (opt.chunk == 0 || opt.chunk >= 64) &&
               (opt.t_medium == 0 || opt.t_large == 0 || opt.t_large > opt.t_medium) &&
               opt.t_medium <= (1 << 30) && opt.t_large <= (1 << 30) && opt.chunk <= (1 << 30),
           SG_ERR_INVALID, "sg_bfs: tuning out of range (t_medium %d, t_large %d, chunk %d, blocks_per_cu %d)",
           opt.t_medium, opt.t_large, opt.chunk, opt.blocks_per_cu);
  SG_CHECK(opt.burst_arcs <= kBurstMaxArcs, SG_ERR_INVALID, "sg_bfs: burst_arcs %d above %d", opt.burst_arcs,
           kBurstMaxArcs);
  SG_CHECK(opt.grid_burst_arcs <= kGbMaxArcs && opt.grid_burst_workgroups >= 0 &&
               opt.grid_burst_workgroups <= kGbMaxWgs,
           SG_ERR_INVALID, "sg_bfs: grid_burst_arcs %d above %lld or grid_burst_workgroups %d outside [0, %d]",
           opt.grid_burst_arcs, (long long)kGbMaxArcs, opt.grid_burst_workgroups, kGbMaxWgs);
  SG_CHECK(opt.policy != SG_POLICY_HYBRID || (opt.alpha > 0.f && opt.beta > 0.f), SG_ERR_INVALID,
           "sg_bfs: alpha and beta must be positive");
  return SG_OK;
}
} // namespace

extern "C" {

int sg_bfs(sg_graph_t g, int32_t source, const sg_bfs_options* opt_in, int32_t* depth, sg_bfs_stats* stats) {
  SG_CHECK(g && depth, SG_ERR_INVALID, "sg_bfs: null argument");
  SG_CHECK(source >= 0 && source < g->n, SG_ERR_INVALID, "sg_bfs: source %d out of range [0,%lld)", source,
           (long long)g->n);
  sg_bfs_options opt;
  SG_TRY(bfs_options_checked(opt_in, opt));
  if (stats) {
    stats->levels = stats->n_push = stats->n_pull = 0;
    stats->reached_vertices = stats->reached_edges = 0;
    stats->kernel_ms = 0.f;
  }
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return run_bfs<int64_t>(g, source, opt, depth, stats);
  return run_bfs<uint32_t>(g, source, opt, depth, stats);
}

/* Single-source BFS from each of nsrc sources, one after another on the
 * device: the next source's start is enqueued as soon as the host sees the
 * previous traversal's last level, with no synchronisation in between. */
int sg_bfs_many(sg_graph_t g, int32_t nsrc, const int32_t* sources, const sg_bfs_options* opt_in, int32_t* depth,
                sg_bfs_many_stats* stats) {
  SG_CHECK(g && depth && (nsrc == 0 || sources), SG_ERR_INVALID, "sg_bfs_many: null argument");
  SG_CHECK(nsrc >= 0, SG_ERR_INVALID, "sg_bfs_many: negative source count");
  for (int32_t i = 0; i < nsrc; ++i)
    SG_CHECK(sources[i] >= 0 && sources[i] < g->n, SG_ERR_INVALID, "sg_bfs_many: source %d out of range [0,%lld)",
             sources[i], (long long)g->n);
  sg_bfs_options opt;
  SG_TRY(bfs_options_checked(opt_in, opt));
  opt.record_levels = 0;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  SG_HIP(hipSetDevice(g->dev->ordinal));
  hipStream_t s = g->dev->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, s));
  }
  int rc = SG_OK;
  for (int32_t i = 0; i < nsrc && rc == SG_OK; ++i) {
    int64_t reached[2] = {0, 0};
    int32_t* row = depth + (int64_t)i * g->n;
    rc = g->offset_bytes == 8 ? run_bfs<int64_t>(g, sources[i], opt, row, nullptr, false, reached)
                              : run_bfs<uint32_t>(g, sources[i], opt, row, nullptr, false, reached);
    if (stats) {
      stats->reached_vertices += reached[0];
      stats->reached_edges += reached[1];
    }
  }
  if (stats) (void)hipEventRecord(e1, s);
  const hipError_t se = hipStreamSynchronize(s);
  if (stats) {
    float ms = 0.f;
    if (rc == SG_OK && se == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) stats->kernel_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (rc != SG_OK) return rc;
  SG_HIP(se);
  return SG_OK;
}

} // extern "C"
