/*
 * sssp.hip -- single-source shortest paths for gfx950 (SSSP::run,
 * algorithms/sssp.hpp:152-210).
 *
 * Label-correcting relaxation, as the reference: a round relaxes all arcs of
 * its frontier with an atomic minimum on the tentative distance
 * (sssp.hpp:180-189, sync/atomics.hpp:76-81) and vertices whose distance
 * dropped become dirty again (the filter of sssp.hpp:192-196 only
 * deduplicates them, which a bitmap does by construction).  The reference
 * relaxes every dirty vertex each round (frontier Bellman-Ford); here a round
 * takes only the dirty vertices with distance below a threshold T (the near
 * set) and defers the rest, raising T to (smallest deferred distance) + delta
 * when the near set runs dry -- the bucket order of delta-stepping, which
 * relaxes far fewer arcs on weighted graphs.  delta = infinity is exactly the
 * reference's schedule.  Within a bucket, light iterations relax only the
 * arcs lighter than delta and one heavy iteration the others; the adjacency is
 * read from a light-first copy (per vertex the light arcs, then the heavy
 * ones) so each phase reads only its own arcs.  Each round reads its frontier
 * through the degree-class queues (thread / wave / workgroup bins), relaxes
 * with a plain-load pre-check before the atomic, and marks improved vertices
 * dirty with a test-then-OR.  Shortest distances are unique, so the result does not
 * depend on delta or on the relaxation order.
 *
 * Distances have the weights' 4-byte type: uint32 weights use integer min,
 * float weights (non-negative) compare through their bit patterns.  The
 * unreached value is n+1 in that type (sssp.hpp:43).
 */
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>

#include "traverse.hpp"

using namespace sg::trav;

namespace {

template<typename W>
struct DistOps;
template<>
struct DistOps<uint32_t> {
  static constexpr bool is_float = false;
  __host__ __device__ static uint32_t bits(uint32_t x) { return x; }
  __host__ __device__ static uint32_t value(uint32_t b) { return b; }
};
template<>
struct DistOps<float> {
  static constexpr bool is_float = true;
  __host__ __device__ static uint32_t bits(float x) { return __builtin_bit_cast(uint32_t, x); }
  __host__ __device__ static float value(uint32_t b) { return __builtin_bit_cast(float, b); }
};

template<typename OffT, typename W>
struct SsspArgs {
  const int32_t* cols;
  const W* w;
  uint32_t* dist; /* distance bits */
  uint64_t* next_bm;
  int64_t w0;
  QEntry* qin[3];
  const QueueHdr* qin_hdr;
  /* heavy rounds that pull (light-first adjacency of an undirected graph) */
  const OffT* off;
  const int32_t* light;      /* light arcs per vertex */
  const uint64_t* settled;   /* every vertex settled by a finished bucket (including the current one) */
  const uint64_t* frontier;  /* the current bucket's settled set */
  int64_t n;
};

/* Device-resident control block of the round loop (written only by
 * the last workgroup of sssp_select_kernel; kernel boundaries order the
 * accesses). */
struct SsspCtl {
  uint32_t T_bits;    /* threshold: near = distance below T */
  int32_t done;       /* no dirty vertex left: later iterations are no-ops */
  int32_t relax;      /* the current iteration relaxes (its near set is not empty) */
  int32_t rounds;     /* iterations that relaxed so far */
  int32_t heavy;      /* the current iteration relaxes the heavy arcs of the bucket's settled set */
  int32_t heavy_next; /* the next iteration is the heavy one */
  int32_t r_pending;  /* vertices were relaxed in this bucket since its heavy iteration */
  int32_t pull;       /* the current (heavy) iteration pulls into the vertices not yet settled */
  int32_t burst;      /* the current (Bellman-Ford) round is small: the write pass's last workgroup runs it
                         and the following small rounds */
  int32_t pad1;
  unsigned long long relaxed;    /* arcs relaxed so far */
  unsigned long long heavy_done; /* heavy arcs of every settled set so far */
};

struct alignas(64) RoundLog {
  int32_t iter, done, relax, rounds;
  int64_t near_v, near_e;
  unsigned long long relaxed;
};

constexpr int kRing = 8;

/* Light iterations relax the arcs lighter than delta, the heavy iteration of
 * a bucket the others (delta-stepping's light / heavy split: a heavy arc
 * from a settled bucket cannot land inside it, so it is relaxed once).  The
 * adjacency is the light-first copy, and the queue entries already cover only
 * the phase's part of each list, so every visited arc is relaxed. */
constexpr int kPullHead = 32; /* heavy arcs a lane scans alone before the wave helps */

/* Heavy round by pull (undirected graphs; in-arcs are the out-arcs and the
 * weights are symmetric): a heavy arc (u, v) can only lower dist[v] if v is
 * not settled yet (dist[u] >= T - delta and w >= delta give at least T), so
 * each vertex not yet settled takes the minimum of dist[u] + w over its heavy
 * arcs from the bucket's settled set and becomes dirty if that is below its
 * distance.  Every such vertex belongs to one lane, so its distance is
 * written without an atomic, and the settled distances it reads are final.
 * A wave loads 64 words of the settled bitmap and compacts their clear bits;
 * lists longer than kPullHead are finished by the whole wave.  Returns the
 * arcs from the settled set that were relaxed. */
template<typename OffT, typename W>
__device__ __forceinline__ unsigned long long heavy_pull(const SsspArgs<OffT, W>& a) {
  unsigned long long cnt = 0;
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const int tail_bits = (int)(a.n & 63);
  for (int64_t base = gwave * 64; base < a.w0; base += nwaves * 64) {
    const int64_t myw = base + lane;
    uint64_t myun = myw < a.w0 ? ~a.settled[myw] : 0ull;
    if (myw == a.w0 - 1 && tail_bits) myun &= (1ull << tail_bits) - 1ull;
    const uint32_t c = (uint32_t)__popcll(myun);
    const uint32_t incl = wave_inclusive_scan(c);
    const uint32_t total = __shfl(incl, 63, 64);
    for (uint32_t t0 = 0; t0 < total; t0 += 64) {
      const uint32_t t = t0 + lane;
      const int owner = scan_owner(incl, t);
      const uint32_t before = __shfl(incl - c, owner, 64);
      const uint64_t un = shfl_u64(myun, owner);
      const bool active = t < total;
      int64_t v = 0, j = 0, e = 0;
      uint32_t dv = 0, best = 0xFFFFFFFFu;
      if (active) {
        v = (base + owner) * 64 + select_bit(un, t - before);
        j = (int64_t)a.off[v] + a.light[v];
        e = (int64_t)a.off[v + 1];
        dv = a.dist[v];
        best = dv;
        const int64_t he = e - j > kPullHead ? j + kPullHead : e;
        for (; j < he; j += 4) { /* four ids, then four bitmap probes, then the hits' distances */
          uint32_t u[4];
          bool h[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) u[k] = j + k < he ? (uint32_t)a.cols[j + k] : 0u;
#pragma unroll
          for (int k = 0; k < 4; ++k) h[k] = j + k < he && test_bit(a.frontier, u[k]);
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (h[k]) {
              ++cnt;
              const uint32_t nb = DistOps<W>::bits(DistOps<W>::value(a.dist[u[k]]) + a.w[j + k]);
              best = nb < best ? nb : best;
            }
        }
        if (j > he) j = he;
      }
      uint64_t pending = ballot(active && j < e);
      while (pending) {
        const int k2 = __ffsll((unsigned long long)pending) - 1;
        pending &= pending - 1;
        const int64_t jb = shfl_i64(j, k2), je = shfl_i64(e, k2);
        uint32_t m = 0xFFFFFFFFu;
        for (int64_t jj = jb + lane; jj < je; jj += 64) {
          const uint32_t u = (uint32_t)a.cols[jj];
          if (test_bit(a.frontier, u)) {
            ++cnt;
            const uint32_t nb = DistOps<W>::bits(DistOps<W>::value(a.dist[u]) + a.w[jj]);
            m = nb < m ? nb : m;
          }
        }
        m = ~(uint32_t)wave_max_u64(~(unsigned long long)m); /* wave minimum */
        if (lane == k2) best = m < best ? m : best;
      }
      if (active && best < dv) {
        a.dist[v] = best;
        atomic_or(&a.next_bm[v >> 6], 1ull << (v & 63));
      }
    }
  }
  return cnt;
}

template<typename OffT, typename W>
__global__ void __launch_bounds__(kBlock) relax_kernel(SsspArgs<OffT, W> a, const SsspCtl* ctl,
                                                       unsigned long long* relax_slots) {
  if (ctl->done || !ctl->relax) return;
  unsigned long long cnt = 0;
  if (ctl->pull) {
    cnt = heavy_pull(a);
  } else {
    for_each_queued_arc<OffT>(a.qin, a.qin_hdr, a.cols, [&](uint32_t u, uint32_t v, int64_t e) {
      const W wt = a.w[e];
      ++cnt;
      /* a stale dist[u] is an upper bound: u is dirty again if it improved */
      const W du = DistOps<W>::value(a.dist[u]);
      const W nd = du + wt;
      const uint32_t nb = DistOps<W>::bits(nd);
      if (nb < a.dist[v]) {
        const uint32_t old = atomic_min(&a.dist[v], nb);
        if (nb < old) {
          const uint64_t bit = 1ull << (v & 63);
          if (!(a.next_bm[v >> 6] & bit)) atomic_or(&a.next_bm[v >> 6], bit);
        }
      }
    });
  }
  /* arcs relaxed in this phase: one atomic per workgroup into 64 slots */
  __shared__ unsigned long long red[kBlock / 64];
  cnt = wave_sum_u64(cnt);
  if (lane_id() == 0) red[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += red[w];
    if (t) atomic_add(&relax_slots[(blockIdx.x % kSlots) * 16], t);
  }
}

/* The light-first adjacency: a wave per vertex writes the arcs lighter than
 * dlt from the front of the vertex's range (in order) and the others from the
 * back, and records the light count. */
template<typename OffT, typename W>
__global__ void __launch_bounds__(kBlock) split_build_kernel(const OffT* off, const int32_t* cols, const W* w,
                                                             int64_t n, W dlt, int32_t* scols, W* sw,
                                                             int32_t* light, unsigned long long* heavy_total) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  unsigned long long heavy = 0; /* lane 0: heavy arcs of the wave's vertices */
  for (int64_t v = gwave; v < n; v += nwaves) {
    const int64_t b = (int64_t)off[v], e = (int64_t)off[v + 1];
    int64_t front = b, back = e;
    for (int64_t j0 = b; j0 < e; j0 += 64) {
      const int64_t j = j0 + lane;
      const bool ok = j < e;
      int32_t c = 0;
      W x = 0;
      if (ok) {
        c = cols[j];
        x = w[j];
      }
      const bool lt = ok && x < dlt, hv = ok && !(x < dlt);
      const uint64_t ml = ballot(lt), mh = ballot(hv);
      if (lt) {
        const int64_t p = front + lanes_below(ml);
        scols[p] = c;
        sw[p] = x;
      }
      if (hv) {
        const int64_t p = back - 1 - lanes_below(mh);
        scols[p] = c;
        sw[p] = x;
      }
      front += __popcll(ml);
      back -= __popcll(mh);
    }
    if (lane == 0) {
      light[v] = (int32_t)(front - b);
      heavy += (unsigned long long)(e - front);
    }
  }
  if (lane == 0 && heavy) atomic_add(heavy_total, heavy);
}

/* per-block counters of the split, one 128-byte line per slot */
struct alignas(128) SplitSlot {
  unsigned long long far_count;
  unsigned int far_min; /* distance bits; 4-byte distances order like their bits */
  unsigned int pad0;
  unsigned long long pad[14];
};

/* near = dirty vertices with distance < T (moved out of the dirty bitmap into
 * the round's frontier bitmap); count and smallest distance of the rest.  One
 * word per thread (word i), its dirty bits walked by the lane itself: on the
 * dense bands of a grid this beats spreading them over the wave.  The heavy
 * iteration's frontier is the bucket's settled set instead.  All threads of
 * the workgroup call. */
template<typename W>
__device__ __forceinline__ void split_word(int64_t i, int64_t w0, bool heavy_next, bool bellman, uint32_t T_bits,
                                           uint64_t* dirty, uint64_t* near_bm, uint64_t* settled,
                                           uint64_t* settled_all, const uint32_t* dist, SplitSlot* slots) {
  if (heavy_next) {
    if (i < w0) {
      const uint64_t r = settled[i];
      near_bm[i] = r;
      if (r) {
        settled[i] = 0;
        settled_all[i] |= r;
      }
    }
    return;
  }
  if (bellman) { /* the reference's schedule: T is the largest weight, so every dirty vertex is near */
    if (i < w0) {
      const uint64_t f = dirty[i];
      near_bm[i] = f;
      if (f) {
        dirty[i] = 0;
        settled[i] |= f;
      }
    }
    return;
  }
  const W T = DistOps<W>::value(T_bits);
  unsigned long long far = 0;
  unsigned int fmin = 0xFFFFFFFFu;
  if (i < w0) {
    const uint64_t f = dirty[i];
    uint64_t nb = 0;
    for (uint64_t rest = f; rest; rest &= rest - 1) {
      const int b = __ffsll((unsigned long long)rest) - 1;
      const uint32_t db = dist[i * 64 + b];
      if (DistOps<W>::value(db) < T) nb |= 1ull << b;
      else fmin = db < fmin ? db : fmin;
    }
    near_bm[i] = nb;
    if (nb) {
      dirty[i] = f & ~nb;
      settled[i] |= nb;
    }
    far = (unsigned long long)__popcll(f & ~nb);
  }
  far = wave_sum_u64(far);
  unsigned long long m = wave_max_u64(~(unsigned long long)fmin); /* min via max of complement */
  fmin = ~(unsigned int)m;
  if (lane_id() == 0 && (far || fmin != 0xFFFFFFFFu)) {
    SplitSlot* sl = slots + (blockIdx.x % kSlots);
    if (far) atomic_add(&sl->far_count, far);
    atomic_min(&sl->far_min, fmin);
  }
}

__global__ void split_slots_reset_kernel(SplitSlot* slots, const SsspCtl* ctl) {
  if (ctl->done) return;
  const int i = threadIdx.x;
  if (i < kSlots) {
    slots[i].far_count = 0;
    slots[i].far_min = 0xFFFFFFFFu;
  }
}

/* ------------------------------------------------------------ bursts --- */
/* Rounds of the reference's schedule (every dirty vertex relaxed each round)
 * on a graph of long diameter have small frontiers, and each costs a
 * selection, a write and a relax launch.  When a round has at most
 * burst_arcs arcs, the last workgroup of the write pass relaxes it and the
 * following rounds back to back, one workgroup barrier per round, for as
 * long as they stay small.  The frontier moves between two lists of (first
 * arc, vertex, degree) entries; a round lays the list's arcs end to end (an
 * exclusive scan of the lengths in LDS) and each thread takes every
 * kBlock-th arc, finding its entry by binary search.  A relaxation that
 * lowers dist[v] sets v's dirty bit; the lane that set it appends v to the
 * next list.  Before a round the bits of its list are cleared (the split's
 * job in the launched rounds), so a vertex lowered again during its own
 * round is relaxed again in the next.  The burst stops when a round lowers
 * nothing or the next list is too large: its vertices keep their dirty bits
 * and the next iteration's selection takes them from the bitmap. */
constexpr int kSsspBurstDefault = 4096;
constexpr int kSsspBurstMax = 8192; /* entries (and arcs) of a burst round; LDS for its scan */

template<typename OffT, typename W>
struct SsspBurst {
  const OffT* off;
  const int32_t* cols;
  const W* w;
  uint32_t* dist;
  uint64_t* dirty;
  QEntry* q[3]; /* the round's queues (written by this launch) */
  const QueueHdr* hdr;
  QEntry* list[2]; /* kSsspBurstMax entries each */
  int64_t burst_arcs;
  unsigned long long* relax_slots;
  uint32_t* arrive; /* kArriveWords zeroed words */
};

template<typename OffT, typename W>
__device__ __forceinline__ void sssp_burst(const SsspBurst<OffT, W>& b, SsspCtl* ctl) {
  __shared__ uint32_t s_pref[kSsspBurstMax + 1];
  __shared__ uint32_t s_wtot[kBlock / 64];
  __shared__ uint32_t s_next, s_next_arcs;
  const QueueHdr h = *b.hdr;
  const uint32_t n0 = h.count[0], n1 = h.count[1];
  uint32_t cnt = h.count[0] + h.count[1] + h.count[2];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  /* the first round's entries, the three class queues, gathered into the
   * list that the second round writes */
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock)
    b.list[1][i] = i < n0 ? b.q[0][i] : (i < n0 + n1 ? b.q[1][i - n0] : b.q[2][i - n0 - n1]);
  __syncthreads();
  unsigned long long arcs = 0;
  int32_t rounds = 0;
  for (int it = 0;; ++it) {
    const QEntry* in = b.list[(it + 1) & 1];
    QEntry* out = b.list[it & 1];
    auto entry = [&](uint32_t i) -> QEntry { return in[i]; };
    /* exclusive scan of the lengths: thread t owns a run of kPer entries */
    constexpr int kPer = kSsspBurstMax / kBlock;
    const uint32_t first = threadIdx.x * kPer;
    uint32_t run = 0;
    for (int k = 0; k < kPer; ++k)
      if (first + k < cnt) run += (uint32_t)entry(first + k).len;
    const uint32_t incl = wave_inclusive_scan(run);
    if (lane == 63) s_wtot[wv] = incl;
    if (threadIdx.x == 0) s_next = s_next_arcs = 0;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      base += w < wv ? s_wtot[w] : 0u;
      total += s_wtot[w];
    }
    uint32_t pos = base + incl - run;
    for (int k = 0; k < kPer; ++k)
      if (first + k < cnt) {
        s_pref[first + k] = pos;
        pos += (uint32_t)entry(first + k).len;
      }
    if (threadIdx.x == 0) s_pref[cnt] = total;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < total; t += kBlock) {
      uint32_t lo = 0, hi = cnt; /* s_pref[lo] <= t < s_pref[hi] */
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pref[mid] <= t) lo = mid;
        else hi = mid;
      }
      const QEntry e = entry(lo);
      const int64_t j = e.start + (int64_t)(t - s_pref[lo]);
      const uint32_t u = (uint32_t)e.v, v = (uint32_t)b.cols[j];
      /* distances lowered earlier in the burst were lowered by atomics at L2:
       * read them with agent-scope loads, past any older copy in L1 */
      const uint32_t du = __hip_atomic_load(&b.dist[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t nb = DistOps<W>::bits(DistOps<W>::value(du) + b.w[j]);
      if (nb < __hip_atomic_load(&b.dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
          nb < atomic_min(&b.dist[v], nb)) {
        const uint64_t bit = 1ull << (v & 63);
        if (!(atomic_or(&b.dirty[v >> 6], bit) & bit)) {
          const int64_t s0 = (int64_t)b.off[v], d = (int64_t)(b.off[v + 1] - b.off[v]);
          if (d > 0) { /* a vertex without arcs stays dirty for the selection */
            const uint32_t p = atomicAdd(&s_next, 1u);
            if (p < (uint32_t)kSsspBurstMax) {
              QEntry q;
              q.start = s0;
              q.v = (int32_t)v;
              q.len = (int32_t)d;
              out[p] = q;
            }
            atomicAdd(&s_next_arcs, (uint32_t)(d < (int64_t)kSsspBurstMax ? d : kSsspBurstMax));
          }
        }
      }
    }
    __syncthreads();
    arcs += total;
    ++rounds;
    const uint32_t nc = s_next, na = s_next_arcs;
    /* hand back (the next list keeps its dirty bits) or stop */
    if (nc == 0 || nc > (uint32_t)kSsspBurstMax || (int64_t)na > b.burst_arcs) break;
    for (uint32_t i = threadIdx.x; i < nc; i += kBlock) {
      const uint32_t v = (uint32_t)out[i].v;
      atomic_and(&b.dirty[v >> 6], ~(1ull << (v & 63)));
    }
    __syncthreads();
    cnt = nc;
  }
  if (threadIdx.x == 0) {
    /* the selection counted the first round; the relax launch has nothing left */
    ctl->rounds += rounds - 1;
    ctl->relaxed += arcs - (unsigned long long)h.deg;
    ctl->relax = 0;
    ctl->burst = 0;
    atomic_add(&b.relax_slots[0], arcs);
  }
}

template<typename OffT, typename W>
__global__ void __launch_bounds__(kBlock) near_write_kernel(CompactArgs<OffT> a, SsspBurst<OffT, W> b, SsspCtl* ctl) {
  if (ctl->done || !ctl->relax || ctl->pull) return;
  const bool burst = ctl->burst != 0;
  const bool wrote = compact_write_body<OffT>(a, ctl->heavy ? 1 : 0);
  /* a burst round has at most burst_arcs arcs, so few waves write entries and publish them */
  if (!burst || !last_to_arrive<false>(b.arrive, wrote)) return;
  sssp_burst(b, ctl);
}

/* Threshold after a dry iteration: the smallest deferred distance plus
 * delta, at least one representable step above it. */
template<typename W>
__device__ __forceinline__ uint32_t next_threshold(uint32_t fmin_bits, double delta, double w_max) {
  const W dmin = DistOps<W>::value(fmin_bits);
  const double t = (double)dmin + delta;
  W tn = (W)(t < w_max ? t : w_max);
  if (!(tn > dmin)) {
    /* non-negative finite floats order like their bit patterns: +1 is the next float up */
    if constexpr (DistOps<W>::is_float) tn = DistOps<W>::value(fmin_bits + 1u);
    else tn = dmin + 1u;
  }
  return DistOps<W>::bits(tn);
}

/* Where a selection launch keeps its scratch and writes its decision. */
struct SelectArgs {
  uint64_t *dirty, *near_bm, *settled, *settled_all;
  const uint32_t* dist;
  SplitSlot* slots;
  int64_t w0, nblocks; /* bitmap words; count-pass workgroups (sums to scan) */
  uint32_t* arrive;    /* kArriveWords zeroed words */
  RoundLog* log;
  const unsigned long long* heavy_total;
  double delta, w_max;
  int32_t pull_mode;
  int32_t bellman;    /* the reference's schedule (no buckets) */
  int64_t burst_arcs; /* rounds of the reference's schedule with at most this many arcs run as a burst (0: never) */
};

/* One iteration's selection, count and decision in one launch.  Each
 * workgroup splits the 256 words of the dirty bitmap that its four count
 * tiles cover (thread t splits the word that it then counts, so no other
 * workgroup's near bits are needed), counts the near set's queue entries,
 * and the last workgroup to finish scans the counts and decides: relax /
 * advance the threshold / stop, and the iteration's record into the pinned
 * ring.  A heavy iteration pulls (pull_mode 1), pushes (-1), or (0) pulls
 * when the heavy arcs of the vertices not yet settled -- all heavy arcs less
 * those of every settled set so far, in-arcs being out-arcs on an undirected
 * graph -- are fewer than those of the bucket's settled set.  The far
 * statistics are added into `slots` with atomics; the last workgroup reads
 * them through agent-coherent loads and clears them for the next
 * iteration. */
template<typename OffT, typename W>
__global__ void __launch_bounds__(kBlock) sssp_select_kernel(CompactArgs<OffT> ca, SelectArgs sa, SsspCtl* ctl,
                                                             int32_t iter) {
  __shared__ QueueHdr tot;
  if (ctl->done) return;
  const bool heavy_next = ctl->heavy_next != 0;
  split_word<W>((int64_t)blockIdx.x * kBlock + threadIdx.x, sa.w0, heavy_next, sa.bellman != 0, ctl->T_bits,
                sa.dirty, sa.near_bm, sa.settled, sa.settled_all, sa.dist, sa.slots);
  compact_count_body<OffT, kFromBitmap>(ca, heavy_next ? 1 : 0);
  if (!last_to_arrive<true>(sa.arrive, false)) return;
  compact_scan_body_t<kBlock, 8, true>(ca.blocks, sa.nblocks, ca.hdr, &tot);
  __syncthreads();
  if (threadIdx.x != 0) return;
  RoundLog e{};
  e.iter = iter;
  e.near_v = (int64_t)tot.nvert;
  e.near_e = (int64_t)tot.deg;
  ctl->pull = 0;
  ctl->burst = 0;
  if (heavy_next) { /* this iteration relaxes the heavy arcs of the settled set */
    ctl->heavy_next = 0;
    ctl->r_pending = 0;
    ctl->heavy = 1;
    ctl->relax = tot.nvert > 0 ? 1 : 0;
    const unsigned long long done_arcs = ctl->heavy_done + tot.deg;
    ctl->heavy_done = done_arcs;
    if (tot.nvert > 0) {
      const unsigned long long ht = *sa.heavy_total;
      const unsigned long long pull_arcs = ht > done_arcs ? ht - done_arcs : 0ull;
      const bool pull = sa.pull_mode > 0 || (sa.pull_mode == 0 && pull_arcs < tot.deg);
      ctl->pull = pull ? 1 : 0;
      ctl->rounds += 1;
      ctl->relaxed += pull ? pull_arcs : tot.deg;
    }
  } else if (tot.nvert > 0) {
    ctl->heavy = 0;
    ctl->relax = 1;
    ctl->r_pending = 1;
    ctl->rounds += 1;
    ctl->relaxed += tot.deg;
    ctl->burst = (sa.bellman && (int64_t)tot.deg <= sa.burst_arcs) ? 1 : 0;
  } else {
    ctl->heavy = 0;
    ctl->relax = 0;
    if (ctl->r_pending && sa.delta > 0.0) {
      ctl->heavy_next = 1; /* the bucket's light phase is over */
    } else {
      unsigned long long far = 0;
      uint32_t fmin = 0xFFFFFFFFu;
      for (int k = 0; k < kSlots; ++k) {
        far += __hip_atomic_load(&sa.slots[k].far_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t mk = __hip_atomic_load(&sa.slots[k].far_min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fmin = mk < fmin ? mk : fmin;
      }
      if (far == 0) ctl->done = 1;
      else ctl->T_bits = next_threshold<W>(fmin, sa.delta, sa.w_max);
      ctl->r_pending = 0;
    }
  }
  for (int k = 0; k < kSlots; ++k) { /* the next iteration's split adds into cleared slots */
    sa.slots[k].far_count = 0;
    sa.slots[k].far_min = 0xFFFFFFFFu;
  }
  e.done = ctl->done;
  e.relax = ctl->relax;
  e.rounds = ctl->rounds;
  e.relaxed = ctl->relaxed;
  /* every field, then (after a system-scope fence) the iteration number the
   * host polls for */
  RoundLog* slot = &sa.log[iter % kRing];
  e.iter = -1;
  *slot = e;
  __threadfence_system();
  __hip_atomic_store(&slot->iter, iter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __threadfence_system();
}

template<typename OffT>
__global__ void sssp_init_kernel(const OffT* off, uint32_t* dist, uint32_t unreached, int64_t n, int32_t src,
                                 QEntry* q0, QEntry* q1, QEntry* q2, QueueHdr* hdr, Tuning t) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dist[i] = i == src ? 0u : unreached;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    QEntry* const q[3] = {q0, q1, q2};
    enqueue_single(off, t, q, hdr, src);
  }
}

__global__ void set_bit_kernel(uint64_t* bm, int32_t v) { bm[v >> 6] |= 1ull << (v & 63); }

template<typename W>
__global__ void weight_sum_kernel(const W* w, int64_t m, double* out) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)w[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane_id() == 0) atomicAdd(out, acc);
}

template<typename OffT>
__global__ void reached_kernel(const OffT* off, const uint32_t* dist, uint32_t unreached, int64_t n,
                               unsigned long long* out) {
  uint64_t e = 0, c = 0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    if (dist[v] != unreached) {
      e += (uint64_t)(off[v + 1] - off[v]);
      ++c;
    }
  /* one atomic pair per workgroup (not per wave) */
  __shared__ unsigned long long red[2][4];
  e = wave_sum_u64(e);
  c = wave_sum_u64(c);
  if (lane_id() == 0) {
    red[0][threadIdx.x >> 6] = e;
    red[1][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long te = 0, tc = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      te += red[0][w];
      tc += red[1][w];
    }
    atomicAdd(&out[0], te);
    atomicAdd(&out[1], tc);
  }
}

struct SsspWorkspace {
  /* dirty vertices, near set of the round, bucket's settled set, every vertex settled so far */
  uint64_t* bm[4] = {nullptr, nullptr, nullptr, nullptr};
  unsigned long long* acc = nullptr;
  SplitSlot* slots = nullptr;
  SsspCtl* ctl = nullptr;
  unsigned long long* relax_slots = nullptr; /* kSlots x 16 (one cache line each) */
  uint32_t* arrive = nullptr; /* arrival counters of the selection and write launches (kArriveWords each), zero
                                 between launches */
  QEntry* burst_list[2] = {nullptr, nullptr}; /* kSsspBurstMax entries each */
  RoundLog* log = nullptr; /* kRing entries, pinned coherent host memory */
  hipEvent_t ev[kRing] = {};
  hipEvent_t ev_pre[kRing] = {}; /* before each iteration's relax launch (record_relax) */
  QueueSet qs;
  int64_t w0 = 0;
};

void sssp_ws_free(SsspWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->bm[0]);
  (void)hipFree(w->bm[1]);
  (void)hipFree(w->bm[2]);
  (void)hipFree(w->bm[3]);
  (void)hipFree(w->acc);
  (void)hipFree(w->slots);
  (void)hipFree(w->ctl);
  (void)hipFree(w->relax_slots);
  (void)hipFree(w->arrive);
  (void)hipFree(w->burst_list[0]);
  (void)hipFree(w->burst_list[1]);
  if (w->log) (void)hipHostFree(w->log);
  for (hipEvent_t e : w->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : w->ev_pre)
    if (e) (void)hipEventDestroy(e);
  w->qs.release();
  delete w;
}

/* automatic delta = half the mean arc weight: of 0.125-1.0 times the mean,
 * the least geometric mean of time / the run's best (1.026) over the 30
 * training graphs the rule below sends to buckets (0.25: 1.18; 1.0: 1.19) */
constexpr double kAutoDeltaFactor = 0.5;

int sssp_workspace_for(sg_graph_t g, SsspWorkspace** out) {
  auto* ws = static_cast<SsspWorkspace*>(g->sssp_ws);
  if (!ws) {
    sg_device_t dev = g->dev;
    const int64_t n = g->n;
    ws = new SsspWorkspace();
    ws->w0 = (n + 63) / 64;
    int st = SG_OK;
    if ((st = sg::alloc_device(dev, (void**)&ws->bm[0], 8 * (ws->w0 + 1))) ||
        (st = sg::alloc_device(dev, (void**)&ws->bm[1], 8 * (ws->w0 + 1))) ||
        (st = sg::alloc_device(dev, (void**)&ws->bm[2], 8 * (ws->w0 + 1))) ||
        (st = sg::alloc_device(dev, (void**)&ws->bm[3], 8 * (ws->w0 + 1))) ||
        (st = sg::alloc_device(dev, (void**)&ws->acc, 16)) ||
        (st = sg::alloc_device(dev, (void**)&ws->slots, sizeof(SplitSlot) * kSlots)) ||
        (st = sg::alloc_device(dev, (void**)&ws->ctl, sizeof(SsspCtl))) ||
        (st = sg::alloc_device(dev, (void**)&ws->relax_slots, 8 * 16 * kSlots)) ||
        (st = sg::alloc_device(dev, (void**)&ws->arrive, 8 * kArriveWords)) ||
        (st = sg::zero_async(dev, ws->arrive, 8 * kArriveWords)) ||
        (st = sg::alloc_device(dev, (void**)&ws->burst_list[0], sizeof(QEntry) * kSsspBurstMax)) ||
        (st = sg::alloc_device(dev, (void**)&ws->burst_list[1], sizeof(QEntry) * kSsspBurstMax)) ||
        (st = ws->qs.init(dev, n, g->m, default_tuning()))) {
      sssp_ws_free(ws);
      return st;
    }
    bool ok = hipHostMalloc((void**)&ws->log, sizeof(RoundLog) * kRing, hipHostMallocCoherent) == hipSuccess;
    for (int r = 0; r < kRing && ok; ++r)
      ok = hipEventCreate(&ws->ev[r]) == hipSuccess && hipEventCreate(&ws->ev_pre[r]) == hipSuccess;
    if (!ok) {
      sssp_ws_free(ws);
      sg::set_error("sssp workspace: event or pinned log allocation failed");
      return SG_ERR_HIP;
    }
    g->sssp_ws = ws;
  }
  *out = ws;
  return SG_OK;
}

/* The light-first adjacency for bucket width dlt, built on first use and
 * kept with the graph until a run asks for another width. */
template<typename OffT, typename W>
int ensure_split(sg_graph_t g, W dlt) {
  const uint32_t bits = DistOps<W>::bits(dlt);
  if (g->sssp_split_valid && g->sssp_split_bits == bits) return SG_OK;
  sg_device_t dev = g->dev;
  if (!g->sssp_cols) {
    SG_TRY(sg::alloc_device(dev, (void**)&g->sssp_cols, 4 * (size_t)(g->m + 1)));
    SG_TRY(sg::alloc_device(dev, (void**)&g->sssp_w, 4 * (size_t)(g->m + 1)));
    SG_TRY(sg::alloc_device(dev, (void**)&g->sssp_light, 4 * (size_t)(g->n + 1)));
  }
  unsigned long long* heavy_total = nullptr; /* heavy arcs of the split, read by the heavy rounds' decision */
  SG_TRY(sg::graph_buffer(g, "sssp_heavy_total", 8, &heavy_total));
  SG_TRY(sg::zero_async(dev, heavy_total, 8));
  split_build_kernel<OffT, W><<<dev->cus * 8, kBlock, 0, dev->stream>>>(
      (const OffT*)g->offsets, g->cols, (const W*)g->weights, g->n, dlt, g->sssp_cols, (W*)g->sssp_w,
      g->sssp_light, heavy_total);
  SG_HIP(hipGetLastError());
  g->sssp_split_bits = bits;
  g->sssp_split_valid = true;
  return SG_OK;
}

template<typename OffT, typename W>
int run_sssp(sg_graph_t g, int32_t src, double delta, int32_t heavy_pull, int64_t burst_arcs, bool record_relax,
             W* dist_out, sg_sssp_stats* stats) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n;
  const Tuning tune = default_tuning();
  SsspWorkspace* ws = nullptr;
  SG_TRY(sssp_workspace_for(g, &ws));
  QueueSet& Q = ws->qs;
  const OffT* off = (const OffT*)g->offsets;
  uint32_t* dist = reinterpret_cast<uint32_t*>(dist_out);
  const uint32_t unreached = DistOps<W>::bits((W)(n + 1));
  const int grid_main = dev->cus * 8;

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, s));
  }
  uint64_t* dirty = ws->bm[0];
  uint64_t* near_bm = ws->bm[1];
  uint64_t* settled = ws->bm[2];
  uint64_t* settled_all = ws->bm[3];
  SG_TRY(sg::zero_async(dev, dirty, 8 * ws->w0));
  SG_TRY(sg::zero_async(dev, settled, 8 * ws->w0));
  /* vertices without arcs never change: a pull skips them from the start */
  const bool directed = (g->flags & SG_GRAPH_DIRECTED) != 0;
  if (!directed) {
    SG_TRY(sg::graph_ensure_isolated(g));
    SG_HIP(hipMemcpyAsync(settled_all, g->isolated, 8 * ws->w0, hipMemcpyDeviceToDevice, s));
  } else {
    SG_TRY(sg::zero_async(dev, settled_all, 8 * ws->w0));
  }
  SG_TRY(sg::zero_async(dev, ws->relax_slots, 8 * 16 * kSlots));
  sssp_init_kernel<OffT><<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(off, dist, unreached, n, src, Q.q[0][0],
                                                                       Q.q[0][1], Q.q[0][2], Q.hdr, tune);
  set_bit_kernel<<<1, 1, 0, s>>>(dirty, src);
  SG_HIP(hipGetLastError());
  const bool bellman = !(delta > 0.0) || !std::isfinite(delta);
  const double w_max = DistOps<W>::is_float ? 3.0e38 : 4294967295.0;
  {
    SsspCtl c{};
    c.T_bits = DistOps<W>::bits((W)(bellman ? w_max : std::min(delta, w_max)));
    SG_HIP(hipMemcpyAsync(ws->ctl, &c, sizeof(c), hipMemcpyHostToDevice, s));
  }
  /* arcs of weight >= dlt are heavy (none under the reference's schedule) */
  const W dlt = bellman ? (W)w_max : (W)std::min(std::max(delta, DistOps<W>::is_float ? 0.0 : 1.0), w_max);
  /* bucketed runs read the light-first adjacency, the entries covering one
   * phase's part of each list */
  if (!bellman) SG_TRY((ensure_split<OffT, W>(g, dlt)));
  CompactArgs<OffT> ca = make_compact_args<OffT>(Q, 0, off, nullptr, nullptr, near_bm, tune);
  ca.split = bellman ? nullptr : g->sssp_light;
  SsspArgs<OffT, W> a{};
  a.cols = bellman ? g->cols : g->sssp_cols;
  a.w = bellman ? (const W*)g->weights : (const W*)g->sssp_w;
  a.dist = dist;
  a.next_bm = dirty;
  a.w0 = ws->w0;
  for (int k = 0; k < 3; ++k) a.qin[k] = Q.q[0][k];
  a.qin_hdr = Q.hdr;
  a.off = off;
  a.light = g->sssp_light;
  a.settled = settled_all;
  a.frontier = near_bm;
  a.n = n;
  /* a heavy round pulls only on an undirected graph (in-arcs = out-arcs, symmetric weights) */
  const int32_t pull_mode = directed ? -1 : heavy_pull;
  unsigned long long* heavy_total = nullptr;
  if (!bellman) SG_TRY(sg::graph_buffer(g, "sssp_heavy_total", 8, &heavy_total));
  SelectArgs sa{};
  sa.dirty = dirty;
  sa.near_bm = near_bm;
  sa.settled = settled;
  sa.settled_all = settled_all;
  sa.dist = dist;
  sa.slots = ws->slots;
  sa.w0 = ws->w0;
  sa.nblocks = Q.nblocks;
  sa.arrive = ws->arrive;
  sa.log = ws->log;
  sa.heavy_total = heavy_total;
  sa.delta = bellman ? 0.0 : delta;
  sa.w_max = w_max;
  sa.pull_mode = pull_mode;
  sa.bellman = bellman ? 1 : 0;
  sa.burst_arcs = burst_arcs;
  SsspBurst<OffT, W> sb{};
  sb.off = off;
  sb.cols = g->cols;
  sb.w = (const W*)g->weights;
  sb.dist = dist;
  sb.dirty = dirty;
  for (int k = 0; k < 3; ++k) sb.q[k] = Q.q[0][k];
  sb.hdr = Q.hdr;
  sb.list[0] = ws->burst_list[0];
  sb.list[1] = ws->burst_list[1];
  sb.burst_arcs = burst_arcs;
  sb.relax_slots = ws->relax_slots;
  sb.arrive = ws->arrive + kArriveWords;
  split_slots_reset_kernel<<<1, 64, 0, s>>>(ws->slots, ws->ctl); /* later iterations clear them as they decide */
  /* one iteration, three launches: split + near-set count + decision (relax /
   * next bucket / stop); queue write; relax.  Everything the host needs is in
   * the iteration's record. */
  auto enqueue = [&](int32_t it) -> int {
    sssp_select_kernel<OffT, W><<<Q.grid_tiles(), kBlock, 0, s>>>(ca, sa, ws->ctl, it);
    near_write_kernel<OffT, W><<<Q.grid_tiles(), kBlock, 0, s>>>(ca, sb, ws->ctl);
    if (record_relax) SG_HIP(hipEventRecord(ws->ev_pre[it % kRing], s));
    relax_kernel<OffT, W><<<grid_main, kBlock, 0, s>>>(a, ws->ctl, ws->relax_slots);
    SG_HIP(hipGetLastError());
    /* untimed iterations record no event (a stream barrier per iteration):
     * the host polls the iteration's record, written by its selection */
    if (record_relax) SG_HIP(hipEventRecord(ws->ev[it % kRing], s));
    return SG_OK;
  };
  for (int k = 0; k < kRing; ++k) __atomic_store_n(&ws->log[k].iter, -1, __ATOMIC_RELEASE);
  SG_TRY(enqueue(0));
  SG_TRY(enqueue(1));
  int32_t round = 0;
  int64_t relaxed = 0;
  float relax_ms = 0.f;
  int32_t relax_launches = 0;
  const int64_t max_iter = 4 * n + 256; /* every iteration relaxes or advances T past a distance */
  for (int32_t it = 0;; ++it) {
    if (record_relax) SG_HIP(hipEventSynchronize(ws->ev[it % kRing]));
    else SG_TRY(sg::wait_host_flag(&ws->log[it % kRing].iter, it, s, "sg_sssp"));
    RoundLog e;
    std::memcpy(&e, (const void*)&ws->log[it % kRing], sizeof(e));
    if (e.iter != it) {
      SG_HIP(hipStreamSynchronize(s));
      sg::set_error("sg_sssp: iteration record %d out of order (found %d)", it, e.iter);
      return SG_ERR_STATE;
    }
    if (e.relax && stats && e.rounds - 1 < SG_MAX_LEVELS_RECORDED) {
      stats->round_vertices[e.rounds - 1] = e.near_v;
      stats->round_edges[e.rounds - 1] = e.near_e;
    }
    if (e.relax && record_relax) { /* the relax launch of a relaxing iteration, on its own */
      float ms = 0.f;
      SG_HIP(hipEventElapsedTime(&ms, ws->ev_pre[it % kRing], ws->ev[it % kRing]));
      relax_ms += ms;
      ++relax_launches;
    }
    round = e.rounds;
    relaxed = (int64_t)e.relaxed;
    if (e.done) break;
    if (it + 2 > max_iter) {
      SG_HIP(hipStreamSynchronize(s));
      sg::set_error("sg_sssp: no termination after %lld iterations", (long long)it);
      return SG_ERR_STATE;
    }
    SG_TRY(enqueue(it + 2));
  }
  if (stats) {
    SG_HIP(hipEventRecord(e1, s));
    SG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    stats->kernel_ms = ms;
    stats->rounds = round;
    stats->relax_ms = relax_ms;
    stats->relax_launches = relax_launches;
    {
      unsigned long long rs[16 * kSlots];
      SG_HIP(hipMemcpyAsync(rs, ws->relax_slots, sizeof(rs), hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      unsigned long long t = 0;
      for (int k = 0; k < kSlots; ++k) t += rs[k * 16];
      stats->relaxations = (int64_t)t; /* arcs relaxed in their phase (examined: `relaxed`) */
      (void)relaxed;
    }
    SG_TRY(sg::zero_async(dev, ws->acc, 16));
    reached_kernel<OffT><<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(off, dist, unreached, n, ws->acc);
    unsigned long long h[2];
    SG_HIP(hipMemcpyAsync(h, ws->acc, 16, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    stats->reached_edges = (int64_t)h[0];
    stats->reached_vertices = (int64_t)h[1];
  }
  SG_HIP(hipStreamSynchronize(s));
  return SG_OK;
}

} // namespace

namespace sg {
void sssp_workspace_release(sg_graph_t g) {
  if (g && g->sssp_ws) {
    sssp_ws_free(static_cast<SsspWorkspace*>(g->sssp_ws));
    g->sssp_ws = nullptr;
  }
}
} // namespace sg

/* mean arc weight, computed once per graph (automatic delta) */
template<typename W>
int mean_weight(sg_graph_t g, double* out) {
  if (g->mean_weight >= 0.0) {
    *out = g->mean_weight;
    return SG_OK;
  }
  SsspWorkspace* ws = nullptr;
  SG_TRY(sssp_workspace_for(g, &ws));
  double* acc = reinterpret_cast<double*>(ws->acc);
  SG_TRY(sg::zero_async(g->dev, acc, 16));
  weight_sum_kernel<W><<<sg::grid_for(g->dev, g->m, 256, 4), 256, 0, g->dev->stream>>>((const W*)g->weights, g->m,
                                                                                       acc);
  double h = 0.0;
  SG_HIP(hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, g->dev->stream));
  SG_HIP(hipStreamSynchronize(g->dev->stream));
  g->mean_weight = g->m > 0 ? h / (double)g->m : 1.0;
  *out = g->mean_weight;
  return SG_OK;
}

/* The automatic schedule: buckets of half the mean arc weight on a graph of
 * at least kAutoBucketArcs arcs and kAutoBucketDegree arcs per vertex, the
 * reference's frontier Bellman-Ford otherwise.  Below the arc count an
 * iteration's fixed cost (its launches) outweighs the relaxations that buckets
 * save; on sparse, long-diameter graphs (grids: 3-4 arcs per vertex) buckets
 * multiply the iterations, each near that fixed cost (3-4x slower there).  The
 * pair has the least geometric mean of time / the faster schedule's time over
 * 100 generated graphs (RMAT scale 16-22, grids 512^2-4096^2, ER; weights
 * U[1,255]; four sources each) among arc counts 2^18..2^28 and degrees 0..32:
 * 3.4 % over choosing per run (18 % for the reference's schedule everywhere),
 * total time 1.8 % under the reference's schedule everywhere. */
constexpr int64_t kAutoBucketArcs = int64_t(1) << 25;
constexpr int64_t kAutoBucketDegree = 8;

template<typename OffT, typename W>
int run_sssp_auto(sg_graph_t g, int32_t source, double delta, int32_t heavy_pull, int64_t burst_arcs,
                  bool record_relax, W* dist, sg_sssp_stats* stats) {
  if (delta == 0.0 && (g->m < kAutoBucketArcs || g->m < kAutoBucketDegree * (int64_t)g->n)) {
    delta = -1.0;
  } else if (delta == 0.0) {
    double mw = 1.0;
    SG_TRY(mean_weight<W>(g, &mw));
    delta = std::max(kAutoDeltaFactor * mw, DistOps<W>::is_float ? 1e-30 : 1.0);
  }
  return run_sssp<OffT, W>(g, source, delta, heavy_pull, burst_arcs, record_relax, dist, stats);
}

extern "C" {

void sg_sssp_options_default(sg_sssp_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->load_balancer = SG_LB_BUCKETING;
  opt->delta = 0.0;
}

int sg_sssp_ex(sg_graph_t g, int32_t source, const sg_sssp_options* opt_in, void* dist, sg_sssp_stats* stats) {
  SG_CHECK(g && dist, SG_ERR_INVALID, "sg_sssp: null argument");
  SG_CHECK(source >= 0 && source < g->n, SG_ERR_INVALID, "sg_sssp: source %d out of range [0,%lld)", source,
           (long long)g->n);
  SG_CHECK(g->weights, SG_ERR_STATE, "sg_sssp: graph has no edge weights");
  sg_sssp_options opt;
  if (opt_in) opt = *opt_in;
  else sg_sssp_options_default(&opt);
  SG_CHECK(opt.load_balancer >= SG_LB_WORKITEM && opt.load_balancer <= SG_LB_BUCKETING, SG_ERR_INVALID,
           "sg_sssp: unknown load balancer %d", opt.load_balancer);
  SG_CHECK(opt.heavy_pull >= -1 && opt.heavy_pull <= 1, SG_ERR_INVALID, "sg_sssp: heavy_pull %d not in {-1, 0, 1}",
           opt.heavy_pull);
  SG_CHECK(opt.burst_arcs <= kSsspBurstMax, SG_ERR_INVALID, "sg_sssp: burst_arcs %d above %d", opt.burst_arcs,
           kSsspBurstMax);
  if (stats) {
    stats->rounds = 0;
    stats->reached_vertices = stats->reached_edges = stats->relaxations = 0;
    stats->kernel_ms = 0.f;
    stats->relax_ms = 0.f;
    stats->relax_launches = 0;
  }
  SG_HIP(hipSetDevice(g->dev->ordinal));
  const bool fw = (g->flags & SG_GRAPH_WEIGHT_FLOAT) != 0;
  const int64_t ba = opt.burst_arcs < 0 ? 0 : (opt.burst_arcs == 0 ? kSsspBurstDefault : opt.burst_arcs);
  const bool rr = opt.record_relax != 0 && stats != nullptr;
  if (g->offset_bytes == 8)
    return fw ? run_sssp_auto<int64_t, float>(g, source, opt.delta, opt.heavy_pull, ba, rr, (float*)dist, stats)
              : run_sssp_auto<int64_t, uint32_t>(g, source, opt.delta, opt.heavy_pull, ba, rr, (uint32_t*)dist,
                                                 stats);
  return fw ? run_sssp_auto<uint32_t, float>(g, source, opt.delta, opt.heavy_pull, ba, rr, (float*)dist, stats)
            : run_sssp_auto<uint32_t, uint32_t>(g, source, opt.delta, opt.heavy_pull, ba, rr, (uint32_t*)dist, stats);
}

int sg_sssp(sg_graph_t g, int32_t source, int32_t load_balancer, void* dist, sg_sssp_stats* stats) {
  sg_sssp_options opt;
  sg_sssp_options_default(&opt);
  opt.load_balancer = load_balancer;
  return sg_sssp_ex(g, source, &opt, dist, stats);
}

} // extern "C"
