/*
 * msbfs.hip -- multi-source BFS and batched betweenness centrality, 64
 * sources per batch (one bit-lane each), the whole level loop of a batch in
 * one cooperative launch.
 *
 * State per vertex: vis (lanes that reached it) and four rotating level masks
 * M[d & 3] (lanes for which it lies at depth d).  Each level keeps a list of
 * its vertices (each once, whatever number of lanes), so no pass ever sweeps
 * all vertices on a sparse level: a 4096 x 4096 grid runs its ~8,000 levels
 * at the cost of its frontiers plus one grid barrier per level, not of a
 * launch or a host round trip per level.
 *
 * Level d, one phase between grid barriers:
 *  - expand: push over the out-arcs of L_d (lane per arc), or pull over the
 *    in-arcs of every vertex still missing lanes when L_d holds more than
 *    1/20 of the vertices; for an arc u -> v the new lanes are
 *    M_d[u] & ~(vis[v] | M_d[v]), ORed into M_{d+1}[v].  The first lane to
 *    touch v appends it to L_{d+1}.  Adjacency lists longer than kHub arcs
 *    are split into chunks of kHub arcs, listed when the vertex is appended;
 *  - visit L_d: vis |= M_d, depth rows (MS-BFS) or the level's masks kept
 *    for the backward pass (BC), and the masks of L_{d-2} cleared so that
 *    the array serves as M_{d+2};
 *  - BC: path counts of L_d, pulled from in-neighbours at depth d - 1.
 * Depths do not depend on push or pull (the level at which a lane first
 * reaches a vertex is fixed), so they are bit-exact with one BFS per source.
 *
 * Brandes (algorithms/bc.hpp:118-200), lane l = source l:
 *   sigma_l(v) = sum of sigma_l(u) over in-arcs u -> v with u at depth d-1;
 *   delta_l(v) = sum over out-arcs v -> w with w at depth d+1 of
 *                sigma_l(v) / sigma_l(w) * (1 + delta_l(w));
 *   bc(v) += sum of delta_l(v) over the lanes with v at depth >= 1.
 * Path counts grow like 2^depth on a grid (~2^8190 on 4096 x 4096) and
 * differ by thousands of binary orders within one level, so each (vertex,
 * lane) keeps sigma as an FP32 mantissa in [0.5, 1) and an int16 exponent;
 * sums align exponents as they go.  The backward pass overwrites sigma with
 * R = (1 + delta) / sigma in the same form, so an arc's term is one product
 * and one ldexp; delta itself is at most the vertex count and stays FP32.
 * The backward pass walks the levels deepest first; membership of level d+1
 * is rebuilt from its list into one of two mask arrays (the OR of L_d's masks
 * and the AND-NOT of L_{d+2}'s act on disjoint lanes, so they commute).
 *
 * Work mapping of the sigma and delta passes: a wave takes 64 list entries
 * and walks their (vertex, lane) pairs -- 64 pairs per step from several
 * vertices when masks are sparse (grids), a vertex per step with lane =
 * source when they are dense (power-law graphs).  Each pair sums its own
 * arcs in CSR order; lists longer than kHub are summed per chunk into
 * partial rows that a following phase adds in chunk order.  Nothing is
 * summed with atomics, so the results do not depend on scheduling.
 */
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "policy.hpp"
#include "traverse.hpp"

using namespace sg::trav;

namespace {

/* Data one workgroup writes and another reads later in the same launch (the
 * masks, level lists, chunk records, path-count and dependency rows, partial
 * rows, bc sums) goes to and comes from the point of coherence that every
 * XCD shares -- agent-scope relaxed atomic loads and stores -- so the grid
 * barrier between phases neither writes back nor invalidates any L2 (with
 * fences, a 4,096-level batch on a grid spent ~100 us per barrier). */
template<typename T>
__device__ __forceinline__ T cld(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename T, typename U>
__device__ __forceinline__ void cst(T* p, U v) {
  __hip_atomic_store(p, static_cast<T>(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template<typename S>
__device__ __forceinline__ void cst16(S* p, const S& v) {
  static_assert(sizeof(S) == 16, "16-byte records");
  uint64_t w[2];
  __builtin_memcpy(w, &v, 16);
  cst(reinterpret_cast<uint64_t*>(p), w[0]);
  cst(reinterpret_cast<uint64_t*>(p) + 1, w[1]);
}
template<typename S>
__device__ __forceinline__ S cld16(const S* p) {
  static_assert(sizeof(S) == 16, "16-byte records");
  uint64_t w[2];
  w[0] = cld(reinterpret_cast<const uint64_t*>(p));
  w[1] = cld(reinterpret_cast<const uint64_t*>(p) + 1);
  S v;
  __builtin_memcpy(&v, w, 16);
  return v;
}

constexpr int kHub = 256;     /* adjacency lists longer than this are chunked */
constexpr int kLiRing = 8;    /* level records of the MS-BFS ring */
constexpr int kBarGroups = 8; /* arrival counters, one per XCD */
constexpr int kBarStride = 32; /* 128 bytes apart */
constexpr uint32_t kOvList = 1, kOvHubV = 2, kOvHubC = 4, kOvPart = 8, kOvLevels = 16;
constexpr uint32_t kTimeout = 32, kRange = 64;
constexpr uint32_t kStop = kOvList | kOvHubV | kOvHubC | kOvPart | kOvLevels | kTimeout | kRange;
constexpr int kExpMax = 32000; /* |exponent| of sigma and R kept in int16 */

/* One level: its vertices and the chunks of its long lists (out-arcs; in-arcs
 * as well on directed graphs).  Counts grow with atomics while the level is
 * being discovered; bases are positions in the list arrays. */
struct alignas(128) LevelInfo {
  unsigned long long count, arcs, base;
  unsigned long long hov, hov_base, hoc, hoc_base; /* out-hubs: vertices, chunks */
  unsigned long long hiv, hiv_base, hic, hic_base; /* in-hubs (directed BC) */
  unsigned long long mx;                           /* largest out-degree among the level's vertices */
  unsigned long long pad[4];
};
constexpr int kLevelInfoZero = 12; /* leading words of a LevelInfo cleared before the level fills */

/* a vertex with a long list: its chunks are [first, first + nch) of the level's chunk list */
struct HubV {
  int32_t v, nch, first, pos; /* pos: the vertex's index in its level list */
};
struct HubC {
  int32_t v, c, pos, pad;
};

struct Ctl {
  uint32_t grp[kBarGroups * kBarStride];
  uint32_t top[kBarStride];
  uint32_t gen[kBarStride];
  uint32_t flags;                /* polled by every waiting workgroup */
  int32_t levels;                /* levels of the batch (depths 0 .. levels-1) */
  int32_t n_pull;                /* levels expanded by pulling */
  uint32_t pad0[kBarStride - 3];
  /* on a line of its own: atomics on the polled flags line queue behind the polls */
  unsigned long long full_deg;   /* out-degrees of the vertices every lane has reached */
  uint32_t pad1[kBarStride - 2];
  /* a level's arcs and largest degree, added by the workgroups that discover
   * it into eight lines (workgroup % 8) instead of the list counter's line,
   * three rotating copies (level % 3) */
  struct alignas(128) LevelAcc {
    unsigned long long arcs, mx, pad[14];
  } acc[3][kBarGroups];
};
static_assert(sizeof(Ctl) % 8 == 0, "Ctl is cleared in 4-byte words");

template<typename OffT>
struct MsArgs {
  int64_t n;
  uint64_t full;
  int32_t k;        /* sources in the batch */
  int32_t bc;       /* 1: batched BC, 0: MS-BFS */
  int32_t ring;     /* 1: level lists in a ring of four slots (MS-BFS) */
  int32_t directed; /* in-arcs are not the out-arcs */
  const OffT* off;
  const int32_t* cols;
  const OffT* ioff;
  const int32_t* icols;
  uint64_t* vis;
  uint64_t* M; /* four arrays of n words */
  int32_t* lv;
  uint64_t* lm;     /* BC: each list entry's mask */
  int64_t lcap;     /* list entries (ring: per slot) */
  HubV* hov;
  HubC* hoc;
  HubV* hiv;
  HubC* hic;
  int64_t hv_cap, hc_cap; /* per slot in the ring */
  LevelInfo* li;
  int64_t li_cap;
  int32_t* depth; /* MS-BFS: rows of the batch's sources */
  float* sm;      /* BC: sigma / R mantissas, [n][64] */
  int16_t* se;    /* BC: exponents, [n][64] */
  float* dl;      /* BC: delta, [n][64] */
  float* pS;      /* BC: partial rows of chunked lists, [p_cap][64] */
  int32_t* pE;
  int64_t p_cap;
  double* bc64;
  const int32_t* src;
  Ctl* ctl;
  uint64_t wait_ticks;
  /* direction of each level's expansion: -1 pull when the level holds more
   * than n/20 vertices; SG_POLICY_PUSH / _PULL always; SG_POLICY_HYBRID the
   * reference's alpha / beta rule (algorithms/bfs.hpp:341-350) on the batch's
   * totals; SG_POLICY_LEARNED the learned policy on the same totals */
  int32_t dir_policy;
  float alpha, beta;
  const float* mlp;
  float f6, f7;
  unsigned long long total_degree;
};

/* --------------------------------------------------------- device helpers --- */
__device__ __forceinline__ uint32_t load_flags(const Ctl* c) { return atomic_load(&c->flags); }
__device__ __forceinline__ void raise_flag(Ctl* c, uint32_t f) { atomic_or(&c->flags, f); }

/* Grid barrier over a launch whose workgroups are all resident.  Arrival
 * counters only grow (barrier t completes at t arrivals per member), so
 * nothing is reset between barriers.  No fences: everything the workgroups
 * exchange goes through agent-coherent loads and stores (cld / cst and the
 * atomics), so a wave only waits for its own stores to complete before its
 * workgroup arrives.  The wait is bounded: past wait_ticks of the
 * constant-rate clock the kTimeout flag is raised and every workgroup leaves.
 * Returns false when the run must stop (a flag is set). */
__device__ __noinline__ bool grid_sync(Ctl* c, uint32_t& t, uint64_t wait_ticks) {
  __shared__ uint32_t s_stop;
  wait_stores_complete();
  __syncthreads();
  ++t;
  if (threadIdx.x == 0) {
    const uint32_t grid = gridDim.x, g = blockIdx.x % kBarGroups;
    const uint32_t ngroups = grid < (uint32_t)kBarGroups ? grid : (uint32_t)kBarGroups;
    const uint32_t in_group = (grid - g + kBarGroups - 1) / kBarGroups;
    bool released = false;
    if (atomic_add(&c->grp[g * kBarStride], 1u) == t * in_group - 1) {
      if (atomic_add(&c->top[0], 1u) == t * ngroups - 1) {
        atomic_store(&c->gen[0], t);
        released = true;
      }
    }
    if (!released) {
      const uint64_t t0 = wall_clock64();
      while (atomic_load(&c->gen[0]) < t) {
        if (load_flags(c) & kStop) break;
        if ((uint64_t)(wall_clock64() - t0) > wait_ticks) {
          raise_flag(c, kTimeout);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    s_stop = (load_flags(c) & kStop) ? 1u : 0u;
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();
  return s_stop == 0;
}

/* this lane's first slot of k, with one atomic per wave (all lanes call) */
__device__ __forceinline__ uint64_t wave_reserve(unsigned long long* ctr, uint32_t k) {
  const uint32_t incl = wave_inclusive_scan(k);
  const uint32_t total = __shfl(incl, 63, 64);
  unsigned long long base = 0;
  if (total) {
    if (lane_id() == 63) base = atomic_add(ctr, (unsigned long long)total);
    base = (unsigned long long)shfl_i64((int64_t)base, 63);
  }
  return base + incl - k;
}

template<typename OffT>
__device__ __forceinline__ LevelInfo* level(const MsArgs<OffT>& a, int32_t d) {
  return a.ring ? a.li + (d & (kLiRing - 1)) : a.li + d;
}

template<typename OffT>
__device__ __forceinline__ int64_t list_base(const MsArgs<OffT>& a, int32_t d) {
  return a.ring ? (int64_t)(d & 3) * a.lcap : (int64_t)atomic_load(&level(a, d)->base);
}

/* Bases of level d+1's lists: the ring's slot, or right after level d's. */
struct Bases {
  int64_t l, hov, hoc, hiv, hic;
};
template<typename OffT>
__device__ __forceinline__ Bases next_bases(const MsArgs<OffT>& a, int32_t d) {
  Bases b;
  if (a.ring) {
    const int64_t s = (d + 1) & 3;
    b.l = s * a.lcap;
    b.hov = b.hiv = s * a.hv_cap;
    b.hoc = b.hic = s * a.hc_cap;
  } else {
    LevelInfo* L = level(a, d);
    b.l = (int64_t)(atomic_load(&L->base) + atomic_load(&L->count));
    b.hov = (int64_t)(atomic_load(&L->hov_base) + atomic_load(&L->hov));
    b.hoc = (int64_t)(atomic_load(&L->hoc_base) + atomic_load(&L->hoc));
    b.hiv = (int64_t)(atomic_load(&L->hiv_base) + atomic_load(&L->hiv));
    b.hic = (int64_t)(atomic_load(&L->hic_base) + atomic_load(&L->hic));
  }
  return b;
}

/* Write the entries whose list positions `rel` (relative to the level's
 * base) are known: the vertex id and, for lists longer than kHub arcs, the
 * hub and chunk records.  All lanes call; lanes with ok == false only take
 * part in the wave's reservations. */
template<typename OffT>
__device__ __forceinline__ void place(const MsArgs<OffT>& a, LevelInfo* nx, const Bases& b, bool ok, int32_t v,
                                     uint64_t rel) {
  const int64_t cap_l = a.ring ? a.lcap : a.lcap - b.l;
  const int64_t cap_hv = a.ring ? a.hv_cap : a.hv_cap - b.hov;
  const int64_t cap_hc = a.ring ? a.hc_cap : a.hc_cap - b.hoc;
  int64_t deg = 0;
  if (ok) {
    if ((int64_t)rel < cap_l) cst(&a.lv[b.l + (int64_t)rel], v);
    else raise_flag(a.ctl, kOvList);
    deg = (int64_t)(a.off[v + 1] - a.off[v]);
  }
  uint32_t nch = (ok && deg > kHub) ? (uint32_t)((deg + kHub - 1) / kHub) : 0u;
  if (ballot(nch != 0)) {
    const uint64_t hv = wave_reserve(&nx->hov, nch ? 1u : 0u);
    const uint64_t hc = wave_reserve(&nx->hoc, nch);
    if (nch) {
      if ((int64_t)hv < cap_hv && (int64_t)(hc + nch) <= cap_hc) {
        cst16(&a.hov[b.hov + (int64_t)hv], HubV{v, (int32_t)nch, (int32_t)hc, (int32_t)rel});
        for (uint32_t c = 0; c < nch; ++c) cst16(&a.hoc[b.hoc + (int64_t)hc + c], HubC{v, (int32_t)c, (int32_t)rel, 0});
      } else {
        raise_flag(a.ctl, (int64_t)hv < cap_hv ? kOvHubC : kOvHubV);
      }
    }
  }
  if (a.bc && a.directed) { /* in-arc chunks for the path counts */
    const int64_t ideg = ok ? (int64_t)(a.ioff[v + 1] - a.ioff[v]) : 0;
    nch = ideg > kHub ? (uint32_t)((ideg + kHub - 1) / kHub) : 0u;
    if (ballot(nch != 0)) {
      const int64_t cap_iv = a.hv_cap - b.hiv, cap_ic = a.hc_cap - b.hic;
      const uint64_t hv = wave_reserve(&nx->hiv, nch ? 1u : 0u);
      const uint64_t hc = wave_reserve(&nx->hic, nch);
      if (nch) {
        if ((int64_t)hv < cap_iv && (int64_t)(hc + nch) <= cap_ic) {
          cst16(&a.hiv[b.hiv + (int64_t)hv], HubV{v, (int32_t)nch, (int32_t)hc, (int32_t)rel});
          for (uint32_t c = 0; c < nch; ++c) cst16(&a.hic[b.hic + (int64_t)hc + c], HubC{v, (int32_t)c, (int32_t)rel, 0});
        } else {
          raise_flag(a.ctl, (int64_t)hv < cap_iv ? kOvHubC : kOvHubV);
        }
      }
    }
  }
}

/* Append the vertices this wave touched first (first == true) to level
 * nx's list, with their out-degrees and, for long lists, their chunks.
 * All lanes call.  One atomic per wave and call: the level-0 seeds only; the
 * expansions stage their discoveries (Stage below). */
template<typename OffT>
__device__ __forceinline__ void append(const MsArgs<OffT>& a, LevelInfo* nx, const Bases& b, bool first,
                                       int32_t v) {
  const uint64_t any = ballot(first);
  if (!any) return;
  const uint64_t rel = wave_reserve(&nx->count, first ? 1u : 0u);
  place(a, nx, b, first, v, rel);
  const uint64_t deg = first ? (uint64_t)(a.off[v + 1] - a.off[v]) : 0ull;
  const uint64_t arcs = wave_sum_u64(deg), mx = wave_max_u64(deg);
  if (lane_id() == 0 && arcs) {
    atomic_add(&nx->arcs, (unsigned long long)arcs);
    atomic_max(&nx->mx, (unsigned long long)mx);
  }
}

/* A wave's discoveries of one phase, staged in LDS and written to the level
 * list kStage at a time with one atomic, so that a dense level (millions of
 * discoveries) does not serialize on its list counter; the level's arc count
 * and largest degree are summed per workgroup and added once. */
constexpr uint32_t kStage = 448;
struct Stage {
  int32_t* buf; /* this wave's LDS region, kStage ids */
  uint32_t n;   /* staged ids (the same in every lane) */
  uint64_t arcs, mx; /* per lane */
};

/* Place the wave's staged ids at list positions base, base + 1, ... */
template<typename OffT>
__device__ __forceinline__ void stage_place(const MsArgs<OffT>& a, LevelInfo* nx, const Bases& b, Stage& st,
                                            unsigned long long base) {
  const int lane = lane_id();
  for (uint32_t i0 = 0; i0 < st.n; i0 += 64) {
    const uint32_t i = i0 + lane;
    const bool ok = i < st.n;
    const int32_t v = ok ? st.buf[i] : 0;
    if (ok) {
      const uint64_t deg = (uint64_t)(a.off[v + 1] - a.off[v]);
      st.arcs += deg;
      st.mx = deg > st.mx ? deg : st.mx;
    }
    place(a, nx, b, ok, v, base + i);
  }
  st.n = 0;
}

/* A full stage in the middle of a phase: one reservation for the wave. */
template<typename OffT>
__device__ __forceinline__ void stage_flush(const MsArgs<OffT>& a, LevelInfo* nx, const Bases& b, Stage& st) {
  if (st.n == 0) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomic_add(&nx->count, (unsigned long long)st.n);
  base = (unsigned long long)shfl_i64((int64_t)base, 0);
  stage_place(a, nx, b, st, base);
}

/* Stage the lanes' first-touched vertices (all lanes call). */
template<typename OffT>
__device__ __forceinline__ void stage_add(const MsArgs<OffT>& a, LevelInfo* nx, const Bases& b, Stage& st, bool first,
                                          int32_t v) {
  const uint64_t m = ballot(first);
  if (!m) return;
  const uint32_t c = (uint32_t)__popcll(m);
  if (st.n + c > kStage) stage_flush(a, nx, b, st);
  if (first) st.buf[st.n + lanes_below(m)] = v;
  st.n += c;
}

/* End of a phase: the workgroup's waves place what they still stage under
 * one reservation (one atomic on the list counter per workgroup, not per
 * wave: on a level of a few hundred thousand entries spread over every wave
 * those atomics queue on one line), then one add (arcs) and one max (largest
 * degree) per workgroup.  Every thread of the workgroup calls. */
template<typename OffT>
__device__ __forceinline__ void stage_finish(const MsArgs<OffT>& a, LevelInfo* nx, const Bases& b, Stage& st,
                                             unsigned long long* s_red, Ctl::LevelAcc* acc) {
  __shared__ unsigned long long s_at[kBlock / 64];
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 0) s_at[wv] = st.n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      const unsigned long long c = s_at[w];
      s_at[w] = tot;
      tot += c;
    }
    const unsigned long long base = tot ? atomic_add(&nx->count, tot) : 0ull;
    for (int w = 0; w < kBlock / 64; ++w) s_at[w] += base;
  }
  __syncthreads();
  if (st.n) stage_place(a, nx, b, st, s_at[wv]);
  const uint64_t arcs = wave_sum_u64(st.arcs), mx = wave_max_u64(st.mx);
  if (lane_id() == 0) {
    s_red[2 * wv] = arcs;
    s_red[2 * wv + 1] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0, x = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      t += s_red[2 * w];
      x = s_red[2 * w + 1] > x ? s_red[2 * w + 1] : x;
    }
    Ctl::LevelAcc* my = acc + (blockIdx.x % kBarGroups);
    if (t) atomic_add(&my->arcs, t);
    if (x) atomic_max(&my->mx, x);
  }
  __syncthreads();
  st.arcs = st.mx = 0;
}

/* New lanes for a lane's four heads v[k] (< 0: none) from frontier masks
 * fm[k]: the lanes not in vis or M_d are ORed into M_{d+1}[v]; first[k]: this
 * OR was the first to give v[k] any lane at depth d+1.  The loads of the four
 * arcs are in flight together, one round trip per stage: vis and M_d of every
 * head, then M_{d+1} of the heads that would gain lanes (most arcs of a
 * power-law graph stop at the first stage), then the ORs that still add. */
__device__ __forceinline__ void relax4(const uint64_t (&fm)[4], const int32_t (&v)[4], const uint64_t* vis,
                                       const uint64_t* Mc, uint64_t* Mn, bool (&first)[4]) {
  uint64_t x[4], y[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = v[k] >= 0 ? cld(&vis[v[k]]) : ~0ull;
    y[k] = v[k] >= 0 ? cld(&Mc[v[k]]) : ~0ull;
  }
  uint64_t add[4], z[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    add[k] = fm[k] & ~(x[k] | y[k]);
    z[k] = add[k] ? cld(&Mn[v[k]]) : ~0ull;
  }
  uint64_t old[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    old[k] = ~0ull;
    if (add[k] && (z[k] & add[k]) != add[k]) old[k] = atomic_or(&Mn[v[k]], add[k]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) first[k] = old[k] == 0ull;
}

/* A wave's walk over the (entry, lane) pairs of 64 entry masks (lane i holds
 * entry i's mask and its list length len): an entry per step (lane = bit,
 * every lane walking the same list) or 64 pairs per step from several
 * entries (each lane its own list, the step as long as its longest list) --
 * whichever takes fewer arc steps: the flat walk suits sparse masks over
 * lists of similar length (grids), the entry walk dense masks and lists of
 * very different lengths (power-law graphs). */
struct PairWalk {
  uint64_t words;
  uint32_t incl, pc, total, steps;
  bool flat;
};
__device__ __forceinline__ PairWalk pair_walk(uint64_t mine, int64_t len) {
  PairWalk w;
  w.words = ballot(mine != 0);
  w.pc = (uint32_t)__popcll(mine);
  w.incl = wave_inclusive_scan(w.pc);
  w.total = __shfl(w.incl, 63, 64);
  const uint32_t flat_steps = (w.total + 63) / 64, word_steps = (uint32_t)__popcll(w.words);
  const uint64_t l = mine ? (uint64_t)len + 1 : 0ull;
  const uint64_t word_cost = wave_sum_u64(l), flat_cost = (uint64_t)flat_steps * wave_max_u64(l);
  w.flat = flat_cost < word_cost;
  w.steps = w.flat ? flat_steps : word_steps;
  return w;
}
/* this lane's pair at `step`: owner entry (lane index) and source lane */
__device__ __forceinline__ bool pair_item(PairWalk& w, uint64_t mine, uint32_t step, int& owner, int& bit) {
  const int lane = lane_id();
  if (w.flat) {
    const uint32_t t = step * 64 + lane;
    owner = scan_owner(w.incl, t);
    const uint32_t before = __shfl(w.incl - w.pc, owner, 64);
    const uint64_t word = shfl_u64(mine, owner);
    bit = t < w.total ? (int)select_bit(word, t - before) : 0;
    return t < w.total;
  }
  owner = __ffsll((unsigned long long)w.words) - 1;
  w.words &= w.words - 1;
  const uint64_t word = shfl_u64(mine, owner);
  bit = lane;
  return (word >> lane) & 1ull;
}

/* running sum S * 2^E of terms a * 2^e (exponents aligned as they come) */
__device__ __forceinline__ void xadd(float& S, int& E, float a, int e) {
  if (e > E) {
    const int sh = E - e;
    S = __builtin_amdgcn_ldexpf(S, sh < -200 ? -200 : sh) + a;
    E = e;
  } else {
    const int sh = e - E;
    S += __builtin_amdgcn_ldexpf(a, sh < -200 ? -200 : sh);
  }
}
constexpr int kNoExp = -(1 << 30);

/* store S * 2^E as mantissa in [0.5, 1) and int16 exponent */
template<typename OffT>
__device__ __forceinline__ void store_norm(const MsArgs<OffT>& a, int64_t at, float S, int E) {
  int k = 0;
  float m = 0.f;
  if (S > 0.f) {
    m = __builtin_amdgcn_frexp_mantf(S);
    k = __builtin_amdgcn_frexp_expf(S);
  }
  const int ex = S > 0.f ? E + k : 0;
  if (ex > kExpMax || ex < -kExpMax) raise_flag(a.ctl, kRange);
  cst(&a.sm[at], m);
  cst(&a.se[at], (int16_t)(ex > kExpMax ? kExpMax : (ex < -kExpMax ? -kExpMax : ex)));
}

/* ------------------------------------------------------------ level phase --- */
/* Push expansion of L_d: lane per arc over groups of 64 entries (long lists
 * are left to their chunks), four arcs per lane in flight. */
template<typename OffT>
__device__ __forceinline__ void expand_push(const MsArgs<OffT>& a, int32_t d, int64_t base, int64_t cnt,
                                            const uint64_t* Mc, uint64_t* Mn, LevelInfo* nx, const Bases& nb,
                                            Stage& stg, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t g = gwave; g * 64 < cnt; g += nwaves) {
    const int64_t i = g * 64 + lane;
    uint64_t fm = 0;
    int64_t st = 0;
    uint32_t len = 0;
    if (i < cnt) {
      const int32_t u = cld(&a.lv[base + i]);
      st = (int64_t)a.off[u];
      const int64_t dg = (int64_t)a.off[u + 1] - st;
      if (dg <= kHub) {
        len = (uint32_t)dg;
        fm = cld(&Mc[u]);
      }
    }
    const uint32_t incl = wave_inclusive_scan(len);
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - len;
    for (uint32_t t0 = 0; t0 < total; t0 += 4 * 64) {
      int32_t v[4];
      uint64_t f[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t t = t0 + 64 * k + lane;
        const int owner = scan_owner(incl, t);
        const uint32_t before = __shfl(excl, owner, 64);
        const int64_t s0 = shfl_i64(st, owner);
        f[k] = shfl_u64(fm, owner);
        v[k] = t < total ? a.cols[s0 + (t - before)] : -1;
      }
      bool first[4];
      relax4(f, v, a.vis, Mc, Mn, first);
#pragma unroll
      for (int k = 0; k < 4; ++k) stage_add(a, nx, nb, stg, first[k], v[k]);
    }
  }
}

/* Push expansion of the chunks of L_d's long lists: a wave per chunk. */
template<typename OffT>
__device__ __forceinline__ void expand_chunks(const MsArgs<OffT>& a, const HubC* hc, int64_t nc, const uint64_t* Mc,
                                              uint64_t* Mn, LevelInfo* nx, const Bases& nb, Stage& st, int64_t gwave,
                                              int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t c = gwave; c < nc; c += nwaves) {
    const HubC h = cld16(&hc[c]);
    const uint64_t fm = cld(&Mc[h.v]);
    const int64_t s = (int64_t)a.off[h.v] + (int64_t)h.c * kHub;
    const int64_t e = std::min<int64_t>(s + kHub, (int64_t)a.off[h.v + 1]);
    for (int64_t j0 = s; j0 < e; j0 += 4 * 64) {
      int32_t v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t j = j0 + 64 * k + lane;
        v[k] = j < e ? a.cols[j] : -1;
      }
      bool first[4];
      const uint64_t f4[4] = {fm, fm, fm, fm};
      relax4(f4, v, a.vis, Mc, Mn, first);
#pragma unroll
      for (int k = 0; k < 4; ++k) stage_add(a, nx, nb, st, first[k], v[k]);
    }
  }
}

/* Pull expansion: every vertex still missing lanes ORs the level masks of its
 * in-neighbours, stopping once the missing lanes are covered; lists longer
 * than 64 arcs are ORed by the whole wave. */
template<typename OffT>
__device__ __forceinline__ void expand_pull(const MsArgs<OffT>& a, const uint64_t* Mc, uint64_t* Mn, LevelInfo* nx,
                                            const Bases& nb, Stage& st, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t base = gwave * 64; base < a.n; base += nwaves * 64) {
    const int64_t v = base + lane;
    uint64_t want = 0, acc = 0;
    int64_t b = 0, e = 0;
    if (v < a.n) {
      want = a.full & ~(cld(&a.vis[v]) | cld(&Mc[v]));
      if (want) {
        b = (int64_t)a.ioff[v];
        e = (int64_t)a.ioff[v + 1];
      }
    }
    const bool coop = e - b > 64;
    if (!coop)
      for (int64_t j = b; j < e; ++j) {
        acc |= cld(&Mc[a.icols[j]]);
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
        if (j0 + lane < ke) part |= cld(&Mc[a.icols[j0 + lane]]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part |= shfl_u64(part, lane ^ o);
        if ((part & kwant) == kwant) break;
      }
      if (lane == k) acc = part;
    }
    const uint64_t add = acc & want;
    const bool first = add != 0;
    if (first) cst(&Mn[v], add);
    stage_add(a, nx, nb, st, first, (int32_t)v);
  }
}

/* Visit L_d: vis |= M_d, depth rows (MS-BFS) or the kept masks (BC). */
template<typename OffT>
__device__ __forceinline__ void visit_level(const MsArgs<OffT>& a, int32_t d, int64_t base, int64_t cnt,
                                            const uint64_t* Mc, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  uint64_t fdeg = 0; /* arcs of the vertices this lane saw every lane reach */
  for (int64_t g = gwave; g * 64 < cnt; g += nwaves) {
    const int64_t i = g * 64 + lane;
    if (i >= cnt) continue;
    const int32_t u = cld(&a.lv[base + i]);
    const uint64_t m = cld(&Mc[u]);
    const uint64_t old = cld(&a.vis[u]);
    cst(&a.vis[u], old | m);
    if (a.bc) {
      cst(&a.lm[base + i], m);
    } else {
      /* one store per source that reached u at this depth (a sparse mask
       * costs its own bits, not the wave's union) */
      for (uint64_t r = m; r; r &= r - 1) a.depth[(int64_t)(__ffsll((unsigned long long)r) - 1) * a.n + u] = d;
    }
    if (a.dir_policy >= SG_POLICY_HYBRID && old != a.full && (old | m) == a.full)
      fdeg += (uint64_t)(a.off[u + 1] - a.off[u]);
  }
  if (a.dir_policy >= SG_POLICY_HYBRID) { /* one add per workgroup and level */
    __shared__ unsigned long long s_fd[kBlock / 64];
    fdeg = wave_sum_u64(fdeg);
    if (lane == 0) s_fd[threadIdx.x >> 6] = fdeg;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int w = 0; w < kBlock / 64; ++w) t += s_fd[w];
      if (t) atomic_add(&a.ctl->full_deg, t);
    }
  }
}

/* Clear the level masks of L_{d-2}, so that their array serves as M_{d+2}. */
template<typename OffT>
__device__ __forceinline__ void clear_level(const MsArgs<OffT>& a, int64_t base, int64_t cnt, uint64_t* Mold,
                                            int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t i = gwave * 64 + lane; i < cnt; i += nwaves * 64) cst(&Mold[cld(&a.lv[base + i])], 0ull);
}

/* One list [vb, ve) walked by the whole wave, lane = source l: the wave
 * loads the ids and neighbour masks of 64 arcs at once (lane k: arc k), then
 * takes the arcs in list order, eight at a time, so each lane issues the
 * row loads of eight arcs before it adds any (the walk is latency-bound:
 * four arcs in flight left it three times slower).  hit[u]: arc u's
 * neighbour has bit l in its mask and this lane takes part (ok).  All lanes
 * call; `term(x, hit)` loads and adds the eight arcs' terms in order. */
template<typename Term>
__device__ __forceinline__ void wave_list_walk(const int32_t* ids, int64_t vb, int64_t ve, const uint64_t* masks,
                                               int l, bool ok, Term&& term) {
  const int lane = lane_id();
  for (int64_t j0 = vb; j0 < ve; j0 += 64) {
    const int64_t j = j0 + lane;
    const int32_t xk = j < ve ? ids[j] : -1;
    const uint64_t mk = xk >= 0 ? cld(&masks[xk]) : 0ull;
    const int cnt = (int)(ve - j0 < 64 ? ve - j0 : 64);
    for (int t0 = 0; t0 < cnt; t0 += 8) {
      int32_t x[8];
      bool hit[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = t0 + u;
        x[u] = __shfl(xk, t, 64);
        const uint64_t m = shfl_u64(mk, t);
        hit[u] = ok && t < cnt && ((m >> l) & 1ull);
      }
      term(x, hit);
    }
  }
}

/* Path counts of the pairs of L_d whose in-lists are short: each pair sums
 * sigma over its in-arcs from depth d-1 (M_{d-1}). */
template<typename OffT>
__device__ __forceinline__ void sigma_pairs(const MsArgs<OffT>& a, int64_t base, int64_t cnt, const uint64_t* Mc,
                                            const uint64_t* Mp, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t g = gwave; g * 64 < cnt; g += nwaves) {
    const int64_t i = g * 64 + lane;
    int32_t u = 0;
    uint64_t own = 0;
    int64_t b = 0, e = 0;
    if (i < cnt) {
      u = cld(&a.lv[base + i]);
      b = (int64_t)a.ioff[u];
      e = (int64_t)a.ioff[u + 1];
      if (e - b <= kHub) own = cld(&Mc[u]);
    }
    PairWalk w = pair_walk(own, e - b);
    for (uint32_t step = 0; step < w.steps; ++step) {
      int owner, l;
      const bool ok = pair_item(w, own, step, owner, l);
      const int32_t v = __shfl(u, owner, 64);
      const int64_t vb = shfl_i64(b, owner), ve = shfl_i64(e, owner);
      float S = 0.f;
      int E = kNoExp;
      if (!w.flat) { /* one vertex, every lane a source */
        wave_list_walk(a.icols, vb, ve, Mp, l, ok, [&](const int32_t (&x)[8], const bool (&hit)[8]) {
          float sv[8];
          int ev[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            sv[k] = hit[k] ? cld(&a.sm[(int64_t)x[k] * 64 + l]) : 0.f;
            ev[k] = hit[k] ? (int)cld(&a.se[(int64_t)x[k] * 64 + l]) : 0;
          }
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (hit[k]) xadd(S, E, sv[k], ev[k]);
        });
        if (ok) store_norm(a, (int64_t)v * 64 + l, S, E);
        continue;
      }
      if (!ok) continue;
      for (int64_t j = vb; j < ve; j += 4) {
        int32_t x[4];
        bool hit[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = j + k < ve ? a.icols[j + k] : -1;
#pragma unroll
        for (int k = 0; k < 4; ++k) hit[k] = x[k] >= 0 && ((cld(&Mp[x[k]]) >> l) & 1ull);
        float sv[4];
        int ev[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          sv[k] = hit[k] ? cld(&a.sm[(int64_t)x[k] * 64 + l]) : 0.f;
          ev[k] = hit[k] ? (int)cld(&a.se[(int64_t)x[k] * 64 + l]) : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (hit[k]) xadd(S, E, sv[k], ev[k]);
      }
      store_norm(a, (int64_t)v * 64 + l, S, E);
    }
  }
}

/* Partial path counts over one chunk of a long in-list: a wave per chunk,
 * lane = source. */
template<typename OffT>
__device__ __forceinline__ void sigma_chunks(const MsArgs<OffT>& a, const HubC* hc, int64_t nc, const uint64_t* Mc,
                                             const uint64_t* Mp, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t c = gwave; c < nc; c += nwaves) {
    const HubC h = cld16(&hc[c]);
    const uint64_t own = cld(&Mc[h.v]);
    const bool mine = (own >> lane) & 1ull;
    const int64_t s = (int64_t)a.ioff[h.v] + (int64_t)h.c * kHub;
    const int64_t e = std::min<int64_t>(s + kHub, (int64_t)a.ioff[h.v + 1]);
    float S = 0.f;
    int E = kNoExp;
    wave_list_walk(a.icols, s, e, Mp, lane, mine, [&](const int32_t (&x)[8], const bool (&hit)[8]) {
      float sv[8];
      int ev[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        sv[k] = hit[k] ? cld(&a.sm[(int64_t)x[k] * 64 + lane]) : 0.f;
        ev[k] = hit[k] ? (int)cld(&a.se[(int64_t)x[k] * 64 + lane]) : 0;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (hit[k]) xadd(S, E, sv[k], ev[k]);
    });
    if (c < a.p_cap) {
      cst(&a.pS[c * 64 + lane], S);
      cst(&a.pE[c * 64 + lane], E);
    } else if (lane == 0) {
      raise_flag(a.ctl, kOvPart);
    }
  }
}

/* Chunked vertices: add the partial rows in chunk order. */
template<typename OffT>
__device__ __forceinline__ void sigma_combine(const MsArgs<OffT>& a, const HubV* hv, int64_t nv, const uint64_t* Mc,
                                              int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t i = gwave; i < nv; i += nwaves) {
    const HubV h = cld16(&hv[i]);
    if (!((cld(&Mc[h.v]) >> lane) & 1ull)) continue;
    float S = 0.f;
    int E = kNoExp;
    for (int32_t c = 0; c < h.nch; ++c) {
      const int64_t slot = (int64_t)h.first + c;
      const float s = cld(&a.pS[slot * 64 + lane]);
      if (s > 0.f) xadd(S, E, s, cld(&a.pE[slot * 64 + lane]));
    }
    store_norm(a, (int64_t)h.v * 64 + lane, S, E);
  }
}

/* ------------------------------------------------------- backward phase --- */
/* delta of a pair from its arc sum: R = (1 + delta) / sigma replaces sigma */
template<typename OffT>
__device__ __forceinline__ void finish_delta(const MsArgs<OffT>& a, int64_t at, float delta) {
  const float m = cld(&a.sm[at]);
  const int e = (int)cld(&a.se[at]);
  cst(&a.dl[at], delta);
  const float r = (1.0f + delta) / m;
  store_norm(a, at, r, -e);
}

/* One arc's term sigma_l(v) / sigma_l(w) * (1 + delta_l(w)) = m_v 2^e_v R_l(w). */
__device__ __forceinline__ float delta_term(float mv, int ev, float mr, int er) {
  const int sh = ev + er;
  return __builtin_amdgcn_ldexpf(mv * mr, sh < -200 ? -200 : (sh > 100 ? 100 : sh));
}

template<typename OffT>
__device__ __forceinline__ void delta_pairs(const MsArgs<OffT>& a, int64_t base, int64_t cnt, const uint64_t* Xa,
                                            int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t g = gwave; g * 64 < cnt; g += nwaves) {
    const int64_t i = g * 64 + lane;
    int32_t u = 0;
    uint64_t own = 0;
    int64_t b = 0, e = 0;
    if (i < cnt) {
      u = cld(&a.lv[base + i]);
      b = (int64_t)a.off[u];
      e = (int64_t)a.off[u + 1];
      if (e - b <= kHub) own = cld(&a.lm[base + i]);
    }
    PairWalk w = pair_walk(own, e - b);
    for (uint32_t step = 0; step < w.steps; ++step) {
      int owner, l;
      const bool ok = pair_item(w, own, step, owner, l);
      const int32_t v = __shfl(u, owner, 64);
      const int64_t vb = shfl_i64(b, owner), ve = shfl_i64(e, owner);
      const int64_t at = (int64_t)v * 64 + l;
      if (!w.flat) { /* one vertex, every lane a source */
        const float mv = ok ? cld(&a.sm[at]) : 0.f;
        const int ev = ok ? (int)cld(&a.se[at]) : 0;
        float acc = 0.f;
        wave_list_walk(a.cols, vb, ve, Xa, l, ok, [&](const int32_t (&x)[8], const bool (&hit)[8]) {
          float r[8];
          int er[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            r[k] = hit[k] ? cld(&a.sm[(int64_t)x[k] * 64 + l]) : 0.f;
            er[k] = hit[k] ? (int)cld(&a.se[(int64_t)x[k] * 64 + l]) : 0;
          }
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (hit[k]) acc += delta_term(mv, ev, r[k], er[k]);
        });
        if (ok) finish_delta(a, at, acc);
        continue;
      }
      if (!ok) continue;
      const float mv = cld(&a.sm[at]);
      const int ev = (int)cld(&a.se[at]);
      float acc = 0.f;
      for (int64_t j = vb; j < ve; j += 4) {
        int32_t x[4];
        bool hit[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = j + k < ve ? a.cols[j + k] : -1;
#pragma unroll
        for (int k = 0; k < 4; ++k) hit[k] = x[k] >= 0 && ((cld(&Xa[x[k]]) >> l) & 1ull);
        float r[4];
        int er[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          r[k] = hit[k] ? cld(&a.sm[(int64_t)x[k] * 64 + l]) : 0.f;
          er[k] = hit[k] ? (int)cld(&a.se[(int64_t)x[k] * 64 + l]) : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (hit[k]) acc += delta_term(mv, ev, r[k], er[k]);
      }
      finish_delta(a, at, acc);
    }
  }
}

template<typename OffT>
__device__ __forceinline__ void delta_chunks(const MsArgs<OffT>& a, const HubC* hc, int64_t nc, int64_t lbase,
                                             const uint64_t* Xa, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t c = gwave; c < nc; c += nwaves) {
    const HubC h = cld16(&hc[c]);
    const uint64_t own = cld(&a.lm[lbase + h.pos]);
    const bool mine = (own >> lane) & 1ull;
    const int64_t at = (int64_t)h.v * 64 + lane;
    const float mv = mine ? cld(&a.sm[at]) : 0.f;
    const int ev = mine ? (int)cld(&a.se[at]) : 0;
    const int64_t s = (int64_t)a.off[h.v] + (int64_t)h.c * kHub;
    const int64_t e = std::min<int64_t>(s + kHub, (int64_t)a.off[h.v + 1]);
    float acc = 0.f;
    wave_list_walk(a.cols, s, e, Xa, lane, mine, [&](const int32_t (&x)[8], const bool (&hit)[8]) {
      float r[8];
      int er[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        r[k] = hit[k] ? cld(&a.sm[(int64_t)x[k] * 64 + lane]) : 0.f;
        er[k] = hit[k] ? (int)cld(&a.se[(int64_t)x[k] * 64 + lane]) : 0;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (hit[k]) acc += delta_term(mv, ev, r[k], er[k]);
    });
    if (c < a.p_cap) cst(&a.pS[c * 64 + lane], acc);
    else if (lane == 0) raise_flag(a.ctl, kOvPart);
  }
}

template<typename OffT>
__device__ __forceinline__ void delta_combine(const MsArgs<OffT>& a, const HubV* hv, int64_t nv, int64_t lbase,
                                              int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t i = gwave; i < nv; i += nwaves) {
    const HubV h = cld16(&hv[i]);
    if (!((cld(&a.lm[lbase + h.pos]) >> lane) & 1ull)) continue;
    float acc = 0.f;
    for (int32_t c = 0; c < h.nch; ++c) acc += cld(&a.pS[((int64_t)h.first + c) * 64 + lane]);
    finish_delta(a, (int64_t)h.v * 64 + lane, acc);
  }
}

/* Level membership for the next backward step: OR L_d's masks into Xb and
 * take L_{d+2}'s out (disjoint lanes, so the two commute). */
template<typename OffT>
__device__ __forceinline__ void scatter_masks(const MsArgs<OffT>& a, int64_t base, int64_t cnt, uint64_t* Xb,
                                              bool set, int64_t gwave, int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t i = gwave * 64 + lane; i < cnt; i += nwaves * 64) {
    const int32_t v = cld(&a.lv[base + i]);
    const uint64_t m = cld(&a.lm[base + i]);
    if (set) atomic_or(&Xb[v], m);
    else atomic_and(&Xb[v], ~m);
  }
}

/* bc(v) += sum of delta over the lanes of v's entry of one level, in lane order */
template<typename OffT>
__device__ __forceinline__ void accumulate(const MsArgs<OffT>& a, int64_t base, int64_t cnt, int64_t gwave,
                                           int64_t nwaves) {
  const int lane = lane_id();
  for (int64_t i = gwave * 64 + lane; i < cnt; i += nwaves * 64) {
    const int32_t v = cld(&a.lv[base + i]);
    uint64_t m = cld(&a.lm[base + i]);
    double s = 0.0;
    while (m) {
      const int l = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      s += (double)cld(&a.dl[(int64_t)v * 64 + l]);
    }
    cst(&a.bc64[v], cld(&a.bc64[v]) + s);
  }
}

/* ------------------------------------------------------------- the kernel --- */
/* Direction of level d's expansion (the same in every workgroup: the inputs
 * are the totals every workgroup reads after the barrier).  dir: the
 * previous level's direction; cnt / fe / mx: the level's vertices, arcs and
 * largest degree; full_deg: the arcs of the vertices every lane has reached. */
template<typename OffT>
__device__ __forceinline__ int32_t ms_direction(const MsArgs<OffT>& a, const sg::policy::Staged* pol, int32_t d,
                                                int32_t dir, int64_t cnt, unsigned long long fe, unsigned long long mx,
                                                unsigned long long full_deg) {
  switch (a.dir_policy) {
  case SG_POLICY_PUSH: return 0;
  case SG_POLICY_PULL: return 1;
  case SG_POLICY_HYBRID:
  case SG_POLICY_LEARNED: {
    if (d == 0) return 0; /* the sources push, as a single-source traversal starts */
    const unsigned long long unexplored = a.total_degree - full_deg;
    if (a.dir_policy == SG_POLICY_LEARNED) {
      if (fe * 1024ull < a.total_degree) return 0; /* the single-source BFS's default gate */
      return sg::policy::infer_staged(pol, (double)cnt, (double)fe, (double)unexplored, d, (double)mx, a.f6, a.f7) / 3;
    }
    if (dir == 0) return ((float)(uint32_t)fe > (float)(uint32_t)unexplored / a.alpha) ? 1 : 0;
    return ((float)cnt < (float)a.n / a.beta) ? 0 : 1;
  }
  default: return cnt * 20 > a.n ? 1 : 0;
  }
}

/* A level's header words (LevelInfo's first kLevelInfoZero), read once per
 * workgroup by its first lanes and shared through LDS: every thread loading
 * them from the point of coherence put hundreds of thousands of loads on a
 * few lines per level. */
enum : int { kHCount, kHArcs, kHBase, kHHov, kHHovBase, kHHoc, kHHocBase, kHHiv, kHHivBase, kHHic, kHHicBase, kHMx };
__device__ __forceinline__ void load_head(unsigned long long* dst, const LevelInfo* L, int lane0) {
  const int k = (int)threadIdx.x - lane0;
  if (k >= 0 && k < kLevelInfoZero) dst[k] = atomic_load(reinterpret_cast<const unsigned long long*>(L) + k);
}

/* three waves per SIMD: the kernel holds a level loop's state, the pair
 * walks and four arcs' loads in flight; left alone the compiler takes 176
 * registers (two waves), which costs the power-law graphs' big levels 7 % */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) ms_kernel(MsArgs<OffT> a) {
  __shared__ sg::policy::Staged s_pol;
  __shared__ int32_t s_stage[kBlock / 64][kStage];
  __shared__ unsigned long long s_red[2 * (kBlock / 64)];
  __shared__ unsigned long long s_h[3][kLevelInfoZero];
  __shared__ unsigned long long s_full;
  __shared__ unsigned long long s_acc[2 * kBarGroups];
  Stage st{s_stage[threadIdx.x >> 6], 0u, 0ull, 0ull};
  /* wave ranks run across the workgroups first, so a small level's few busy
   * waves spread over as many compute units as there are of them */
  const int64_t gwave = (int64_t)(threadIdx.x >> 6) * gridDim.x + blockIdx.x;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const int64_t n = a.n;
  uint32_t t = 0;
  int32_t d = 0, dir = 0, n_pull = 0;
  if (a.dir_policy == SG_POLICY_LEARNED) sg::policy::stage(&s_pol, a.mlp, (int)threadIdx.x, kBlock);
  int64_t max_chunks = 0; /* largest chunk list of a level (the backward pass's partial rows) */
  for (;; ++d) {
    LevelInfo* L = level(a, d);
    load_head(s_h[0], L, 0);
    if (d >= 2) load_head(s_h[1], level(a, d - 2), 32);
    if (threadIdx.x == 63) s_full = atomic_load(&a.ctl->full_deg);
    if (threadIdx.x >= 64 && threadIdx.x < 64 + 2 * kBarGroups) { /* the level's arcs and largest degree */
      const int k = threadIdx.x - 64;
      const Ctl::LevelAcc& c = a.ctl->acc[d % 3][k >> 1];
      s_acc[k] = atomic_load((k & 1) ? &c.mx : &c.arcs);
    }
    __syncthreads();
    const unsigned long long* h = s_h[0];
    const int64_t cnt = (int64_t)h[kHCount];
    if (cnt == 0) break;
    if (!a.ring && d + 2 >= a.li_cap) { /* every workgroup takes this exit */
      if (blockIdx.x == 0 && threadIdx.x == 0) raise_flag(a.ctl, kOvLevels);
      return;
    }
    const int64_t base = a.ring ? (int64_t)(d & 3) * a.lcap : (int64_t)h[kHBase];
    uint64_t* Mc = a.M + (int64_t)(d & 3) * n;
    uint64_t* Mn = a.M + (int64_t)((d + 1) & 3) * n;
    const uint64_t* Mp = a.M + (int64_t)((d + 3) & 3) * n;
    LevelInfo* nx = level(a, d + 1);
    Bases nb;
    if (a.ring) {
      const int64_t sl = (d + 1) & 3;
      nb.l = sl * a.lcap;
      nb.hov = nb.hiv = sl * a.hv_cap;
      nb.hoc = nb.hic = sl * a.hc_cap;
    } else {
      nb.l = (int64_t)(h[kHBase] + h[kHCount]);
      nb.hov = (int64_t)(h[kHHovBase] + h[kHHov]);
      nb.hoc = (int64_t)(h[kHHocBase] + h[kHHoc]);
      nb.hiv = (int64_t)(h[kHHivBase] + h[kHHiv]);
      nb.hic = (int64_t)(h[kHHicBase] + h[kHHic]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      LevelInfo* z = level(a, d + 2); /* the level after next starts from zero */
      unsigned long long* zp = reinterpret_cast<unsigned long long*>(z);
      for (int k = 0; k < kLevelInfoZero; ++k) atomic_store(zp + k, 0ull);
      for (int k = 0; k < kBarGroups; ++k) { /* read by every workgroup at level d - 1, all past its barrier */
        atomic_store(&a.ctl->acc[(d + 2) % 3][k].arcs, 0ull);
        atomic_store(&a.ctl->acc[(d + 2) % 3][k].mx, 0ull);
      }
      atomic_store(&nx->base, (unsigned long long)nb.l);
      atomic_store(&nx->hov_base, (unsigned long long)nb.hov);
      atomic_store(&nx->hoc_base, (unsigned long long)nb.hoc);
      atomic_store(&nx->hiv_base, (unsigned long long)nb.hiv);
      atomic_store(&nx->hic_base, (unsigned long long)nb.hic);
    }
    const int64_t hoc = (int64_t)h[kHHoc];
    max_chunks = hoc > max_chunks ? hoc : max_chunks;
    const HubC* hocp = a.hoc + (a.ring ? (int64_t)(d & 3) * a.hc_cap : (int64_t)h[kHHocBase]);
    unsigned long long lv_arcs = 0, lv_mx = 0;
    for (int k = 0; k < kBarGroups; ++k) {
      lv_arcs += s_acc[2 * k];
      lv_mx = s_acc[2 * k + 1] > lv_mx ? s_acc[2 * k + 1] : lv_mx;
    }
    if (d == 0) { /* the sources' arcs, counted by the seeding */
      lv_arcs = h[kHArcs];
      lv_mx = h[kHMx];
    }
    dir = ms_direction(a, &s_pol, d, dir, cnt, lv_arcs, lv_mx, s_full);
    n_pull += dir;
    const int64_t o_cnt = d >= 2 ? (int64_t)s_h[1][kHCount] : 0;
    const int64_t o_base = d >= 2 ? (a.ring ? (int64_t)((d - 2) & 3) * a.lcap : (int64_t)s_h[1][kHBase]) : 0;
    int64_t hiv = 0, hic = 0;
    const HubV* hivp = nullptr;
    const HubC* hicp = nullptr;
    if (a.bc && d >= 1) {
      hicp = a.directed ? a.hic + (int64_t)h[kHHicBase] : hocp;
      hic = a.directed ? (int64_t)h[kHHic] : hoc;
      hiv = a.directed ? (int64_t)h[kHHiv] : (int64_t)h[kHHov];
      hivp = a.directed ? a.hiv + (int64_t)h[kHHivBase] : a.hov + (int64_t)h[kHHovBase];
    }
    if (dir) {
      expand_pull(a, Mc, Mn, nx, nb, st, gwave, nwaves);
    } else {
      expand_push(a, d, base, cnt, Mc, Mn, nx, nb, st, gwave, nwaves);
      expand_chunks(a, hocp, hoc, Mc, Mn, nx, nb, st, gwave, nwaves);
    }
    stage_finish(a, nx, nb, st, s_red, a.ctl->acc[(d + 1) % 3]);
    visit_level(a, d, base, cnt, Mc, gwave, nwaves);
    if (d >= 2) clear_level(a, o_base, o_cnt, a.M + (int64_t)((d + 2) & 3) * n, gwave, nwaves);
    if (a.bc && d >= 1) {
      sigma_pairs(a, base, cnt, Mc, Mp, gwave, nwaves);
      sigma_chunks(a, hicp, hic, Mc, Mp, gwave, nwaves);
    }
    if (!grid_sync(a.ctl, t, a.wait_ticks)) return;
    if (hiv > 0) {
      sigma_combine(a, hivp, hiv, Mc, gwave, nwaves);
      if (!grid_sync(a.ctl, t, a.wait_ticks)) return;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctl->levels = d;
    a.ctl->n_pull = n_pull;
  }
  if (!a.bc) return;
  /* the partial rows must hold every level's chunks before anything is added
   * to bc (a batch that stops here runs again with more rows) */
  if (max_chunks > a.p_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) raise_flag(a.ctl, kOvPart);
    return;
  }

  /* backward: levels d-1 .. 1; X[j] = M[j] hold the memberships */
  const int32_t deepest = d - 1;
  for (int64_t i = gwave * 64 + lane_id(); i < n; i += nwaves * 64) {
    cst(&a.M[i], 0ull);
    cst(&a.M[n + i], 0ull);
  }
  if (!grid_sync(a.ctl, t, a.wait_ticks)) return;
  for (int32_t dd = deepest; dd >= 1; --dd) {
    load_head(s_h[0], level(a, dd), 0);
    if (dd + 1 <= deepest) load_head(s_h[1], level(a, dd + 1), 32);
    if (dd + 2 <= deepest) load_head(s_h[2], level(a, dd + 2), 64);
    __syncthreads();
    const unsigned long long* h = s_h[0];
    const int64_t base = (int64_t)h[kHBase], cnt = (int64_t)h[kHCount];
    const uint64_t* Xa = a.M + (int64_t)((dd + 1) & 1) * n;
    uint64_t* Xb = a.M + (int64_t)(dd & 1) * n;
    const int64_t hov = (int64_t)h[kHHov], hoc = (int64_t)h[kHHoc];
    const HubV* hovp = a.hov + (int64_t)h[kHHovBase];
    const HubC* hocp = a.hoc + (int64_t)h[kHHocBase];
    const int64_t b1 = (int64_t)s_h[1][kHBase], c1 = (int64_t)s_h[1][kHCount];
    const int64_t b2 = (int64_t)s_h[2][kHBase], c2 = (int64_t)s_h[2][kHCount];
    delta_pairs(a, base, cnt, Xa, gwave, nwaves);
    delta_chunks(a, hocp, hoc, base, Xa, gwave, nwaves);
    scatter_masks(a, base, cnt, Xb, true, gwave, nwaves);
    if (dd + 2 <= deepest) scatter_masks(a, b2, c2, Xb, false, gwave, nwaves);
    if (dd + 1 <= deepest) accumulate(a, b1, c1, gwave, nwaves);
    if (!grid_sync(a.ctl, t, a.wait_ticks)) return;
    if (hov > 0) {
      delta_combine(a, hovp, hov, base, gwave, nwaves);
      if (!grid_sync(a.ctl, t, a.wait_ticks)) return;
    }
  }
  if (deepest >= 1) {
    LevelInfo* L1 = level(a, 1);
    accumulate(a, (int64_t)atomic_load(&L1->base), (int64_t)atomic_load(&L1->count), gwave, nwaves);
  }
}

/* ----------------------------------------------------------- batch set-up --- */
__global__ void ms_clear_kernel(int64_t n, uint64_t* vis, uint64_t* M, LevelInfo* li, int32_t nli, Ctl* ctl) {
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = tid; i < n; i += stride) {
    vis[i] = 0ull;
    M[i] = 0ull;
    M[n + i] = 0ull;
    M[2 * n + i] = 0ull;
    M[3 * n + i] = 0ull;
  }
  unsigned long long* lp = reinterpret_cast<unsigned long long*>(li);
  for (int64_t i = tid; i < (int64_t)nli * (int64_t)(sizeof(LevelInfo) / 8); i += stride) lp[i] = 0ull;
  uint32_t* cp = reinterpret_cast<uint32_t*>(ctl);
  for (int64_t i = tid; i < (int64_t)(sizeof(Ctl) / 4); i += stride) cp[i] = 0u;
}

/* Level 0: the batch's distinct sources (one wave; lane i = source i). */
template<typename OffT>
__global__ void __launch_bounds__(64) ms_seed_kernel(MsArgs<OffT> a) {
  const int i = threadIdx.x;
  bool first = false;
  int32_t s = 0;
  if (i < a.k) {
    s = a.src[i];
    atomic_or(&a.vis[s], 1ull << i);
    first = atomic_or(&a.M[s], 1ull << i) == 0ull;
    if (a.bc) {
      a.sm[(int64_t)s * 64 + i] = 0.5f; /* sigma = 1 = 0.5 * 2^1 */
      a.se[(int64_t)s * 64 + i] = 1;
    }
  }
  Bases b{0, 0, 0, 0, 0};
  append(a, a.li, b, first, s);
}

__global__ void ms_fill_i32_kernel(int32_t* p, int32_t val, int64_t cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = val;
}

__global__ void ms_zero_f64_kernel(double* p, int64_t cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void ms_add_to_float_kernel(const double* acc, float* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)((double)out[i] + acc[i]);
}

/* ------------------------------------------------------------------ host --- */
/* Capacities of the growable lists (entries); kept with the graph under "msb_caps". */
struct Caps {
  int64_t lcap = 0, hv_cap = 0, hc_cap = 0, p_cap = 0, li_cap = 0;
};

template<typename OffT>
int launch_batches(sg_graph_t g, bool bc, int32_t nsrc, const int32_t* sources, int32_t* depth, float* out,
                   const sg_msbfs_options& opt) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n, m = g->m;
  const bool directed = g->inv_offsets != nullptr;
  const int64_t hubs_bound = m / kHub + 64, chunks_bound = 2 * (m / kHub) + 64;
  Caps caps;
  /* the growable capacities are the sizes of the buffers kept with the graph */
  auto cap_of = [&](const char* name, int64_t dflt) -> int64_t {
    auto it = g->buffers.find(name);
    return it == g->buffers.end() ? dflt : (int64_t)it->second.second;
  };
  const int32_t kmax = std::min<int32_t>(64, nsrc);
  if (bc) {
    caps.lcap = std::max<int64_t>(cap_of("msb_lv", 0) / 4, std::min<int64_t>(16, kmax) * n + 64);
    caps.hv_cap = std::max<int64_t>(cap_of("msb_hov", 0) / (int64_t)sizeof(HubV), 4 * hubs_bound);
    caps.hc_cap = std::max<int64_t>(cap_of("msb_hoc", 0) / (int64_t)sizeof(HubC), 4 * chunks_bound);
    caps.p_cap = std::max<int64_t>(cap_of("msb_pS", 0) / (64 * 4), chunks_bound / 4 + 64);
    caps.li_cap = std::max<int64_t>(cap_of("msb_li", 0) / (int64_t)sizeof(LevelInfo), std::min<int64_t>(n + 3, 65536));
  } else {
    caps.lcap = n + 64;
    caps.hv_cap = hubs_bound;
    caps.hc_cap = chunks_bound;
    caps.li_cap = kLiRing;
  }
  static int occ = 0; /* per offset width (one instantiation each) */
  if (occ == 0) {
    int nb = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ms_kernel<OffT>, kBlock, 0));
    occ = std::max(1, std::min(nb, 4));
  }
  const int grid = occ * dev->cus;
  const int grid_n = sg::grid_for(dev, n, kBlock, 8);

  if (!bc) {
    ms_fill_i32_kernel<<<sg::grid_for(dev, (int64_t)nsrc * n, 256, 8), 256, 0, s>>>(depth, (int32_t)(n + 1),
                                                                                   (int64_t)nsrc * n);
    SG_HIP(hipGetLastError());
  }
  double* bc64 = nullptr;
  if (bc) {
    SG_TRY(sg::graph_buffer(g, "msb_bc64", 8 * (size_t)(n + 1), &bc64));
    ms_zero_f64_kernel<<<grid_n, kBlock, 0, s>>>(bc64, n + 1);
  }
  Ctl* hctl = static_cast<Ctl*>(dev->host_scratch);
  for (int32_t b0 = 0; b0 < nsrc;) {
    const int32_t k = std::min<int32_t>(64, nsrc - b0);
    MsArgs<OffT> a{};
    a.n = n;
    a.full = k == 64 ? ~0ull : ((1ull << k) - 1ull);
    a.k = k;
    a.bc = bc ? 1 : 0;
    a.ring = bc ? 0 : 1;
    a.directed = (bc && directed) ? 1 : 0;
    a.off = (const OffT*)g->offsets;
    a.cols = g->cols;
    a.ioff = (const OffT*)(directed ? g->inv_offsets : g->offsets);
    a.icols = directed ? g->inv_cols : g->cols;
    const int64_t lslots = bc ? caps.lcap : 4 * caps.lcap;
    const int64_t hvslots = bc ? caps.hv_cap : 4 * caps.hv_cap;
    const int64_t hcslots = bc ? caps.hc_cap : 4 * caps.hc_cap;
    SG_TRY(sg::graph_buffer(g, "msb_vis", 8 * (size_t)(n + 1), &a.vis));
    SG_TRY(sg::graph_buffer(g, "msb_M", 8 * 4 * (size_t)(n + 1), &a.M));
    SG_TRY(sg::graph_buffer(g, "msb_lv", 4 * (size_t)lslots, &a.lv));
    SG_TRY(sg::graph_buffer(g, "msb_hov", sizeof(HubV) * (size_t)hvslots, &a.hov));
    SG_TRY(sg::graph_buffer(g, "msb_hoc", sizeof(HubC) * (size_t)hcslots, &a.hoc));
    SG_TRY(sg::graph_buffer(g, "msb_li", sizeof(LevelInfo) * (size_t)caps.li_cap, &a.li));
    SG_TRY(sg::graph_buffer(g, "msb_ctl", sizeof(Ctl), &a.ctl));
    SG_TRY(sg::graph_buffer(g, "msb_src", 4 * 64, &a.src));
    a.lcap = caps.lcap;
    a.hv_cap = caps.hv_cap;
    a.hc_cap = caps.hc_cap;
    a.li_cap = caps.li_cap;
    if (bc) {
      SG_TRY(sg::graph_buffer(g, "msb_lm", 8 * (size_t)lslots, &a.lm));
      if (directed) {
        SG_TRY(sg::graph_buffer(g, "msb_hiv", sizeof(HubV) * (size_t)hvslots, &a.hiv));
        SG_TRY(sg::graph_buffer(g, "msb_hic", sizeof(HubC) * (size_t)hcslots, &a.hic));
      }
      SG_TRY(sg::graph_buffer(g, "msb_sm", 4 * 64 * (size_t)(n + 1), &a.sm));
      SG_TRY(sg::graph_buffer(g, "msb_se", 2 * 64 * (size_t)(n + 1), &a.se));
      SG_TRY(sg::graph_buffer(g, "msb_dl", 4 * 64 * (size_t)(n + 1), &a.dl));
      SG_TRY(sg::graph_buffer(g, "msb_pS", 4 * 64 * (size_t)caps.p_cap, &a.pS));
      SG_TRY(sg::graph_buffer(g, "msb_pE", 4 * 64 * (size_t)caps.p_cap, &a.pE));
      a.p_cap = caps.p_cap;
      a.bc64 = bc64;
    } else {
      a.depth = depth + (int64_t)b0 * n;
    }
    a.wait_ticks = dev->wait_ticks;
    a.dir_policy = opt.policy;
    a.alpha = opt.alpha;
    a.beta = opt.beta;
    a.mlp = opt.policy == SG_POLICY_LEARNED ? sg::policy_device_params(opt.learned) : nullptr;
    a.f6 = sg::policy::feature(6, 0, 0, 0, 0, 0, (double)n, (double)m);
    a.f7 = sg::policy::feature(7, 0, 0, 0, 0, 0, (double)n, (double)m);
    a.total_degree = (unsigned long long)g->total_degree;
    SG_HIP(hipMemcpyAsync((void*)a.src, sources + b0, 4 * (size_t)k, hipMemcpyHostToDevice, s));
    ms_clear_kernel<<<grid_n, kBlock, 0, s>>>(n, a.vis, a.M, a.li, bc ? 2 : kLiRing, a.ctl);
    ms_seed_kernel<OffT><<<1, 64, 0, s>>>(a);
    SG_HIP(hipGetLastError());
    void* args[] = {&a};
    SG_HIP(launch_resident(reinterpret_cast<const void*>(ms_kernel<OffT>), grid, kBlock, args, s));
    SG_HIP(hipMemcpyAsync(hctl, a.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    const uint32_t fl = hctl->flags;
    SG_CHECK(!(fl & kTimeout), SG_ERR_HIP, "%s: a grid barrier waited longer than the device's wait limit",
             bc ? "sg_bc_batched" : "sg_bfs_multi");
    SG_CHECK(!(fl & kRange), SG_ERR_UNSUPPORTED, "sg_bc_batched: path counts beyond 2^%d", kExpMax);
    if (fl & (kOvList | kOvHubV | kOvHubC | kOvPart | kOvLevels)) {
      /* grow what overflowed to its bound and run the batch again (the
       * bounds hold for any batch: each (vertex, lane) pair lies at one depth) */
      SG_CHECK(bc, SG_ERR_HIP, "sg_bfs_multi: level list overflow");
      bool grew = false;
      auto grow = [&](int64_t& c, int64_t bound) {
        if (c < bound) {
          c = bound;
          grew = true;
        }
      };
      if (fl & kOvList) grow(caps.lcap, (int64_t)kmax * n + 64);
      if (fl & kOvHubV) grow(caps.hv_cap, (int64_t)kmax * hubs_bound);
      if (fl & kOvHubC) grow(caps.hc_cap, (int64_t)kmax * chunks_bound);
      if (fl & kOvPart) grow(caps.p_cap, chunks_bound);
      if (fl & kOvLevels) grow(caps.li_cap, std::min<int64_t>(n + 3, 8 * caps.li_cap));
      SG_CHECK(grew, SG_ERR_HIP, "sg_bc_batched: list capacity exceeded its bound (flags %u)", fl);
      continue; /* same batch, larger lists */
    }
    b0 += k;
  }
  if (bc) {
    ms_add_to_float_kernel<<<grid_n, kBlock, 0, s>>>(bc64, out, n);
    SG_HIP(hipGetLastError());
    SG_HIP(hipStreamSynchronize(s));
  }
  return SG_OK;
}

int check_sources(sg_graph_t g, int32_t nsrc, const int32_t* sources, const char* who) {
  SG_CHECK(nsrc >= 0, SG_ERR_INVALID, "%s: negative source count", who);
  for (int32_t i = 0; i < nsrc; ++i)
    SG_CHECK(sources[i] >= 0 && sources[i] < g->n, SG_ERR_INVALID, "%s: source %d out of range [0,%lld)", who,
             sources[i], (long long)g->n);
  return SG_OK;
}

} // namespace

extern "C" void sg_msbfs_options_default(sg_msbfs_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->policy = -1;
  opt->alpha = 15.0f;
  opt->beta = 24.0f;
}

/* Multi-source BFS: 64 sources per batch, one bit-lane each, with a choice
 * of the per-level direction rule. */
extern "C" int sg_bfs_multi_ex(sg_graph_t g, int32_t nsrc, const int32_t* sources_host,
                               const sg_msbfs_options* opt_in, int32_t* depth) {
  SG_CHECK(g && depth && (nsrc == 0 || sources_host), SG_ERR_INVALID, "sg_bfs_multi: null argument");
  SG_TRY(check_sources(g, nsrc, sources_host, "sg_bfs_multi"));
  sg_msbfs_options opt;
  if (opt_in) opt = *opt_in;
  else sg_msbfs_options_default(&opt);
  SG_CHECK(opt.policy >= -1 && opt.policy <= SG_POLICY_LEARNED, SG_ERR_INVALID, "sg_bfs_multi: policy %d out of range",
           opt.policy);
  SG_CHECK(opt.policy != SG_POLICY_LEARNED || opt.learned, SG_ERR_INVALID,
           "sg_bfs_multi: the learned policy needs a trained sg_policy_t in options.learned");
  SG_CHECK(opt.policy != SG_POLICY_HYBRID || (opt.alpha > 0.f && opt.beta > 0.f), SG_ERR_INVALID,
           "sg_bfs_multi: alpha and beta must be positive");
  if (nsrc == 0) return SG_OK;
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return launch_batches<int64_t>(g, false, nsrc, sources_host, depth, nullptr, opt);
  return launch_batches<uint32_t>(g, false, nsrc, sources_host, depth, nullptr, opt);
}

/* Multi-source BFS: 64 sources per batch, one bit-lane each. */
extern "C" int sg_bfs_multi(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, int32_t* depth) {
  return sg_bfs_multi_ex(g, nsrc, sources_host, nullptr, depth);
}

/* Betweenness centrality over many sources, 64 per lane batch; the same sums
 * as sg_bc (bc.hpp:118-200 per source) up to FP32 rounding. */
extern "C" int sg_bc_batched(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, float* bc) {
  SG_CHECK(g && bc && (nsrc == 0 || sources_host), SG_ERR_INVALID, "sg_bc_batched: null argument");
  SG_TRY(check_sources(g, nsrc, sources_host, "sg_bc_batched"));
  if (nsrc == 0) return SG_OK;
  SG_HIP(hipSetDevice(g->dev->ordinal));
  sg_msbfs_options opt;
  sg_msbfs_options_default(&opt);
  if (g->offset_bytes == 8) return launch_batches<int64_t>(g, true, nsrc, sources_host, nullptr, bc, opt);
  return launch_batches<uint32_t>(g, true, nsrc, sources_host, nullptr, bc, opt);
}
