/*
 * cc.hip -- connected components for gfx950 (CC::run, algorithms/cc.hpp:103-169).
 *
 * The reference propagates the larger label along arcs until nothing changes,
 * so on an undirected graph every vertex ends with the largest vertex id of
 * its component.  That fixpoint is reached here without per-level rounds:
 *  - undirected graphs: concurrent union-find.  An arc (u,v) hooks the
 *    smaller of the two roots under the larger with a CAS, so parent[x] >= x
 *    always holds and the surviving root of a component is its largest id
 *    (the largest vertex can never be hooked under a larger root).  Work is
 *    O(m) regardless of the graph's diameter, and most arcs are skipped by
 *    sampling (the Afforest scheme of Sutton, Ben-Nun and Barak, IPDPS 2018):
 *    every vertex first links its first kSampleArcs arcs; after a flatten the
 *    most frequent root among a fixed sample of vertices is taken as the
 *    giant component; then only vertices outside it link their remaining
 *    arcs (an arc into the giant component is seen from its other end, since
 *    both directions are stored), and a last flatten points every vertex at
 *    its root.
 *  - directed graphs: label propagation along out-arcs with atomic max
 *    (labels[v] = largest id that reaches v), frontier by frontier, which is
 *    exactly the reference's fixpoint on a directed graph.  An undirected
 *    graph whose CSR does not hold the reverse of every arc takes this path
 *    too (the reference propagates along out-arcs whatever the graph's
 *    flag); symmetry is checked once per graph by an order-independent hash
 *    of the arcs against that of their reverses (sym_hash_kernel).
 * Arcs are visited through the degree-class queues of the whole vertex set
 * (thread / wave / workgroup bins), so hubs are split across the grid.
 */
#include <algorithm>
#include <cstring>
#include <vector>

#include "traverse.hpp"

using namespace sg::trav;

namespace {

__device__ __forceinline__ uint32_t find_root(uint32_t* parent, uint32_t x) {
  uint32_t p = atomic_load(&parent[x]);
  while (p != x) {
    const uint32_t gp = atomic_load(&parent[p]);
    if (gp != p) atomic_store(&parent[x], gp); /* path halving; parents only grow */
    x = p;
    p = gp;
  }
  return x;
}

__device__ __forceinline__ void hook(uint32_t* parent, uint32_t u, uint32_t v) {
  uint32_t ru = find_root(parent, u), rv = find_root(parent, v);
  while (ru != rv) {
    const uint32_t lo = ru < rv ? ru : rv, hi = ru < rv ? rv : ru;
    uint32_t expected = lo;
    if (atomic_cas(&parent[lo], expected, hi)) return;
    ru = find_root(parent, expected);
    rv = find_root(parent, hi);
  }
}

template<typename OffT>
__global__ void __launch_bounds__(kBlock) hook_kernel(QEntry* q0, QEntry* q1, QEntry* q2, const QueueHdr* hdr,
                                                      const int32_t* cols, uint32_t* parent) {
  QEntry* const q[3] = {q0, q1, q2};
  for_each_queued_arc<OffT>(q, hdr, cols, [&](uint32_t u, uint32_t v, int64_t) { hook(parent, u, v); });
}

constexpr int kSampleArcs = 2;
constexpr int kRootSamples = 1024;

/* every vertex links its first kSampleArcs arcs */
template<typename OffT>
__global__ void sample_link_kernel(const OffT* off, const int32_t* cols, int64_t n, uint32_t* parent) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = (int64_t)off[u], e = (int64_t)off[u + 1];
    for (int64_t j = b; j < e && j < b + kSampleArcs; ++j) hook(parent, (uint32_t)u, (uint32_t)cols[j]);
  }
}

/* roots of the sampled vertices i * n / kRootSamples (flattened parents) */
__global__ void sample_roots_kernel(const uint32_t* parent, int64_t n, uint32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kRootSamples) out[i] = parent[(int64_t)i * n / kRootSamples];
}

/* vertices outside the giant component with arcs left after the sampled ones */
template<typename OffT>
__global__ void outside_kernel(const OffT* off, const uint32_t* parent, int64_t n, int64_t w0, uint32_t giant,
                               uint64_t* bm) {
  const int lane = lane_id();
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t w = gw; w < w0; w += nw) {
    const int64_t v = w * 64 + lane;
    const bool keep = v < n && parent[v] != giant && (int64_t)(off[v + 1] - off[v]) > kSampleArcs;
    const uint64_t m = ballot(keep);
    if (lane == 0) bm[w] = m;
  }
}

/* Symmetry check: the sum over arcs (u,v) of H(u,v) - H(v,u) (mod 2^64), H a
 * 64-bit mix of the ordered pair, is exactly zero when every arc has its
 * reverse (the arc multiset equals its transpose) and nonzero otherwise
 * except with probability ~2^-64.  A wave takes 64 vertices; lists longer
 * than 64 arcs are summed by the whole wave. */
__device__ __forceinline__ uint64_t pair_mix(uint32_t a, uint32_t b) {
  uint64_t x = ((uint64_t)a << 32) | b; /* splitmix64 finaliser */
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

template<typename OffT>
__global__ void __launch_bounds__(kBlock) sym_hash_kernel(const OffT* off, const int32_t* cols, int64_t n,
                                                          unsigned long long* acc) {
  const int lane = lane_id();
  const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
  uint64_t h = 0;
  for (int64_t base = gw * 64; base < n; base += nw * 64) {
    const int64_t u = base + lane;
    int64_t b = 0, e = 0;
    if (u < n) {
      b = (int64_t)off[u];
      e = (int64_t)off[u + 1];
    }
    const bool coop = e - b > 64;
    if (!coop)
      for (int64_t j = b; j < e; ++j) {
        const uint32_t v = (uint32_t)cols[j];
        h += pair_mix((uint32_t)u, v) - pair_mix(v, (uint32_t)u);
      }
    uint64_t heavy = ballot(coop);
    while (heavy) {
      const int k = __ffsll((unsigned long long)heavy) - 1;
      heavy &= heavy - 1;
      const int64_t kb = shfl_i64(b, k), ke = shfl_i64(e, k);
      const uint32_t ku = (uint32_t)(base + k);
      for (int64_t j = kb + lane; j < ke; j += 64) {
        const uint32_t v = (uint32_t)cols[j];
        h += pair_mix(ku, v) - pair_mix(v, ku);
      }
    }
  }
  h = wave_sum_u64(h);
  if (lane == 0 && h) atomic_add(acc, (unsigned long long)h);
}

__global__ void zero_acc_kernel(unsigned long long* acc) {
  if (threadIdx.x == 0) *acc = 0ull;
}

/* Whether an undirected graph stores both directions of every arc, checked
 * once per graph (acc: one device word to sum into); cached in g. */
template<typename OffT>
int graph_symmetric(sg_graph_t g, unsigned long long* acc, bool* sym) {
  if (g->symmetric < 0) {
    sg_device_t dev = g->dev;
    hipStream_t s = dev->stream;
    auto* host = static_cast<unsigned long long*>(dev->host_scratch);
    zero_acc_kernel<<<1, 64, 0, s>>>(acc);
    sym_hash_kernel<OffT><<<sg::grid_for(dev, (g->n + 63) / 64 * 64, kBlock, 8), kBlock, 0, s>>>(
        (const OffT*)g->offsets, g->cols, g->n, acc);
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpyAsync(host, acc, 8, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    g->symmetric = *host == 0ull ? 1 : 0;
  }
  *sym = g->symmetric == 1;
  return SG_OK;
}

__global__ void iota_kernel(uint32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

/* After hooking every tree is final; each thread follows its own chain
 * without path halving, because a halving write by another thread could
 * replace an already flattened parent[i] with a (stale) non-root ancestor. */
__global__ void flatten_kernel(uint32_t* parent, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t x = atomic_load(&parent[i]);
    for (uint32_t p = atomic_load(&parent[x]); p != x; p = atomic_load(&parent[x])) x = p;
    atomic_store(&parent[i], x);
  }
}

__global__ void fill_words_kernel(uint64_t* bm, int64_t w0, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < w0; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rem = n - i * 64;
    bm[i] = rem >= 64 ? ~0ull : ((1ull << rem) - 1ull);
  }
}

/* directed: push the larger label along out-arcs; changed vertices form the
 * next frontier */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) maxprop_kernel(QEntry* q0, QEntry* q1, QEntry* q2, const QueueHdr* hdr,
                                                         const int32_t* cols, uint32_t* labels, uint64_t* next_bm,
                                                         uint64_t* zero_bm, int64_t w0) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < w0; i += (int64_t)gridDim.x * kBlock) zero_bm[i] = 0ull;
  QEntry* const q[3] = {q0, q1, q2};
  for_each_queued_arc<OffT>(q, hdr, cols, [&](uint32_t u, uint32_t v, int64_t) {
    const uint32_t lu = atomic_load(&labels[u]);
    if (lu > labels[v]) {
      const uint32_t old = atomic_max(&labels[v], lu);
      if (lu > old) {
        const uint64_t bit = 1ull << (v & 63);
        if (!(next_bm[v >> 6] & bit)) atomic_or(&next_bm[v >> 6], bit);
      }
    }
  });
}

template<typename OffT>
int run_cc(sg_graph_t g, uint32_t* labels, int32_t* rounds_out) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n, w0 = (n + 63) / 64;
  const Tuning tune = default_tuning();
  QueueSet* Qp = nullptr;
  uint64_t* bm[2] = {nullptr, nullptr};
  uint32_t* roots = nullptr;
  int st = graph_queue_set(g, &Qp);
  if (st == SG_OK) st = sg::graph_buffer(g, "cc_bm0", 8 * (w0 + 1), &bm[0]);
  if (st == SG_OK) st = sg::graph_buffer(g, "cc_bm1", 8 * (w0 + 1), &bm[1]);
  if (st == SG_OK) st = sg::graph_buffer(g, "cc_roots", 4 * kRootSamples, &roots);
  if (st != SG_OK) return st;
  QueueSet& Q = *Qp;
  const OffT* off = (const OffT*)g->offsets;
  const int grid = dev->cus * 8;
  int32_t rounds = 0;
  bool sym = false;
  if (!(g->flags & SG_GRAPH_DIRECTED)) SG_TRY(graph_symmetric<OffT>(g, reinterpret_cast<unsigned long long*>(roots), &sym));
  iota_kernel<<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(labels, n);
  fill_words_kernel<<<sg::grid_for(dev, w0, 256, 4), 256, 0, s>>>(bm[0], w0, n);
  if (sym) {
    const int grid_n = sg::grid_for(dev, n, 256, 4);
    sample_link_kernel<OffT><<<grid_n, 256, 0, s>>>(off, g->cols, n, labels);
    flatten_kernel<<<grid_n, 256, 0, s>>>(labels, n);
    /* the giant component: the most frequent root of a fixed vertex sample */
    std::vector<uint32_t> h(kRootSamples);
    if (n > 0) {
      sample_roots_kernel<<<kRootSamples / 256, 256, 0, s>>>(labels, n, roots);
      SG_HIP(hipGetLastError());
      SG_HIP(hipMemcpyAsync(h.data(), roots, 4 * kRootSamples, hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
    }
    std::sort(h.begin(), h.end());
    uint32_t giant = h[0];
    for (size_t i = 0, best = 0; i < h.size();) {
      size_t j = i;
      while (j < h.size() && h[j] == h[i]) ++j;
      if (j - i > best) {
        best = j - i;
        giant = h[i];
      }
      i = j;
    }
    outside_kernel<OffT><<<sg::grid_for(dev, w0 * 64, 256, 4), 256, 0, s>>>(off, labels, n, w0, giant, bm[0]);
    st = compact_to_queues<OffT, kFromBitmap>(s, Q, 0, off, nullptr, nullptr, bm[0], tune, false, kSampleArcs);
    if (st == SG_OK) {
      hook_kernel<OffT><<<grid, kBlock, 0, s>>>(Q.q[0][0], Q.q[0][1], Q.q[0][2], Q.hdr, g->cols, labels);
      flatten_kernel<<<grid_n, 256, 0, s>>>(labels, n);
      rounds = 1;
    }
  } else {
    /* every vertex starts active (the all-vertex pass of cc.hpp:135-145) */
    QueueHdr* hh = (QueueHdr*)dev->host_scratch;
    int cur = 0, qs = 0;
    st = compact_to_queues<OffT, kFromBitmap>(s, Q, qs, off, nullptr, nullptr, bm[cur], tune);
    if (st == SG_OK) st = sg::zero_async(dev, bm[1], 8 * w0);
    while (st == SG_OK) {
      maxprop_kernel<OffT><<<grid, kBlock, 0, s>>>(Q.q[qs][0], Q.q[qs][1], Q.q[qs][2], Q.hdr + qs, g->cols, labels,
                                                   bm[cur ^ 1], bm[cur], w0);
      ++rounds;
      st = compact_to_queues<OffT, kFromBitmap>(s, Q, qs ^ 1, off, nullptr, nullptr, bm[cur ^ 1], tune);
      if (st != SG_OK) break;
      SG_HIP(hipMemcpyAsync(hh, Q.hdr + (qs ^ 1), sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      qs ^= 1;
      cur ^= 1;
      if (hh->nvert == 0) break;
    }
  }
  SG_HIP(hipStreamSynchronize(s));
  if (rounds_out) *rounds_out = rounds;
  return st;
}

/* ------------------------------------------------------------ no-alloc CC ---
 * sg_cc_noalloc computes the labels of sg_cc with the same kernels for the
 * hooks and flattens, but takes every buffer beyond the caller's labels from
 * one file-scope array, so it never allocates device memory:
 *   [0, kNaWords)            bitmap A: vertices outside the giant component
 *                            (undirected) or the current frontier (directed)
 *   [kNaWords, 2 kNaWords)   bitmap B: the next frontier (directed)
 *   then kRootSamples roots of the sampled vertices (uint32), the giant root
 *   and two frontier counters.
 * The array holds bitmaps of up to kNaMaxN vertices (32 MiB in all); every
 * device keeps its own copy at its own address, taken once per device when
 * the device is opened (cc_scratch_init).  Calls on one device share it, so
 * calls that overlap on one device race. */
constexpr int64_t kNaMaxN = int64_t(1) << 27;
constexpr int64_t kNaWords = kNaMaxN / 64;
constexpr int64_t kNaRoots = 2 * kNaWords;                  /* kRootSamples uint32 */
constexpr int64_t kNaMisc = kNaRoots + kRootSamples / 2;    /* giant, counter 0, counter 1 */
constexpr int64_t kNaTotalWords = kNaMisc + 32;

static __device__ __align__(256) uint64_t scratch[kNaTotalWords];

/* The most frequent root of the sample (the smallest of equally frequent
 * ones, as the host-side choice of sg_cc): one workgroup of kRootSamples
 * threads sorts the sample in LDS and keeps the longest run. */
__global__ void __launch_bounds__(kRootSamples) giant_kernel(const uint32_t* roots, uint64_t* giant) {
  __shared__ uint32_t key[kRootSamples];
  __shared__ unsigned long long best;
  const int i = threadIdx.x;
  key[i] = roots[i];
  if (i == 0) best = 0ull;
  __syncthreads();
  for (int k = 2; k <= kRootSamples; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = i ^ j;
      if (p > i) {
        const bool up = (i & k) == 0;
        const uint32_t a = key[i], b = key[p];
        if ((a > b) == up) {
          key[i] = b;
          key[p] = a;
        }
      }
      __syncthreads();
    }
  }
  if (i == 0 || key[i] != key[i - 1]) {
    int len = 1;
    while (i + len < kRootSamples && key[i + len] == key[i]) ++len;
    /* longest run first, then the smallest root */
    const unsigned long long score = ((unsigned long long)len << 32) | (0xffffffffull - key[i]);
    atomicMax(&best, score);
  }
  __syncthreads();
  if (i == 0) *giant = 0xffffffffull - (best & 0xffffffffull);
}

/* outside_kernel with the giant root read from device memory */
template<typename OffT>
__global__ void outside_na_kernel(const OffT* off, const uint32_t* parent, int64_t n, int64_t w0,
                                  const uint64_t* giant, uint64_t* bm) {
  const int lane = lane_id();
  const uint32_t gr = (uint32_t)*giant;
  const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t w = gw; w < w0; w += nw) {
    const int64_t v = w * 64 + lane;
    const bool keep = v < n && parent[v] != gr && (int64_t)(off[v + 1] - off[v]) > kSampleArcs;
    const uint64_t m = ballot(keep);
    if (lane == 0) bm[w] = m;
  }
}

/* Visit the arcs [b, e) of the vertices whose bit is set in the wave's word:
 * lists up to 64 arcs by their own lane, longer ones by the whole wave. */
template<typename OffT, typename F>
__device__ __forceinline__ void wave_word_arcs(const OffT* off, const int32_t* cols, int64_t w, uint64_t word,
                                               int64_t skip, F&& f) {
  const int lane = lane_id();
  const int64_t u = w * 64 + lane;
  int64_t b = 0, e = 0;
  if ((word >> lane) & 1ull) {
    b = (int64_t)off[u] + skip;
    e = (int64_t)off[u + 1];
    if (b > e) b = e;
  }
  const bool coop = e - b > 64;
  if (!coop)
    for (int64_t j = b; j < e; ++j) f((uint32_t)u, (uint32_t)cols[j]);
  uint64_t heavy = ballot(coop);
  while (heavy) {
    const int k = __ffsll((unsigned long long)heavy) - 1;
    heavy &= heavy - 1;
    const int64_t kb = shfl_i64(b, k), ke = shfl_i64(e, k);
    const uint32_t ku = (uint32_t)(w * 64 + k);
    for (int64_t j = kb + lane; j < ke; j += 64) f(ku, (uint32_t)cols[j]);
  }
}

/* vertices outside the giant component link their arcs after the sampled ones */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) hook_na_kernel(const OffT* off, const int32_t* cols, const uint64_t* bm,
                                                         int64_t w0, uint32_t* parent) {
  const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t w = gw; w < w0; w += nw) {
    const uint64_t word = bm[w];
    if (word) wave_word_arcs(off, cols, w, word, kSampleArcs, [&](uint32_t u, uint32_t v) { hook(parent, u, v); });
  }
}

/* Directed rounds: every vertex of the frontier pushes its label along its
 * out-arcs with atomic max; the vertices whose label grew form the next
 * frontier, counted in ctr[round & 1] (ctr[(round + 1) & 1] is cleared for the
 * next round); the frontier's words are cleared as they are read. */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) maxprop_na_kernel(const OffT* off, const int32_t* cols, uint32_t* labels,
                                                            uint64_t* cur, uint64_t* next, int64_t w0,
                                                            unsigned long long* ctr, int32_t round) {
  if (blockIdx.x == 0 && threadIdx.x == 0) atomic_store(&ctr[(round + 1) & 1], 0ull);
  const int64_t gw = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t w = gw; w < w0; w += nw) {
    const uint64_t word = cur[w];
    if (!word) continue;
    if (lane_id() == 0) cur[w] = 0ull;
    uint32_t joined = 0;
    wave_word_arcs(off, cols, w, word, 0, [&](uint32_t u, uint32_t v) {
      const uint32_t lu = atomic_load(&labels[u]);
      if (lu > atomic_load(&labels[v]) && lu > atomic_max(&labels[v], lu)) {
        const uint64_t bit = 1ull << (v & 63);
        if (!(atomic_load(&next[v >> 6]) & bit) && !(atomic_or(&next[v >> 6], bit) & bit)) ++joined;
      }
    });
    const uint64_t tot = wave_sum_u64(joined);
    if (lane_id() == 0 && tot) atomic_add(&ctr[round & 1], (unsigned long long)tot);
  }
}

__global__ void zero_words_kernel(uint64_t* p, int64_t cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

template<typename OffT>
int run_cc_noalloc(sg_graph_t g, uint32_t* labels, int32_t* rounds_out) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n, w0 = (n + 63) / 64;
  uint64_t* base = static_cast<uint64_t*>(dev->cc_scratch);
  uint64_t* bmA = base;
  uint64_t* bmB = base + kNaWords;
  uint32_t* roots = reinterpret_cast<uint32_t*>(base + kNaRoots);
  uint64_t* giant = base + kNaMisc;
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base + kNaMisc + 1);
  const OffT* off = (const OffT*)g->offsets;
  const int grid = dev->cus * 8;
  const int grid_n = sg::grid_for(dev, n, 256, 4);
  const int grid_w = sg::grid_for(dev, w0 * 64, kBlock, 8);
  int32_t rounds = 0;
  bool sym = false;
  if (!(g->flags & SG_GRAPH_DIRECTED)) SG_TRY(graph_symmetric<OffT>(g, ctr, &sym));
  iota_kernel<<<grid_n, 256, 0, s>>>(labels, n);
  if (sym) {
    sample_link_kernel<OffT><<<grid_n, 256, 0, s>>>(off, g->cols, n, labels);
    flatten_kernel<<<grid_n, 256, 0, s>>>(labels, n);
    sample_roots_kernel<<<kRootSamples / 256, 256, 0, s>>>(labels, n, roots);
    giant_kernel<<<1, kRootSamples, 0, s>>>(roots, giant);
    outside_na_kernel<OffT><<<grid_w, kBlock, 0, s>>>(off, labels, n, w0, giant, bmA);
    hook_na_kernel<OffT><<<grid, kBlock, 0, s>>>(off, g->cols, bmA, w0, labels);
    flatten_kernel<<<grid_n, 256, 0, s>>>(labels, n);
    SG_HIP(hipGetLastError());
    rounds = 1;
  } else {
    /* every vertex starts active (the all-vertex pass of cc.hpp:135-145) */
    auto* host = static_cast<unsigned long long*>(dev->host_scratch);
    fill_words_kernel<<<sg::grid_for(dev, w0, 256, 4), 256, 0, s>>>(bmA, w0, n);
    zero_words_kernel<<<sg::grid_for(dev, w0, 256, 4), 256, 0, s>>>(bmB, w0);
    zero_words_kernel<<<1, 64, 0, s>>>(reinterpret_cast<uint64_t*>(ctr), 2);
    uint64_t* cur = bmA;
    uint64_t* nxt = bmB;
    for (;;) {
      maxprop_na_kernel<OffT><<<grid, kBlock, 0, s>>>(off, g->cols, labels, cur, nxt, w0, ctr, rounds);
      SG_HIP(hipGetLastError());
      SG_HIP(hipMemcpyAsync(host, ctr + (rounds & 1), 8, hipMemcpyDeviceToHost, s));
      SG_HIP(hipStreamSynchronize(s));
      ++rounds;
      if (*host == 0ull) break;
      uint64_t* t = cur;
      cur = nxt;
      nxt = t;
    }
  }
  SG_HIP(hipStreamSynchronize(s));
  if (rounds_out) *rounds_out = rounds;
  return SG_OK;
}

} // namespace

namespace sg {
/* The device address of this device's copy of the no-alloc scratch array,
 * taken once when the device is opened (the device is current). */
void cc_scratch_init(sg_device_t dev) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  dev->cc_scratch = e == hipSuccess ? p : nullptr;
  dev->cc_scratch_error = e;
}

void aux_queues_release(sg_graph_t g) {
  if (g && g->aux_queues) {
    auto* qs = static_cast<sg::trav::QueueSet*>(g->aux_queues);
    qs->release();
    delete qs;
    g->aux_queues = nullptr;
  }
}
} // namespace sg

extern "C" int sg_cc(sg_graph_t g, uint32_t* labels, int32_t* rounds) {
  SG_CHECK(g && labels, SG_ERR_INVALID, "sg_cc: null argument");
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return run_cc<int64_t>(g, labels, rounds);
  return run_cc<uint32_t>(g, labels, rounds);
}

/* CC::run (algorithms/cc.hpp:103-169) with no device allocation: the labels
 * of sg_cc, with every other buffer in the library's static scratch array. */
extern "C" int sg_cc_noalloc(sg_graph_t g, uint32_t* labels, int32_t* rounds) {
  SG_CHECK(g && labels, SG_ERR_INVALID, "sg_cc_noalloc: null argument");
  SG_CHECK(g->n <= kNaMaxN, SG_ERR_INVALID,
           "sg_cc_noalloc: %lld vertices exceed the static scratch's maximum of %lld (use sg_cc)",
           (long long)g->n, (long long)kNaMaxN);
  SG_HIP(hipSetDevice(g->dev->ordinal));
  SG_CHECK(g->dev->cc_scratch, SG_ERR_HIP, "sg_cc_noalloc: address of the static scratch unavailable: %s",
           hipGetErrorString(g->dev->cc_scratch_error));
  if (g->offset_bytes == 8) return run_cc_noalloc<int64_t>(g, labels, rounds);
  return run_cc_noalloc<uint32_t>(g, labels, rounds);
}
