/*
 * bc.hip -- betweenness centrality (BC::run, algorithms/bc.hpp:118-200) for
 * gfx950.
 *
 * Per source s:
 *  1. depths from the fused BFS (sg_bfs, hybrid push/pull);
 *  2. vertices grouped by level: a radix sort of (level, vertex) with the
 *     level clamped to the number of levels, so only a few key bits are sorted;
 *  3. forward, level by level: sigma(v) = sum of sigma(u) over in-arcs u -> v
 *     with depth(u) = depth(v) - 1, then the level is rescaled by a power of
 *     two so that its largest sigma lies in [0.5, 1) -- path counts overflow
 *     FP64 on long-diameter graphs (a 4096 x 4096 grid has ~1e2464 shortest
 *     paths corner to corner), and power-of-two scaling is exact;
 *  4. backward, deepest level first: delta(u) = sigma(u) * scale(L+1) *
 *     sum over out-arcs u -> w with depth(w) = depth(u) + 1 of
 *     (1 + delta(w)) / sigma(w); bc(u) += delta(u) for u != s (no halving on
 *     undirected graphs, as bc.hpp:186-196).
 * Steps 3 and 4 are pulls: each vertex gathers its own sum, so there are no
 * floating-point atomics and results do not depend on scheduling.  A lane
 * sums a vertex's arcs in CSR order; vertices of degree above kCoop are
 * summed by the whole wave (strided partial sums, fixed butterfly), which
 * changes the rounding order but not the result beyond FP64 rounding.
 * sigma, delta and the per-source accumulation are FP64; the output adds the
 * FP64 total to the caller's float array once at the end.
 */
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.hpp"

using namespace sygraph::hipdev;

namespace {

constexpr int kBlock = 256;
constexpr int kCoop = 16; /* degree above which the wave sums one vertex together */
constexpr int kMaxSlots = 64;   /* per-level maximum of sigma: slots, one cache line apart */
constexpr int kSlotStride = 16; /* unsigned long long per 128-byte line */
constexpr int64_t kLevelBitmapBytes = 1ll << 30; /* per-level bitmaps up to 1 GiB, else depth[] tests */

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ int64_t shfl_i64_bc(int64_t x, int src) {
  const uint32_t lo = __shfl((uint32_t)((uint64_t)x & 0xFFFFFFFFull), src, 64);
  const uint32_t hi = __shfl((uint32_t)((uint64_t)x >> 32), src, 64);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__global__ void level_keys_kernel(const int32_t* depth, int64_t n, int32_t nlevels, uint32_t* keys, int32_t* ids) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = depth[i];
    keys[i] = (d >= 0 && d < nlevels) ? (uint32_t)d : (uint32_t)nlevels;
    ids[i] = (int32_t)i;
  }
}

/* lvl[k] = first position of key k in the sorted keys; lvl[0..nlevels+1]
 * preset to n by the caller (levels 0..nlevels-1 are never empty) */
__global__ void level_starts_kernel(const uint32_t* keys, int64_t n, int64_t* lvl) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (i == 0 || keys[i] != keys[i - 1]) lvl[keys[i]] = i;
}

/* lbm[L * w0 + word]: bit v set iff depth(v) == L, for L < nlevels; a wave
 * owns the word of its 64 vertices, so there are no atomics (the words are
 * zeroed before) */
__global__ void level_bits_kernel(const int32_t* depth, int64_t n, int64_t w0, int32_t nlevels, uint64_t* lbm) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t w = gwave; w < w0; w += nwaves) {
    const int64_t v = w * 64 + lane;
    const int32_t d = v < n ? depth[v] : -1;
    const bool in = d >= 0 && d < nlevels;
    uint64_t todo = ballot(in);
    while (todo) { /* one store per distinct level among the wave's vertices */
      const int k = __ffsll((unsigned long long)todo) - 1;
      const int32_t lev = __shfl(d, k, 64);
      const uint64_t m = ballot(in && d == lev);
      if (lane == 0) lbm[(int64_t)lev * w0 + w] = m;
      todo &= ~m;
    }
  }
}

__device__ __forceinline__ bool bit_of(const uint64_t* bm, int32_t v) { return (bm[v >> 6] >> (v & 63)) & 1ull; }

__global__ void bc_source_init_kernel(double* sigma, double* scale, int32_t s) {
  sigma[s] = 1.0;
  scale[0] = 1.0;
}

/* Sum of f(arc) over the arcs of each vertex in [begin, end) of `order`; the
 * wave takes 64 consecutive entries, lanes with long lists are summed
 * cooperatively.  `store(v, sum)` runs once per vertex. */
template<typename OffT, typename ArcF, typename StoreF>
__device__ __forceinline__ void gather_level(const int32_t* order, int64_t begin, int64_t end, const OffT* off,
                                             ArcF arc, StoreF store) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t base = begin + gwave * 64; base < end; base += nwaves * 64) {
    const int64_t i = base + lane;
    int32_t v = -1;
    int64_t a = 0, b = 0;
    if (i < end) {
      v = order[i];
      a = (int64_t)off[v];
      b = (int64_t)off[v + 1];
    }
    const bool coop = (b - a) > kCoop;
    double acc = 0.0;
    if (v >= 0 && !coop)
      for (int64_t e = a; e < b; ++e) acc += arc(e);
    uint64_t heavy = ballot(coop);
    while (heavy) {
      const int k = __ffsll((unsigned long long)heavy) - 1;
      heavy &= heavy - 1;
      const int64_t ka = shfl_i64_bc(a, k), kb = shfl_i64_bc(b, k);
      double part = 0.0;
      for (int64_t e = ka + lane; e < kb; e += 64) part += arc(e);
      part = wave_sum_f64(part);
      if (lane == k) acc = part;
    }
    if (v >= 0) store(v, acc);
  }
}

/* prev_bm: bitmap of level L-1 (L2-resident), or nullptr to test depth[]:
 * sigma is only loaded for real parents */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) sigma_level_kernel(const int32_t* order, const int64_t* lvl, int32_t L,
                                                             const OffT* in_off, const int32_t* in_cols,
                                                             const int32_t* depth, const uint64_t* prev_bm,
                                                             double* sigma, unsigned long long* level_max) {
  const int64_t begin = lvl[L], end = lvl[L + 1];
  double mx = 0.0;
  gather_level<OffT>(
      order, begin, end, in_off,
      [&](int64_t e) {
        const int32_t u = in_cols[e];
        const bool parent = prev_bm ? bit_of(prev_bm, u) : depth[u] == L - 1;
        double x = 0.0;
        if (parent) x = sigma[u];
        return x;
      },
      [&](int32_t v, double s) {
        sigma[v] = s;
        mx = s > mx ? s : mx;
      });
  /* positive doubles order like their bit patterns; one atomic per workgroup,
   * spread over kMaxSlots cache lines (same-address atomics from every wave
   * serialise across the chip) */
  __shared__ unsigned long long red[kBlock / 64];
  unsigned long long bits = (unsigned long long)__double_as_longlong(mx);
  bits = wave_max_u64(bits);
  if (lane_id() == 0) red[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int w = 0; w < kBlock / 64; ++w) m = red[w] > m ? red[w] : m;
    if (m) atomic_max(&level_max[(blockIdx.x % kMaxSlots) * kSlotStride], m);
  }
}

__global__ void __launch_bounds__(kBlock) sigma_scale_kernel(const int32_t* order, const int64_t* lvl, int32_t L,
                                                             double* sigma, const unsigned long long* level_max,
                                                             double* scale) {
  __shared__ unsigned long long s_max;
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int k = 0; k < kMaxSlots; ++k) m = level_max[k * kSlotStride] > m ? level_max[k * kSlotStride] : m;
    s_max = m;
  }
  __syncthreads();
  const double mx = __longlong_as_double((long long)s_max);
  int ex = 0;
  (void)frexp(mx, &ex);
  const double sc = ldexp(1.0, -ex);
  if (blockIdx.x == 0 && threadIdx.x == 0) scale[L] = sc;
  const int64_t begin = lvl[L], end = lvl[L + 1];
  for (int64_t i = begin + blockIdx.x * (int64_t)kBlock + threadIdx.x; i < end; i += (int64_t)gridDim.x * kBlock)
    sigma[order[i]] *= sc;
}

/* next_bm: bitmap of level L+1, or nullptr to test depth[] */
template<typename OffT>
__global__ void __launch_bounds__(kBlock) delta_level_kernel(const int32_t* order, const int64_t* lvl, int32_t L,
                                                             const OffT* off, const int32_t* cols, const int32_t* depth,
                                                             const uint64_t* next_bm, const double* sigma,
                                                             double* delta, const double* scale, double* acc_bc,
                                                             int32_t s) {
  const int64_t begin = lvl[L], end = lvl[L + 1];
  const double sc = scale[L + 1];
  gather_level<OffT>(
      order, begin, end, off,
      [&](int64_t e) {
        const int32_t w = cols[e];
        const bool child = next_bm ? bit_of(next_bm, w) : depth[w] == L + 1;
        double x = 0.0;
        if (child) x = (1.0 + delta[w]) / sigma[w];
        return x;
      },
      [&](int32_t v, double sum) {
        const double d = sigma[v] * sum * sc;
        delta[v] = d;
        if (v != s) acc_bc[v] += d;
      });
}

__global__ void add_to_float_kernel(const double* acc, float* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)((double)out[i] + acc[i]);
}

__global__ void fill_f64_kernel(double* p, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_i64_kernel(int64_t* p, int64_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

struct BcBuffers {
  int32_t* depth = nullptr;
  uint32_t* keys[2] = {nullptr, nullptr};
  int32_t* ids[2] = {nullptr, nullptr};
  int64_t* lvl = nullptr;
  double* sigma = nullptr;
  double* delta = nullptr;
  double* acc = nullptr;
  double* scale = nullptr;
  unsigned long long* level_max = nullptr;
  int64_t level_max_cap = 0; /* levels the slot array holds */
  uint64_t* lbm = nullptr;   /* per-level vertex bitmaps */
  int64_t lbm_cap = 0;       /* levels lbm holds */
  void* sort_tmp = nullptr;
  size_t sort_bytes = 0;
  ~BcBuffers() {
    void* ps[] = {depth, keys[0], keys[1], ids[0], ids[1], lvl, sigma, delta, acc, scale, level_max, lbm, sort_tmp};
    for (void* p : ps) (void)hipFree(p);
  }
};

template<typename OffT>
int run_bc(sg_graph_t g, int32_t nsrc, const int32_t* sources, float* out) {
  sg_device_t dev = g->dev;
  hipStream_t st = dev->stream;
  const int64_t n = g->n;
  BcBuffers B;
  SG_TRY(sg::alloc_device(dev, (void**)&B.depth, 4 * (n + 1)));
  for (int k = 0; k < 2; ++k) {
    SG_TRY(sg::alloc_device(dev, (void**)&B.keys[k], 4 * (n + 1)));
    SG_TRY(sg::alloc_device(dev, (void**)&B.ids[k], 4 * (n + 1)));
  }
  SG_TRY(sg::alloc_device(dev, (void**)&B.lvl, 8 * (n + 2)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.sigma, 8 * (n + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.delta, 8 * (n + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.acc, 8 * (n + 1)));
  SG_TRY(sg::alloc_device(dev, (void**)&B.scale, 8 * (n + 2)));
  B.level_max_cap = 0; /* grown per source to the number of levels */
  SG_HIP(rocprim::radix_sort_pairs(nullptr, B.sort_bytes, B.keys[0], B.keys[1], B.ids[0], B.ids[1], (size_t)n, 0, 32, st));
  SG_TRY(sg::alloc_device(dev, &B.sort_tmp, B.sort_bytes + 16));

  const OffT* off = (const OffT*)g->offsets;
  const OffT* in_off = (const OffT*)(g->inv_offsets ? g->inv_offsets : g->offsets);
  const int32_t* in_cols = g->inv_cols ? g->inv_cols : g->cols;
  const int grid_n = sg::grid_for(dev, n, kBlock, 4);
  const int grid_lvl = dev->cus * 4;
  fill_f64_kernel<<<grid_n, kBlock, 0, st>>>(B.acc, 0.0, n);

  sg_bfs_options opt;
  sg_bfs_options_default(&opt);
  opt.policy = SG_POLICY_HYBRID;
  opt.record_levels = 0;
  auto stats = std::make_unique<sg_bfs_stats>();
  for (int32_t si = 0; si < nsrc; ++si) {
    const int32_t s = sources[si];
    SG_CHECK(s >= 0 && s < n, SG_ERR_INVALID, "sg_bc: source %d out of range [0,%lld)", s, (long long)n);
    SG_TRY(sg_bfs(g, s, &opt, B.depth, stats.get()));
    const int32_t nl = stats->levels; /* depths 0 .. nl-1 */
    if (nl < 2) continue;             /* nothing reached: no dependencies */
    int bits = 1;
    while ((1u << bits) <= (uint32_t)nl) ++bits;
    level_keys_kernel<<<grid_n, kBlock, 0, st>>>(B.depth, n, nl, B.keys[0], B.ids[0]);
    size_t tb = B.sort_bytes;
    SG_HIP(rocprim::radix_sort_pairs(B.sort_tmp, tb, B.keys[0], B.keys[1], B.ids[0], B.ids[1], (size_t)n, 0, bits, st));
    fill_i64_kernel<<<sg::grid_for(dev, nl + 2, kBlock, 1), kBlock, 0, st>>>(B.lvl, n, nl + 2);
    level_starts_kernel<<<grid_n, kBlock, 0, st>>>(B.keys[1], n, B.lvl);
    if (B.level_max_cap < nl + 1) {
      SG_HIP(hipStreamSynchronize(st));
      (void)hipFree(B.level_max);
      B.level_max = nullptr;
      SG_TRY(sg::alloc_device(dev, (void**)&B.level_max, 8 * (size_t)(nl + 1) * kMaxSlots * kSlotStride));
      B.level_max_cap = nl + 1;
    }
    fill_i64_kernel<<<sg::grid_for(dev, (int64_t)(nl + 1) * kMaxSlots * kSlotStride, kBlock, 4), kBlock, 0, st>>>(
        (int64_t*)B.level_max, 0, (int64_t)(nl + 1) * kMaxSlots * kSlotStride);
    fill_f64_kernel<<<grid_n, kBlock, 0, st>>>(B.delta, 0.0, n);
    bc_source_init_kernel<<<1, 1, 0, st>>>(B.sigma, B.scale, s);
    /* level bitmaps when they take at most kLevelBitmapBytes (long-diameter
     * graphs fall back to testing depth[]) */
    const int64_t w0 = (n + 63) / 64;
    const bool use_lbm = (int64_t)nl * w0 * 8 <= kLevelBitmapBytes;
    if (use_lbm) {
      if (B.lbm_cap < nl) {
        SG_HIP(hipStreamSynchronize(st));
        (void)hipFree(B.lbm);
        B.lbm = nullptr;
        SG_TRY(sg::alloc_device(dev, (void**)&B.lbm, 8 * (size_t)nl * w0));
        B.lbm_cap = nl;
      }
      fill_i64_kernel<<<sg::grid_for(dev, (int64_t)nl * w0, kBlock, 4), kBlock, 0, st>>>((int64_t*)B.lbm, 0,
                                                                                      (int64_t)nl * w0);
      level_bits_kernel<<<sg::grid_for(dev, n, kBlock, 4), kBlock, 0, st>>>(B.depth, n, w0, nl, B.lbm);
    }
    auto lbm_of = [&](int32_t L) -> const uint64_t* { return use_lbm ? B.lbm + (int64_t)L * w0 : nullptr; };
    const int32_t* order = B.ids[1];
    for (int32_t L = 1; L < nl; ++L) {
      unsigned long long* lm = B.level_max + (size_t)L * kMaxSlots * kSlotStride;
      sigma_level_kernel<OffT><<<grid_lvl, kBlock, 0, st>>>(order, B.lvl, L, in_off, in_cols, B.depth, lbm_of(L - 1),
                                                            B.sigma, lm);
      sigma_scale_kernel<<<grid_lvl, kBlock, 0, st>>>(order, B.lvl, L, B.sigma, lm, B.scale);
    }
    for (int32_t L = nl - 2; L >= 0; --L)
      delta_level_kernel<OffT><<<grid_lvl, kBlock, 0, st>>>(order, B.lvl, L, off, g->cols, B.depth, lbm_of(L + 1),
                                                            B.sigma, B.delta, B.scale, B.acc, s);
    SG_HIP(hipGetLastError());
  }
  add_to_float_kernel<<<grid_n, kBlock, 0, st>>>(B.acc, out, n);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(st));
  return SG_OK;
}

} // namespace

extern "C" int sg_bc(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, float* bc) {
  SG_CHECK(g && bc && (nsrc == 0 || sources_host), SG_ERR_INVALID, "sg_bc: null argument");
  SG_CHECK(nsrc >= 0, SG_ERR_INVALID, "sg_bc: negative source count");
  SG_HIP(hipSetDevice(g->dev->ordinal));
  if (g->offset_bytes == 8) return run_bc<int64_t>(g, nsrc, sources_host, bc);
  return run_bc<uint32_t>(g, nsrc, sources_host, bc);
}
