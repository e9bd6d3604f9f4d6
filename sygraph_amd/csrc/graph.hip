/*
 * graph.hip -- CSR graph storage on the device, device-side inverse CSR and
 * the seeded synthetic-graph generators.
 *
 * Reference: graph/impls/graph_csr.hpp:158-439 (GraphCSR owns a device copy of
 * the CSR and, for directed graphs, an inverse CSR built on the host with
 * formats/csr.hpp:179-209; undirected graphs alias it, :423-425).  Here the
 * inverse is built on the device by sorting (dst, src) keys, which yields
 * in-lists ordered by source id exactly like the host transpose.
 *
 * HBM layout of one graph: offsets (n+1 x 4|8 B), cols (m x int32),
 * weights (m x 4 B, optional) -- structure-of-arrays so BFS never touches
 * weights; inverse arrays only when directed.
 */
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cstring>

#include "internal.hpp"

using namespace sygraph::hipdev;

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
/* Position ctr of the SplitMix64 stream that starts at mix64(seed + gamma). */
__host__ __device__ __forceinline__ uint64_t rand64(uint64_t seed, uint64_t ctr) {
  uint64_t s0 = mix64(seed + kGamma);
  return mix64(s0 + (ctr + 1) * kGamma);
}

struct Relabel {
  uint64_t mask;
  int sh;
  uint64_t mul[3], add[3];
};

Relabel make_relabel(int scale, uint64_t seed) {
  Relabel r{};
  r.mask = scale >= 32 ? 0xFFFFFFFFull : ((1ull << scale) - 1ull);
  r.sh = (scale + 1) / 2;
  for (int k = 0; k < 3; ++k) {
    r.mul[k] = rand64(seed ^ 0x5EEDF00Dull, 2 * k) | 1ull;
    r.add[k] = rand64(seed ^ 0x5EEDF00Dull, 2 * k + 1);
  }
  return r;
}

__device__ __forceinline__ uint32_t relabel(uint32_t x0, const Relabel& r) {
  uint64_t x = x0 & r.mask;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x = (x * r.mul[k]) & r.mask;
    x ^= (x >> r.sh);
    x = (x + r.add[k]) & r.mask;
  }
  return static_cast<uint32_t>(x);
}

/* One RMAT tuple per thread; writes both arc keys (src<<b | dst), or the
 * sentinel for self loops. */
__global__ void rmat_keys_kernel(int scale, int64_t ne, uint64_t ta, uint64_t tb, uint64_t tc, uint64_t seed,
                                 int do_relabel, Relabel rl, int b, uint64_t sentinel, uint64_t* keys) {
  const int draws = (scale + 1) / 2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t su = 0, sv = 0;
    for (int d = 0; d < draws; ++d) {
      uint64_t r = rand64(seed, (uint64_t)e * (uint64_t)draws + (uint64_t)d);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int lvl = 2 * d + h;
        if (lvl < scale) {
          uint64_t r32 = h == 0 ? (r >> 32) : (r & 0xFFFFFFFFull);
          uint32_t bu = r32 >= tc ? 1u : (r32 >= tb ? 1u : 0u);
          uint32_t bv = r32 >= tc ? 1u : ((r32 >= ta && r32 < tb) ? 1u : 0u);
          su |= bu << (scale - 1 - lvl);
          sv |= bv << (scale - 1 - lvl);
        }
      }
    }
    if (do_relabel) {
      su = relabel(su, rl);
      sv = relabel(sv, rl);
    }
    if (su == sv) {
      keys[2 * e] = sentinel;
      keys[2 * e + 1] = sentinel;
    } else {
      keys[2 * e] = ((uint64_t)su << b) | sv;
      keys[2 * e + 1] = ((uint64_t)sv << b) | su;
    }
  }
}

/* Grid edges: index < H are horizontal (r,c)-(r,c+1), the rest vertical. */
__global__ void grid_keys_kernel(int rows, int cols, uint64_t thr, int do_delete, uint64_t seed, int b,
                                 uint64_t sentinel, uint64_t* keys) {
  int64_t H = (int64_t)rows * (cols - 1);
  int64_t total = H + (int64_t)(rows - 1) * cols;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    uint64_t r = rand64(seed, (uint64_t)idx);
    if (do_delete && r < thr) {
      keys[2 * idx] = sentinel;
      keys[2 * idx + 1] = sentinel;
      continue;
    }
    uint32_t a, c;
    if (idx < H) {
      int64_t rr = idx / (cols - 1), cc = idx % (cols - 1);
      a = (uint32_t)(rr * cols + cc);
      c = a + 1;
    } else {
      int64_t j = idx - H;
      a = (uint32_t)j;
      c = a + (uint32_t)cols;
    }
    keys[2 * idx] = ((uint64_t)a << b) | c;
    keys[2 * idx + 1] = ((uint64_t)c << b) | a;
  }
}

__global__ void er_keys_kernel(int64_t n, int64_t m, uint64_t seed, int b, uint64_t sentinel, uint64_t* keys) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t u = (uint32_t)__umul64hi(rand64(seed, 2 * (uint64_t)e), (uint64_t)n);
    uint32_t v = (uint32_t)__umul64hi(rand64(seed, 2 * (uint64_t)e + 1), (uint64_t)n);
    if (u == v) {
      keys[2 * e] = sentinel;
      keys[2 * e + 1] = sentinel;
    } else {
      keys[2 * e] = ((uint64_t)u << b) | v;
      keys[2 * e + 1] = ((uint64_t)v << b) | u;
    }
  }
}

__global__ void mark_unique_kernel(const uint64_t* keys, int64_t k, uint64_t sentinel, uint32_t* flags) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t x = keys[i];
    flags[i] = (x != sentinel && (i == 0 || keys[i - 1] != x)) ? 1u : 0u;
  }
}

__global__ void scatter_unique_kernel(const uint64_t* keys, const uint32_t* flags, const uint64_t* pos, int64_t k,
                                      uint64_t* uniq) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x)
    if (flags[i]) uniq[pos[i]] = keys[i];
}

/* offsets[r] = index of the first unique key with row >= r. */
template<typename OffT>
__global__ void offsets_from_sorted_kernel(const uint64_t* uniq, int64_t m, int64_t n, int b, OffT* offsets,
                                           int32_t* cols) {
  const uint64_t mask = (1ull << b) - 1ull;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j <= m; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t row = j < m ? (int64_t)(uniq[j] >> b) : n;
    int64_t prev = j > 0 ? (int64_t)(uniq[j - 1] >> b) : -1;
    for (int64_t r = prev + 1; r <= row; ++r) offsets[r] = (OffT)j;
    if (j < m) cols[j] = (int32_t)(uniq[j] & mask);
  }
}

__device__ __forceinline__ uint32_t edge_weight(uint32_t u, uint32_t v, uint64_t seed) {
  uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
  uint64_t r = rand64(seed, ((uint64_t)lo << 32) | hi);
  return 1u + (uint32_t)(((r >> 32) * 255ull) >> 32);
}

/* f(u, e) for every arc e of every vertex u: a wave takes 64 consecutive
 * vertices, a lane walks a short list, the whole wave a list of more than 64
 * arcs (so hubs do not serialise a lane).  All lanes of the wave call. */
template<typename OffT, typename F>
__device__ __forceinline__ void for_each_arc_wave(const OffT* off, int64_t n, F&& f) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t base = gwave * 64; base < n; base += nwaves * 64) {
    const int64_t u = base + lane;
    int64_t b = 0, e = 0;
    if (u < n) {
      b = (int64_t)off[u];
      e = (int64_t)off[u + 1];
    }
    const bool coop = e - b > 64;
    if (!coop)
      for (int64_t j = b; j < e; ++j) f(u, j);
    uint64_t heavy = ballot(coop);
    while (heavy) {
      const int k = __ffsll((unsigned long long)heavy) - 1;
      heavy &= heavy - 1;
      const int64_t kb = (int64_t)__shfl((long long)b, k, 64), ke = (int64_t)__shfl((long long)e, k, 64);
      const int64_t ku = base + k;
      for (int64_t j = kb + lane; j < ke; j += 64) f(ku, j);
    }
  }
}

template<typename OffT>
__global__ void __launch_bounds__(256) weights_kernel(const OffT* offsets, const int32_t* cols, int64_t n,
                                                      uint64_t seed, uint32_t* w) {
  for_each_arc_wave(offsets, n, [&](int64_t u, int64_t e) { w[e] = edge_weight((uint32_t)u, (uint32_t)cols[e], seed); });
}

/* Transpose helpers: key = (dst << 32) | src for every arc. */
template<typename OffT>
__global__ void __launch_bounds__(256) transpose_keys_kernel(const OffT* offsets, const int32_t* cols, int64_t n,
                                                             uint64_t* keys, uint64_t* idx) {
  for_each_arc_wave(offsets, n, [&](int64_t u, int64_t e) {
    keys[e] = ((uint64_t)(uint32_t)cols[e] << 32) | (uint64_t)u;
    idx[e] = (uint64_t)e;
  });
}

template<typename OffT>
__global__ void transpose_fill_kernel(const uint64_t* keys, const uint64_t* idx, int64_t m, int64_t n,
                                      const uint32_t* w_in, OffT* inv_off, int32_t* inv_cols, uint32_t* inv_w) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j <= m; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t row = j < m ? (int64_t)(keys[j] >> 32) : n;
    int64_t prev = j > 0 ? (int64_t)(keys[j - 1] >> 32) : -1;
    for (int64_t r = prev + 1; r <= row; ++r) inv_off[r] = (OffT)j;
    if (j < m) {
      inv_cols[j] = (int32_t)(keys[j] & 0xFFFFFFFFull);
      if (w_in) inv_w[j] = w_in[idx[j]];
    }
  }
}

template<typename OffT>
__global__ void isolated_kernel(const OffT* offsets, int64_t n, int64_t w0, uint64_t* bm) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < w0; w += (int64_t)gridDim.x * blockDim.x) {
    uint64_t bits = 0;
    for (int i = 0; i < 64; ++i) {
      int64_t v = w * 64 + i;
      if (v < n && offsets[v + 1] == offsets[v]) bits |= 1ull << i;
    }
    bm[w] = bits;
  }
}

/* hint[v] = {in-neighbour of v with the largest out-degree (lowest id on
 * ties, -1 without in-arcs), out-degree of v}: one 8-byte load gives a pull
 * its first probe and, on a hit, the degree it needs, without the offsets.
 * A wave takes 64 consecutive vertices; lists longer than 64 arcs are reduced
 * by the whole wave.  Candidates are ranked by the key
 * (degree << 32) | ~id, so the maximum key is the answer. */
template<typename OffT>
__global__ void __launch_bounds__(256) pull_hint_kernel(const OffT* in_off, const int32_t* in_cols, const OffT* out_off,
                                                        int64_t n, int2* hint, int32_t with_offsets) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  auto key_of = [&](int32_t u) -> uint64_t {
    const uint64_t d = (uint64_t)(out_off[u + 1] - out_off[u]);
    return (d << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)u);
  };
  for (int64_t base = gwave * 64; base < n; base += nwaves * 64) {
    const int64_t v = base + lane;
    int64_t b = 0, e = 0;
    if (v < n) {
      b = (int64_t)in_off[v];
      e = (int64_t)in_off[v + 1];
    }
    const bool coop = e - b > 64;
    uint64_t best = 0;
    if (!coop)
      for (int64_t j = b; j < e; ++j) {
        const uint64_t k = key_of(in_cols[j]);
        best = k > best ? k : best;
      }
    uint64_t heavy = ballot(coop);
    while (heavy) {
      const int k2 = __ffsll((unsigned long long)heavy) - 1;
      heavy &= heavy - 1;
      const int64_t kb = (int64_t)__shfl((long long)b, k2, 64), ke = (int64_t)__shfl((long long)e, k2, 64);
      uint64_t part = 0;
      for (int64_t j = kb + lane; j < ke; j += 64) {
        const uint64_t k = key_of(in_cols[j]);
        part = k > part ? k : part;
      }
      part = wave_max_u64(part);
      if (lane == k2) best = part;
    }
    if (v < n)
      hint[v] = make_int2(e > b ? (int32_t)(0xFFFFFFFFu - (uint32_t)(best & 0xFFFFFFFFull)) : -1,
                          with_offsets ? (int32_t)(uint32_t)b : (int32_t)(out_off[v + 1] - out_off[v]));
    if (with_offsets && v == n - 1) hint[n] = make_int2(-1, (int32_t)(uint32_t)e); /* the end of the last list */
  }
}

template<typename T>
__global__ void digest_kernel(const T* x, int64_t n, unsigned long long* out) {
  uint64_t acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t xv = sizeof(T) == 8 ? (uint64_t)x[i] : (uint64_t)(uint32_t)x[i];
    acc += mix64(((uint64_t)i * kGamma) ^ xv);
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0) atomicAdd(out, (unsigned long long)acc);
}

template<typename OffT>
__global__ void reached_kernel(const OffT* offsets, const int32_t* depth, int64_t n, unsigned long long* out) {
  uint64_t e = 0, c = 0;
  const int32_t unreached = (int32_t)(n + 1);
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    if (depth[v] != unreached) {
      e += (uint64_t)(offsets[v + 1] - offsets[v]);
      c += 1;
    }
  e = wave_sum_u64(e);
  c = wave_sum_u64(c);
  if (lane_id() == 0) {
    atomicAdd(&out[0], (unsigned long long)e);
    atomicAdd(&out[1], (unsigned long long)c);
  }
}

int bits_for(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

void free_graph_arrays(sg_graph_t g) {
  if (!g) return;
  (void)hipFree(g->offsets);
  (void)hipFree(g->cols);
  (void)hipFree(g->weights);
  if (g->owns_inverse) {
    (void)hipFree(g->inv_offsets);
    (void)hipFree(g->inv_cols);
    (void)hipFree(g->inv_weights);
  }
  (void)hipFree(g->isolated);
  (void)hipFree(g->pull_hint);
  g->pull_hint = nullptr;
  (void)hipFree(g->sssp_cols);
  (void)hipFree(g->sssp_w);
  (void)hipFree(g->sssp_light);
  g->sssp_cols = nullptr;
  g->sssp_w = nullptr;
  g->sssp_light = nullptr;
  g->sssp_split_valid = false;
  g->offsets = g->cols = nullptr;
  g->weights = g->inv_offsets = g->inv_weights = g->isolated = nullptr;
  g->inv_cols = nullptr;
}

/* Device-side inverse CSR by a radix sort of (dst, src) keys. */
template<typename OffT>
int build_inverse(sg_graph_t g) {
  sg_device_t dev = g->dev;
  int64_t m = g->m, n = g->n;
  uint64_t *keys = nullptr, *keys2 = nullptr, *idx = nullptr, *idx2 = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int st = SG_OK;
  auto fail = [&](int s) {
    (void)hipFree(keys);
    (void)hipFree(keys2);
    (void)hipFree(idx);
    (void)hipFree(idx2);
    (void)hipFree(tmp);
    return s;
  };
  if ((st = sg::alloc_device(dev, (void**)&keys, 8 * (m + 1))) || (st = sg::alloc_device(dev, (void**)&keys2, 8 * (m + 1))) ||
      (st = sg::alloc_device(dev, (void**)&idx, 8 * (m + 1))) || (st = sg::alloc_device(dev, (void**)&idx2, 8 * (m + 1))))
    return fail(st);
  if ((st = sg::alloc_device(dev, &g->inv_offsets, (size_t)g->offset_bytes * (n + 1))) ||
      (st = sg::alloc_device(dev, (void**)&g->inv_cols, 4 * (m + 1))))
    return fail(st);
  if (g->weights && (st = sg::alloc_device(dev, &g->inv_weights, 4 * (m + 1)))) return fail(st);
  g->owns_inverse = true;
  int grid = sg::grid_for(dev, n, 256);
  transpose_keys_kernel<OffT><<<grid, 256, 0, dev->stream>>>((const OffT*)g->offsets, g->cols, n, keys, idx);
  int bits = 32 + bits_for(n);
  if (rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, idx, idx2, (size_t)m, 0, bits, dev->stream) != hipSuccess)
    return fail((sg::set_error("radix_sort_pairs sizing failed"), SG_ERR_HIP));
  if ((st = sg::alloc_device(dev, &tmp, tmp_bytes))) return fail(st);
  if (rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, idx, idx2, (size_t)m, 0, bits, dev->stream) != hipSuccess)
    return fail((sg::set_error("radix_sort_pairs failed"), SG_ERR_HIP));
  transpose_fill_kernel<OffT><<<sg::grid_for(dev, m + 1, 256), 256, 0, dev->stream>>>(
      keys2, idx2, m, n, (const uint32_t*)g->weights, (OffT*)g->inv_offsets, g->inv_cols, (uint32_t*)g->inv_weights);
  hipError_t e = hipStreamSynchronize(dev->stream);
  fail(SG_OK);
  if (e != hipSuccess) {
    sg::set_error("build_inverse: %s", hipGetErrorString(e));
    return SG_ERR_HIP;
  }
  return SG_OK;
}

/* sort + unique of arc keys -> CSR with int64 offsets (keys is consumed). */
int csr_from_keys(sg_graph_t g, uint64_t* keys, int64_t k, int b) {
  sg_device_t dev = g->dev;
  const uint64_t sentinel = (b >= 32) ? ~0ull : ((1ull << (2 * b)) - 1ull);
  uint64_t *sorted = nullptr, *pos = nullptr, *uniq = nullptr;
  uint32_t* flags = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0, t2 = 0;
  int st = SG_OK;
  auto cleanup = [&]() {
    (void)hipFree(sorted);
    (void)hipFree(pos);
    (void)hipFree(uniq);
    (void)hipFree(flags);
    (void)hipFree(tmp);
  };
  if ((st = sg::alloc_device(dev, (void**)&sorted, 8 * (k + 1)))) return cleanup(), st;
  if (rocprim::radix_sort_keys(nullptr, tmp_bytes, keys, sorted, (size_t)k, 0, 2 * b, dev->stream) != hipSuccess)
    return cleanup(), sg::set_error("radix_sort_keys sizing failed"), SG_ERR_HIP;
  if (rocprim::exclusive_scan(nullptr, t2, (uint32_t*)nullptr, (uint64_t*)nullptr, (uint64_t)0, (size_t)k,
                              rocprim::plus<uint64_t>(), dev->stream) != hipSuccess)
    return cleanup(), sg::set_error("exclusive_scan sizing failed"), SG_ERR_HIP;
  if (t2 > tmp_bytes) tmp_bytes = t2;
  if ((st = sg::alloc_device(dev, &tmp, tmp_bytes))) return cleanup(), st;
  if (rocprim::radix_sort_keys(tmp, tmp_bytes, keys, sorted, (size_t)k, 0, 2 * b, dev->stream) != hipSuccess)
    return cleanup(), sg::set_error("radix_sort_keys failed"), SG_ERR_HIP;
  (void)hipFree(keys);
  if ((st = sg::alloc_device(dev, (void**)&flags, 4 * (k + 1))) || (st = sg::alloc_device(dev, (void**)&pos, 8 * (k + 1))))
    return cleanup(), st;
  int grid = sg::grid_for(dev, k, 256);
  mark_unique_kernel<<<grid, 256, 0, dev->stream>>>(sorted, k, sentinel, flags);
  if (rocprim::exclusive_scan(tmp, tmp_bytes, flags, pos, (uint64_t)0, (size_t)k, rocprim::plus<uint64_t>(),
                              dev->stream) != hipSuccess)
    return cleanup(), sg::set_error("exclusive_scan failed"), SG_ERR_HIP;
  uint64_t last_pos = 0;
  uint32_t last_flag = 0;
  if (k > 0) {
    SG_HIP(hipMemcpyAsync(&last_pos, pos + k - 1, 8, hipMemcpyDeviceToHost, dev->stream));
    SG_HIP(hipMemcpyAsync(&last_flag, flags + k - 1, 4, hipMemcpyDeviceToHost, dev->stream));
  }
  SG_HIP(hipStreamSynchronize(dev->stream));
  int64_t m = (int64_t)(last_pos + last_flag);
  if ((st = sg::alloc_device(dev, (void**)&uniq, 8 * (m + 1)))) return cleanup(), st;
  scatter_unique_kernel<<<grid, 256, 0, dev->stream>>>(sorted, flags, pos, k, uniq);
  (void)hipFree(sorted);
  sorted = nullptr;
  (void)hipFree(pos);
  pos = nullptr;
  (void)hipFree(flags);
  flags = nullptr;
  g->m = m;
  g->offset_bytes = 8;
  if ((st = sg::alloc_device(dev, &g->offsets, 8 * (g->n + 1))) || (st = sg::alloc_device(dev, (void**)&g->cols, 4 * (m + 1))))
    return cleanup(), st;
  offsets_from_sorted_kernel<int64_t><<<sg::grid_for(dev, m + 1, 256), 256, 0, dev->stream>>>(uniq, m, g->n, b,
                                                                                                 (int64_t*)g->offsets, g->cols);
  hipError_t e = hipStreamSynchronize(dev->stream);
  cleanup();
  if (e != hipSuccess) {
    sg::set_error("csr_from_keys: %s", hipGetErrorString(e));
    return SG_ERR_HIP;
  }
  g->inv_offsets = g->offsets;
  g->inv_cols = g->cols;
  return SG_OK;
}

int attach_weights(sg_graph_t g, uint64_t weight_seed) {
  if (weight_seed == 0) return SG_OK;
  SG_TRY(sg::alloc_device(g->dev, &g->weights, 4 * (g->m + 1)));
  weights_kernel<int64_t><<<sg::grid_for(g->dev, g->n, 256), 256, 0, g->dev->stream>>>(
      (const int64_t*)g->offsets, g->cols, g->n, weight_seed, (uint32_t*)g->weights);
  SG_HIP(hipStreamSynchronize(g->dev->stream));
  g->inv_weights = g->weights;
  g->flags |= SG_GRAPH_WEIGHTED;
  return SG_OK;
}

} // namespace

namespace sg {
int graph_ensure_isolated(sg_graph_t g) {
  if (g->isolated) return SG_OK;
  int64_t w0 = (g->n + 63) / 64;
  SG_TRY(alloc_device(g->dev, (void**)&g->isolated, 8 * (w0 + 1)));
  int grid = grid_for(g->dev, w0, 256);
  /* a vertex with no in-edge can only ever be reached as the source */
  if (g->offset_bytes == 8)
    isolated_kernel<int64_t><<<grid, 256, 0, g->dev->stream>>>((const int64_t*)g->inv_offsets, g->n, w0, g->isolated);
  else
    isolated_kernel<uint32_t><<<grid, 256, 0, g->dev->stream>>>((const uint32_t*)g->inv_offsets, g->n, w0, g->isolated);
  SG_HIP(hipGetLastError());
  return SG_OK;
}
int graph_ensure_pull_hint(sg_graph_t g) {
  if (g->pull_hint) return SG_OK;
  SG_TRY(alloc_device(g->dev, (void**)&g->pull_hint, sizeof(int2) * (g->n + 1)));
  const int grid = grid_for(g->dev, (g->n + 3) / 4, 256);
  const void* in_off = g->inv_offsets ? g->inv_offsets : g->offsets;
  const int32_t* in_cols = g->inv_cols ? g->inv_cols : g->cols;
  /* an undirected graph with fewer than 2^32 arcs keeps the low 32 bits of
   * each in-list's start next to its hint (in- and out-lists are the same
   * lists, so the out-degree is the difference of two starts) */
  g->pull_hint_offsets = !(g->flags & SG_GRAPH_DIRECTED) && g->m < ((int64_t)1 << 32) && g->n > 0;
  const int32_t wo = g->pull_hint_offsets ? 1 : 0;
  if (g->offset_bytes == 8)
    pull_hint_kernel<int64_t><<<grid, 256, 0, g->dev->stream>>>((const int64_t*)in_off, in_cols,
                                                                (const int64_t*)g->offsets, g->n, g->pull_hint, wo);
  else
    pull_hint_kernel<uint32_t><<<grid, 256, 0, g->dev->stream>>>((const uint32_t*)in_off, in_cols,
                                                                 (const uint32_t*)g->offsets, g->n, g->pull_hint, wo);
  SG_HIP(hipGetLastError());
  return SG_OK;
}
} // namespace sg

namespace sg {
int graph_buffer(sg_graph_t g, const char* name, size_t bytes, void** out) {
  auto& e = g->buffers[name];
  if (e.first && e.second >= bytes) {
    *out = e.first;
    return SG_OK;
  }
  if (e.first) {
    SG_HIP(hipStreamSynchronize(g->dev->stream)); /* queued work may still use the old buffer */
    (void)hipFree(e.first);
    e.first = nullptr;
    e.second = 0;
  }
  SG_TRY(alloc_device(g->dev, &e.first, bytes > 0 ? bytes : 1));
  e.second = bytes;
  *out = e.first;
  return SG_OK;
}

void graph_buffers_release(sg_graph_t g) {
  for (auto& kv : g->buffers) (void)hipFree(kv.second.first);
  g->buffers.clear();
}
} // namespace sg

extern "C" {

int sg_graph_create(sg_device_t dev, int64_t n, int64_t m, const void* offsets, int offset_bytes, const int32_t* cols,
                    const void* weights, int flags, sg_graph_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_graph_create: null argument");
  *out = nullptr;
  SG_CHECK(n >= 0 && m >= 0, SG_ERR_INVALID, "sg_graph_create: negative size");
  SG_CHECK(n < (int64_t)INT32_MAX, SG_ERR_INVALID, "sg_graph_create: vertex count %lld exceeds 32-bit ids", (long long)n);
  SG_CHECK(offset_bytes == 4 || offset_bytes == 8, SG_ERR_INVALID, "sg_graph_create: offset_bytes must be 4 or 8");
  SG_CHECK(offset_bytes == 8 || m <= (int64_t)UINT32_MAX, SG_ERR_INVALID, "sg_graph_create: m too large for 4-byte offsets");
  SG_CHECK(offsets && (cols || m == 0), SG_ERR_INVALID, "sg_graph_create: null offsets/cols");
  auto* g = new sg_graph_s();
  g->dev = dev;
  g->n = n;
  g->m = m;
  g->offset_bytes = offset_bytes;
  g->flags = flags & ~SG_GRAPH_INPUT_DEVICE;
  if (weights) g->flags |= SG_GRAPH_WEIGHTED;
  hipMemcpyKind kind = (flags & SG_GRAPH_INPUT_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  auto alloc = [&](void** p, size_t bytes) -> int {
    if (bytes == 0) bytes = 8;
    hipError_t e;
    if (flags & SG_GRAPH_SPACE_SHARED)
      e = hipMallocManaged(p, bytes, hipMemAttachGlobal);
    else if (flags & SG_GRAPH_SPACE_HOST)
      e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    else
      e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      sg::set_error("sg_graph_create: allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
      return SG_ERR_NOMEM;
    }
    return SG_OK;
  };
  int st = SG_OK;
  if ((st = alloc(&g->offsets, (size_t)offset_bytes * (n + 1))) || (st = alloc((void**)&g->cols, 4 * (size_t)m)) ||
      (weights && (st = alloc(&g->weights, 4 * (size_t)m)))) {
    free_graph_arrays(g);
    delete g;
    return st;
  }
  hipError_t e = hipMemcpyAsync(g->offsets, offsets, (size_t)offset_bytes * (n + 1), kind, dev->stream);
  if (e == hipSuccess && m > 0) e = hipMemcpyAsync(g->cols, cols, 4 * (size_t)m, kind, dev->stream);
  if (e == hipSuccess && weights && m > 0) e = hipMemcpyAsync(g->weights, weights, 4 * (size_t)m, kind, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  if (e != hipSuccess) {
    sg::set_error("sg_graph_create: upload failed: %s", hipGetErrorString(e));
    free_graph_arrays(g);
    delete g;
    return SG_ERR_HIP;
  }
  if (flags & SG_GRAPH_DIRECTED) {
    st = offset_bytes == 8 ? build_inverse<int64_t>(g) : build_inverse<uint32_t>(g);
    if (st != SG_OK) {
      free_graph_arrays(g);
      delete g;
      return st;
    }
  } else {
    g->inv_offsets = g->offsets;
    g->inv_cols = g->cols;
    g->inv_weights = g->weights;
    g->owns_inverse = false;
  }
  g->total_degree = m;
  *out = g;
  return SG_OK;
}


int sg_graph_destroy(sg_graph_t g) {
  if (!g) return SG_OK;
  (void)hipStreamSynchronize(g->dev->stream);
  sg::bfs_workspace_release(g);
  sg::sssp_workspace_release(g);
  sg::graph_buffers_release(g);
  sg::aux_queues_release(g);
  free_graph_arrays(g);
  delete g;
  return SG_OK;
}

int sg_graph_prepare(sg_graph_t g, float* ms) {
  SG_CHECK(g, SG_ERR_INVALID, "sg_graph_prepare: null graph");
  SG_HIP(hipSetDevice(g->dev->ordinal));
  hipStream_t s = g->dev->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  SG_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    sg::set_error("sg_graph_prepare: event creation failed");
    return SG_ERR_HIP;
  }
  int st = hipEventRecord(e0, s) == hipSuccess ? SG_OK : SG_ERR_HIP;
  if (st == SG_OK) st = sg::graph_ensure_isolated(g);
  if (st == SG_OK) st = sg::graph_ensure_pull_hint(g);
  if (st == SG_OK && hipEventRecord(e1, s) != hipSuccess) st = SG_ERR_HIP;
  if (st == SG_OK && hipEventSynchronize(e1) != hipSuccess) st = SG_ERR_HIP;
  float t = 0.f;
  if (st == SG_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) st = SG_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st == SG_ERR_HIP) sg::set_error("sg_graph_prepare: event timing failed");
  if (st == SG_OK && ms) *ms = t;
  return st;
}

int sg_graph_get_view(sg_graph_t g, sg_graph_view* v) {
  SG_CHECK(g && v, SG_ERR_INVALID, "sg_graph_get_view: null argument");
  v->n = g->n;
  v->m = g->m;
  v->offset_bytes = g->offset_bytes;
  v->flags = g->flags;
  v->offsets = g->offsets;
  v->cols = g->cols;
  v->weights = g->weights;
  v->inv_offsets = g->inv_offsets;
  v->inv_cols = g->inv_cols;
  v->inv_weights = g->inv_weights;
  return SG_OK;
}

int sg_graph_generate_rmat(sg_device_t dev, int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                           int relabel, uint64_t weight_seed, sg_graph_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_graph_generate_rmat: null argument");
  SG_CHECK(scale >= 1 && scale <= 30, SG_ERR_INVALID, "sg_graph_generate_rmat: scale %d outside [1,30]", scale);
  SG_CHECK(edge_factor >= 1, SG_ERR_INVALID, "sg_graph_generate_rmat: edge_factor must be >= 1");
  SG_CHECK(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, SG_ERR_INVALID, "sg_graph_generate_rmat: bad probabilities");
  *out = nullptr;
  const int64_t n = (int64_t)1 << scale;
  const int64_t ne = (int64_t)edge_factor << scale;
  const int bb = bits_for(n);
  const uint64_t sentinel = (1ull << (2 * bb)) - 1ull;
  const double two32 = 4294967296.0;
  uint64_t* keys = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&keys, 8 * (2 * ne + 1)));
  Relabel rl = make_relabel(scale, seed);
  rmat_keys_kernel<<<sg::grid_for(dev, ne, 256), 256, 0, dev->stream>>>(
      scale, ne, (uint64_t)(a * two32), (uint64_t)((a + b) * two32), (uint64_t)((a + b + c) * two32), seed, relabel, rl, bb,
      sentinel, keys);
  SG_HIP(hipGetLastError());
  auto* g = new sg_graph_s();
  g->dev = dev;
  g->n = n;
  int st = csr_from_keys(g, keys, 2 * ne, bb);
  if (st == SG_OK) st = attach_weights(g, weight_seed);
  if (st != SG_OK) {
    free_graph_arrays(g);
    delete g;
    return st;
  }
  g->total_degree = g->m;
  *out = g;
  return SG_OK;
}

int sg_graph_generate_grid(sg_device_t dev, int rows, int cols, double p_delete, uint64_t seed, uint64_t weight_seed,
                           sg_graph_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_graph_generate_grid: null argument");
  SG_CHECK(rows >= 1 && cols >= 1 && (int64_t)rows * cols < (int64_t)INT32_MAX, SG_ERR_INVALID,
           "sg_graph_generate_grid: bad shape %dx%d", rows, cols);
  *out = nullptr;
  const int64_t n = (int64_t)rows * cols;
  const int64_t total = (int64_t)rows * (cols - 1) + (int64_t)(rows - 1) * cols;
  const int bb = bits_for(n);
  const uint64_t sentinel = (1ull << (2 * bb)) - 1ull;
  uint64_t thr = p_delete >= 1.0 ? ~0ull : (uint64_t)(p_delete * 18446744073709551616.0);
  uint64_t* keys = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&keys, 8 * (2 * total + 1)));
  grid_keys_kernel<<<sg::grid_for(dev, total, 256), 256, 0, dev->stream>>>(rows, cols, thr, p_delete > 0.0 ? 1 : 0, seed,
                                                                           bb, sentinel, keys);
  SG_HIP(hipGetLastError());
  auto* g = new sg_graph_s();
  g->dev = dev;
  g->n = n;
  int st = csr_from_keys(g, keys, 2 * total, bb);
  if (st == SG_OK) st = attach_weights(g, weight_seed);
  if (st != SG_OK) {
    free_graph_arrays(g);
    delete g;
    return st;
  }
  g->total_degree = g->m;
  *out = g;
  return SG_OK;
}

int sg_graph_generate_er(sg_device_t dev, int64_t n, int64_t m_pairs, uint64_t seed, uint64_t weight_seed,
                         sg_graph_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_graph_generate_er: null argument");
  SG_CHECK(n >= 1 && n < (int64_t)INT32_MAX && m_pairs >= 0, SG_ERR_INVALID, "sg_graph_generate_er: bad size");
  *out = nullptr;
  const int bb = bits_for(n);
  const uint64_t sentinel = (1ull << (2 * bb)) - 1ull;
  uint64_t* keys = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&keys, 8 * (2 * m_pairs + 1)));
  er_keys_kernel<<<sg::grid_for(dev, m_pairs, 256), 256, 0, dev->stream>>>(n, m_pairs, seed, bb, sentinel, keys);
  SG_HIP(hipGetLastError());
  auto* g = new sg_graph_s();
  g->dev = dev;
  g->n = n;
  int st = csr_from_keys(g, keys, 2 * m_pairs, bb);
  if (st == SG_OK) st = attach_weights(g, weight_seed);
  if (st != SG_OK) {
    free_graph_arrays(g);
    delete g;
    return st;
  }
  g->total_degree = g->m;
  *out = g;
  return SG_OK;
}

int sg_graph_download(sg_graph_t g, int64_t* offsets, int32_t* cols, uint32_t* weights) {
  SG_CHECK(g, SG_ERR_INVALID, "sg_graph_download: null graph");
  sg_device_t dev = g->dev;
  if (offsets) {
    if (g->offset_bytes == 8) {
      SG_HIP(hipMemcpyAsync(offsets, g->offsets, 8 * (g->n + 1), hipMemcpyDeviceToHost, dev->stream));
    } else {
      std::vector<uint32_t> tmp(g->n + 1);
      SG_HIP(hipMemcpyAsync(tmp.data(), g->offsets, 4 * (g->n + 1), hipMemcpyDeviceToHost, dev->stream));
      SG_HIP(hipStreamSynchronize(dev->stream));
      for (int64_t i = 0; i <= g->n; ++i) offsets[i] = tmp[i];
    }
  }
  if (cols && g->m > 0) SG_HIP(hipMemcpyAsync(cols, g->cols, 4 * g->m, hipMemcpyDeviceToHost, dev->stream));
  if (weights && g->m > 0) {
    SG_CHECK(g->weights, SG_ERR_STATE, "sg_graph_download: graph has no weights");
    SG_HIP(hipMemcpyAsync(weights, g->weights, 4 * g->m, hipMemcpyDeviceToHost, dev->stream));
  }
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_digest(sg_device_t dev, const void* ptr, int64_t count, int elem_bytes, uint64_t* digest) {
  SG_CHECK(dev && digest && (ptr || count == 0) && count >= 0, SG_ERR_INVALID, "sg_digest: bad argument");
  SG_CHECK(elem_bytes == 4 || elem_bytes == 8, SG_ERR_INVALID, "sg_digest: element size %d not 4 or 8", elem_bytes);
  unsigned long long* acc = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&acc, 8));
  SG_TRY(sg::zero_async(dev, acc, 8));
  if (count > 0) {
    if (elem_bytes == 8)
      digest_kernel<int64_t><<<sg::grid_for(dev, count, 256), 256, 0, dev->stream>>>((const int64_t*)ptr, count, acc);
    else
      digest_kernel<uint32_t><<<sg::grid_for(dev, count, 256), 256, 0, dev->stream>>>((const uint32_t*)ptr, count, acc);
  }
  unsigned long long h = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  (void)hipFree(acc);
  SG_HIP(e);
  *digest = h;
  return SG_OK;
}

int sg_graph_digest(sg_graph_t g, uint64_t* offsets_digest, uint64_t* cols_digest) {
  SG_CHECK(g && offsets_digest && cols_digest, SG_ERR_INVALID, "sg_graph_digest: null argument");
  sg_device_t dev = g->dev;
  unsigned long long* acc = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&acc, 16));
  SG_TRY(sg::zero_async(dev, acc, 16));
  if (g->offset_bytes == 8)
    digest_kernel<int64_t><<<sg::grid_for(dev, g->n + 1, 256), 256, 0, dev->stream>>>((const int64_t*)g->offsets, g->n + 1, acc);
  else
    digest_kernel<uint32_t><<<sg::grid_for(dev, g->n + 1, 256), 256, 0, dev->stream>>>((const uint32_t*)g->offsets, g->n + 1, acc);
  digest_kernel<int32_t><<<sg::grid_for(dev, g->m, 256), 256, 0, dev->stream>>>(g->cols, g->m, acc + 1);
  unsigned long long h[2];
  hipError_t e = hipMemcpyAsync(h, acc, 16, hipMemcpyDeviceToHost, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  (void)hipFree(acc);
  SG_HIP(e);
  *offsets_digest = h[0];
  *cols_digest = h[1];
  return SG_OK;
}

int sg_graph_reached_edges(sg_graph_t g, const int32_t* depth, int64_t* edges, int64_t* vertices) {
  SG_CHECK(g && depth, SG_ERR_INVALID, "sg_graph_reached_edges: null argument");
  sg_device_t dev = g->dev;
  unsigned long long* acc = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&acc, 16));
  SG_TRY(sg::zero_async(dev, acc, 16));
  int grid = sg::grid_for(dev, g->n, 256);
  if (g->offset_bytes == 8)
    reached_kernel<int64_t><<<grid, 256, 0, dev->stream>>>((const int64_t*)g->offsets, depth, g->n, acc);
  else
    reached_kernel<uint32_t><<<grid, 256, 0, dev->stream>>>((const uint32_t*)g->offsets, depth, g->n, acc);
  unsigned long long h[2];
  hipError_t e = hipMemcpyAsync(h, acc, 16, hipMemcpyDeviceToHost, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  (void)hipFree(acc);
  SG_HIP(e);
  if (edges) *edges = (int64_t)h[0];
  if (vertices) *vertices = (int64_t)h[1];
  return SG_OK;
}

} // extern "C"
