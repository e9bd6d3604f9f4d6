/*
 * frontier.hip -- the multi-level bitmap (MLB) frontier on the device.
 *
 * Reference: frontier/impls/mlb_frontier.hpp:155-449 (FrontierMLB<T, 2>).
 * Level 0: one bit per element in uint64 words (one word per wavefront).
 * Level 1: one bit per level-0 word, set when that word may be non-empty.
 *
 * computeActiveFrontier (mlb_frontier.hpp:388-442) becomes an ordered
 * set-scan-compact in three passes: per-tile counts (one 256-word tile per
 * workgroup, empty tiles skipped through level 1), one exclusive scan of the
 * tile counts, and a write pass in which every thread emits the ids of its
 * word at its scanned position.  Output is ascending and deterministic (the
 * reference's order depends on atomic timing).  The count stays on the device
 * (list_size), so the advance that follows needs no host round trip.
 */
#include <utility>

#include "internal.hpp"

using namespace sygraph::hipdev;

namespace {

constexpr int kTile = 256; /* level-0 words per compaction tile (one per thread) */

/* per-thread item count for word w under mode 0 (bits), 1 (non-empty words),
 * 2 (words that are not all ones). */
__device__ __forceinline__ uint32_t item_count(const uint64_t* l0, const uint64_t* l1, int64_t w, int64_t w0, int mode,
                                               uint64_t* word_out) {
  if (w >= w0) {
    *word_out = 0;
    return 0;
  }
  if (mode != 2 && !((l1[w >> 6] >> (w & 63)) & 1ull)) {
    *word_out = 0;
    return 0;
  }
  uint64_t x = l0[w];
  *word_out = x;
  if (mode == 0) return (uint32_t)__popcll(x);
  if (mode == 1) return x != 0 ? 1u : 0u;
  return x != ~0ull ? 1u : 0u;
}

/* Block-wide exclusive scan of one uint32 per thread (blockDim == 256). */
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wave_tot[4];
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o, 64);
    if (lane_id() >= o) x += t;
  }
  int wv = threadIdx.x >> 6;
  if (lane_id() == 63) wave_tot[wv] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t t = wave_tot[i];
    if (i < wv) base += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__global__ void __launch_bounds__(256) compact_count_kernel(const uint64_t* l0, const uint64_t* l1, int64_t w0, int mode,
                                                            uint32_t* tile_counts) {
  int64_t w = (int64_t)blockIdx.x * kTile + threadIdx.x;
  uint64_t word;
  uint32_t c = item_count(l0, l1, w, w0, mode, &word);
  uint32_t tot;
  block_exclusive_scan(c, &tot);
  if (threadIdx.x == 0) tile_counts[blockIdx.x] = tot;
}

/* Single-workgroup exclusive scan of the tile counts (in place); writes the
 * grand total to *total. */
__global__ void __launch_bounds__(256) compact_scan_kernel(uint32_t* tile_counts, int64_t ntiles, uint32_t* total) {
  uint32_t carry = 0;
  for (int64_t base = 0; base < ntiles; base += 256) {
    int64_t i = base + threadIdx.x;
    uint32_t v = i < ntiles ? tile_counts[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, &tot);
    if (i < ntiles) tile_counts[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) compact_write_kernel(const uint64_t* l0, const uint64_t* l1, int64_t w0, int64_t n,
                                                            int mode, const uint32_t* tile_offsets, int32_t* list) {
  int64_t w = (int64_t)blockIdx.x * kTile + threadIdx.x;
  uint64_t word;
  uint32_t c = item_count(l0, l1, w, w0, mode, &word);
  uint32_t tot;
  uint32_t pos = tile_offsets[blockIdx.x] + block_exclusive_scan(c, &tot);
  if (c == 0) return;
  if (mode == 0) {
    while (word) {
      int b = __builtin_ctzll(word);
      word &= word - 1;
      int64_t v = w * 64 + b;
      if (v < n) list[pos++] = (int32_t)v;
    }
  } else {
    list[pos] = (int32_t)w;
  }
}

/* Rebuild level 1 from level 0 after the elementwise op: one wave per
 * level-1 word, lane i combines level-0 word 64*t+i. op: 0 or, 1 and, 2 copy-and. */
__global__ void __launch_bounds__(256) combine_kernel(uint64_t* out0, uint64_t* out1, const uint64_t* a, const uint64_t* b,
                                                      int64_t w0, int64_t w1, int op) {
  int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (; t < w1; t += stride) {
    int64_t w = t * 64 + lane_id();
    uint64_t r = 0;
    if (w < w0) {
      uint64_t x = a[w], y = b[w];
      r = op == 0 ? (x | y) : (x & y);
      out0[w] = r;
    }
    uint64_t m = ballot(r != 0);
    if (lane_id() == 0) out1[t] = m;
  }
}

__global__ void any_kernel(const uint64_t* l1, int64_t w1, uint32_t* flag) {
  bool any = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < w1; i += (int64_t)gridDim.x * blockDim.x)
    any |= l1[i] != 0;
  if (ballot(any) != 0 && lane_id() == 0) atomic_or(flag, 1u);
}

__global__ void popcount_kernel(const uint64_t* l0, int64_t w0, unsigned long long* out) {
  uint64_t c = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < w0; i += (int64_t)gridDim.x * blockDim.x)
    c += (uint64_t)__popcll(l0[i]);
  c = wave_sum_u64(c);
  if (lane_id() == 0) atomicAdd(out, (unsigned long long)c);
}

/* MLBDevice::insert / remove (mlb_frontier.hpp:86-103) for one element. */
__global__ void insert_one_kernel(uint64_t* l0, uint64_t* l1, int64_t idx) {
  uint64_t b = 1ull << (idx & 63);
  uint64_t old = atomic_or(&l0[idx >> 6], b);
  if (old == 0) atomic_or(&l1[idx >> 12], 1ull << ((idx >> 6) & 63));
}
__global__ void remove_one_kernel(uint64_t* l0, int64_t idx) { atomic_and(&l0[idx >> 6], ~(1ull << (idx & 63))); }

} // namespace

namespace sg {
int frontier_compact_launch(sg_frontier_t f, int mode) {
  sg_device_t dev = f->dev;
  if (f->ntiles == 0) {
    return zero_async(dev, f->list_size, 4);
  }
  compact_count_kernel<<<(unsigned)f->ntiles, 256, 0, dev->stream>>>(f->l0, f->l1, f->w0, mode, f->tile_counts);
  compact_scan_kernel<<<1, 256, 0, dev->stream>>>(f->tile_counts, f->ntiles, f->list_size);
  compact_write_kernel<<<(unsigned)f->ntiles, 256, 0, dev->stream>>>(f->l0, f->l1, f->w0, f->n, mode, f->tile_counts,
                                                                      f->list);
  SG_HIP(hipGetLastError());
  return SG_OK;
}
} // namespace sg

extern "C" {

int sg_frontier_create(sg_device_t dev, int64_t n, sg_frontier_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_frontier_create: null argument");
  SG_CHECK(n >= 0 && n < (int64_t)INT32_MAX, SG_ERR_INVALID, "sg_frontier_create: bad size %lld", (long long)n);
  *out = nullptr;
  auto* f = new sg_frontier_s();
  f->dev = dev;
  f->n = n;
  f->w0 = (n + 63) / 64;
  f->w1 = (f->w0 + 63) / 64;
  f->ntiles = (f->w0 + kTile - 1) / kTile;
  int st = SG_OK;
  if ((st = sg::alloc_device(dev, (void**)&f->l0, 8 * (f->w0 + 1))) || (st = sg::alloc_device(dev, (void**)&f->l1, 8 * (f->w1 + 1))) ||
      (st = sg::alloc_device(dev, (void**)&f->list, 4 * (n + 1))) || (st = sg::alloc_device(dev, (void**)&f->list_size, 8)) ||
      (st = sg::alloc_device(dev, (void**)&f->tile_counts, 4 * (f->ntiles + 1)))) {
    sg_frontier_destroy(f);
    return st;
  }
  if ((st = sg_frontier_clear(f)) != SG_OK) {
    sg_frontier_destroy(f);
    return st;
  }
  *out = f;
  return SG_OK;
}

int sg_frontier_destroy(sg_frontier_t f) {
  if (!f) return SG_OK;
  (void)hipStreamSynchronize(f->dev->stream);
  (void)hipFree(f->l0);
  (void)hipFree(f->l1);
  (void)hipFree(f->list);
  (void)hipFree(f->list_size);
  (void)hipFree(f->tile_counts);
  delete f;
  return SG_OK;
}

int sg_frontier_get_view(sg_frontier_t f, sg_frontier_view* v) {
  SG_CHECK(f && v, SG_ERR_INVALID, "sg_frontier_get_view: null argument");
  v->n = f->n;
  v->words0 = f->w0;
  v->words1 = f->w1;
  v->l0 = f->l0;
  v->l1 = f->l1;
  v->list = f->list;
  v->list_size = f->list_size;
  return SG_OK;
}

int sg_frontier_check(sg_frontier_t f, int64_t idx, int* present) {
  SG_CHECK(f && present, SG_ERR_INVALID, "sg_frontier_check: null argument");
  SG_CHECK(idx >= 0 && idx < f->n, SG_ERR_INVALID, "sg_frontier_check: index %lld out of range", (long long)idx);
  uint64_t w = 0;
  SG_HIP(hipMemcpyAsync(&w, f->l0 + (idx >> 6), 8, hipMemcpyDeviceToHost, f->dev->stream));
  SG_HIP(hipStreamSynchronize(f->dev->stream));
  *present = (int)((w >> (idx & 63)) & 1ull);
  return SG_OK;
}

int sg_frontier_insert(sg_frontier_t f, int64_t idx) {
  SG_CHECK(f, SG_ERR_INVALID, "sg_frontier_insert: null frontier");
  SG_CHECK(idx >= 0 && idx < f->n, SG_ERR_INVALID, "sg_frontier_insert: index %lld out of range", (long long)idx);
  insert_one_kernel<<<1, 1, 0, f->dev->stream>>>(f->l0, f->l1, idx);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(f->dev->stream));
  return SG_OK;
}

int sg_frontier_remove(sg_frontier_t f, int64_t idx) {
  SG_CHECK(f, SG_ERR_INVALID, "sg_frontier_remove: null frontier");
  SG_CHECK(idx >= 0 && idx < f->n, SG_ERR_INVALID, "sg_frontier_remove: index %lld out of range", (long long)idx);
  remove_one_kernel<<<1, 1, 0, f->dev->stream>>>(f->l0, idx);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(f->dev->stream));
  return SG_OK;
}

int sg_frontier_clear_async(sg_frontier_t f) {
  SG_CHECK(f, SG_ERR_INVALID, "sg_frontier_clear: null frontier");
  SG_TRY(sg::zero_async(f->dev, f->l0, 8 * f->w0));
  SG_TRY(sg::zero_async(f->dev, f->l1, 8 * f->w1));
  SG_TRY(sg::zero_async(f->dev, f->list_size, 8));
  return SG_OK;
}

int sg_frontier_clear(sg_frontier_t f) {
  SG_TRY(sg_frontier_clear_async(f));
  SG_HIP(hipStreamSynchronize(f->dev->stream));
  return SG_OK;
}

int sg_frontier_empty(sg_frontier_t f, int* empty) {
  SG_CHECK(f && empty, SG_ERR_INVALID, "sg_frontier_empty: null argument");
  sg_device_t dev = f->dev;
  uint32_t* flag = (uint32_t*)dev->host_scratch;
  uint32_t* dflag = f->list_size + 1; /* second word of the 8-byte counter slot */
  SG_TRY(sg::zero_async(dev, dflag, 4));
  any_kernel<<<sg::grid_for(dev, f->w1, 256, 1), 256, 0, dev->stream>>>(f->l1, f->w1, dflag);
  SG_HIP(hipGetLastError());
  SG_HIP(hipMemcpyAsync(flag, dflag, 4, hipMemcpyDeviceToHost, dev->stream));
  SG_HIP(hipStreamSynchronize(dev->stream));
  *empty = (*flag == 0) ? 1 : 0;
  return SG_OK;
}

int sg_frontier_size(sg_frontier_t f, int64_t* size) {
  SG_CHECK(f && size, SG_ERR_INVALID, "sg_frontier_size: null argument");
  sg_device_t dev = f->dev;
  unsigned long long* acc = nullptr;
  SG_TRY(sg::alloc_device(dev, (void**)&acc, 8));
  int st = sg::zero_async(dev, acc, 8);
  if (st == SG_OK) {
    popcount_kernel<<<sg::grid_for(dev, f->w0, 256, 4), 256, 0, dev->stream>>>(f->l0, f->w0, acc);
    unsigned long long h = 0;
    hipError_t e = hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, dev->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) {
      sg::set_error("sg_frontier_size: %s", hipGetErrorString(e));
      st = SG_ERR_HIP;
    }
    *size = (int64_t)h;
  }
  (void)hipFree(acc);
  return st;
}

static int combine(sg_frontier_t out, sg_frontier_t a, sg_frontier_t b, int op) {
  SG_CHECK(out && a && b, SG_ERR_INVALID, "frontier combine: null frontier");
  SG_CHECK(out->n == a->n && a->n == b->n, SG_ERR_INVALID, "frontier combine: size mismatch (%lld, %lld, %lld)",
           (long long)out->n, (long long)a->n, (long long)b->n);
  sg_device_t dev = out->dev;
  if (out->w1 > 0)
    combine_kernel<<<sg::grid_for(dev, out->w1 * 64, 256, 4), 256, 0, dev->stream>>>(out->l0, out->l1, a->l0, b->l0, out->w0,
                                                                                      out->w1, op);
  SG_HIP(hipGetLastError());
  SG_HIP(hipStreamSynchronize(dev->stream));
  return SG_OK;
}

int sg_frontier_merge(sg_frontier_t f, sg_frontier_t other) { return combine(f, f, other, 0); }
int sg_frontier_intersect(sg_frontier_t f, sg_frontier_t other) { return combine(f, f, other, 1); }
int sg_frontier_and(sg_frontier_t out, sg_frontier_t a, sg_frontier_t b) { return combine(out, a, b, 1); }

int sg_frontier_compact_async(sg_frontier_t f, int mode) {
  SG_CHECK(f, SG_ERR_INVALID, "sg_frontier_compact: null frontier");
  SG_CHECK(mode >= 0 && mode <= 2, SG_ERR_INVALID, "sg_frontier_compact: mode must be 0, 1 or 2");
  return sg::frontier_compact_launch(f, mode);
}

int sg_frontier_compact(sg_frontier_t f, int mode, int64_t* count) {
  SG_TRY(sg_frontier_compact_async(f, mode));
  sg_device_t dev = f->dev;
  if (count) {
    uint32_t* h = (uint32_t*)dev->host_scratch;
    SG_HIP(hipMemcpyAsync(h, f->list_size, 4, hipMemcpyDeviceToHost, dev->stream));
    SG_HIP(hipStreamSynchronize(dev->stream));
    *count = *h;
  } else {
    SG_HIP(hipStreamSynchronize(dev->stream));
  }
  return SG_OK;
}

int sg_frontier_save(sg_frontier_t f, uint64_t* host_words) {
  SG_CHECK(f && host_words, SG_ERR_INVALID, "sg_frontier_save: null argument");
  SG_HIP(hipMemcpyAsync(host_words, f->l0, 8 * f->w0, hipMemcpyDeviceToHost, f->dev->stream));
  SG_HIP(hipMemcpyAsync(host_words + f->w0, f->l1, 8 * f->w1, hipMemcpyDeviceToHost, f->dev->stream));
  SG_HIP(hipStreamSynchronize(f->dev->stream));
  return SG_OK;
}

int sg_frontier_load(sg_frontier_t f, const uint64_t* host_words) {
  SG_CHECK(f && host_words, SG_ERR_INVALID, "sg_frontier_load: null argument");
  SG_HIP(hipMemcpyAsync(f->l0, host_words, 8 * f->w0, hipMemcpyHostToDevice, f->dev->stream));
  SG_HIP(hipMemcpyAsync(f->l1, host_words + f->w0, 8 * f->w1, hipMemcpyHostToDevice, f->dev->stream));
  SG_HIP(hipStreamSynchronize(f->dev->stream));
  return SG_OK;
}

int sg_frontier_swap(sg_frontier_t a, sg_frontier_t b) {
  SG_CHECK(a && b, SG_ERR_INVALID, "sg_frontier_swap: null frontier");
  SG_CHECK(a->n == b->n, SG_ERR_INVALID, "sg_frontier_swap: size mismatch");
  SG_CHECK(a->dev == b->dev, SG_ERR_INVALID, "sg_frontier_swap: frontiers on different devices");
  std::swap(a->l0, b->l0);
  std::swap(a->l1, b->l1);
  std::swap(a->list, b->list);
  std::swap(a->list_size, b->list_size);
  std::swap(a->tile_counts, b->tile_counts);
  return SG_OK;
}

} // extern "C"
