/*
 * tc.hip -- triangle counting for gfx950 (TC::run + getNumTriangles,
 * algorithms/tc.hpp:67-138).
 *
 * The reference merges N(u) and N(v) for every arc u < v and divides the sum
 * by three; its running count is added inside the merge loop (tc.hpp:104), so
 * it over-counts.  This path computes the exact count on the
 * (degree, id)-oriented DAG: u -> v iff (deg u, u) < (deg v, v), each
 * triangle is then found once, from its lowest-ranked corner, and
 * out-degrees are O(sqrt(m)).
 *   1. orient: one wave per vertex keeps, in ascending id order, the
 *      neighbours that rank higher (ballot + prefix to preserve order);
 *   2. intersect: one wave per vertex u puts N+(u) in an LDS hash table,
 *      then walks the lists N+(v), v in N+(u), 64 lists laid end to end at a
 *      time, a lane per arc (v, w) probing w in the table (segmented
 *      intersection).  Lists longer than the LDS stage are binary-searched in
 *      global memory, where they are sorted.
 * The count is uint64 (the reference's uint32 would overflow at scale 20).
 */
#include <rocprim/device/device_scan.hpp>

#include <cstring>

#include "traverse.hpp"

using namespace sg::trav;

namespace {

constexpr int kStage = 1024; /* longest N+(u) hashed in LDS (table of 2 * kStage slots per wave) */
constexpr int kTcUnroll = 8;  /* arcs (v, w) per lane in flight */

template<typename OffT>
__device__ __forceinline__ bool ranks_higher(const OffT* off, uint32_t u, int64_t du, uint32_t v) {
  const int64_t dv = (int64_t)(off[v + 1] - off[v]);
  return du < dv || (du == dv && u < v);
}

/* pass 1 (count) and pass 2 (fill) of the orientation; one wave per vertex */
template<typename OffT, bool kFill>
__global__ void __launch_bounds__(kBlock) orient_kernel(const OffT* off, const int32_t* cols, int64_t n,
                                                        uint32_t* outdeg, const int64_t* dag_off, int32_t* dag_cols) {
  const int lane = lane_id();
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  for (int64_t u = gwave; u < n; u += nwaves) {
    const int64_t b = (int64_t)off[u], e = (int64_t)off[u + 1], du = e - b;
    uint32_t cnt = 0;
    int64_t pos = kFill ? dag_off[u] : 0;
    for (int64_t j0 = b; j0 < e; j0 += 64) {
      const int64_t j = j0 + lane;
      bool keep = false;
      uint32_t v = 0;
      if (j < e) {
        v = (uint32_t)cols[j];
        keep = ranks_higher(off, (uint32_t)u, du, v);
      }
      const uint64_t m = ballot(keep);
      if (kFill && keep) dag_cols[pos + lanes_below(m)] = (int32_t)v;
      pos += __popcll(m);
      cnt += (uint32_t)__popcll(m);
    }
    if (!kFill && lane == 0) outdeg[u] = cnt;
  }
}

__device__ __forceinline__ bool search(const int32_t* list, int32_t len, int32_t x) {
  int32_t lo = 0, hi = len;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    const int32_t y = list[mid];
    if (y == x) return true;
    if (y < x) lo = mid + 1;
    else hi = mid;
  }
  return false;
}

__device__ __forceinline__ uint32_t hash_slot(int32_t x, uint32_t mask) {
  return ((uint32_t)x * 2654435761u >> 7) & mask; /* Knuth's multiplicative hash */
}

/* Open-addressing membership in the wave's LDS table (empty slots hold -1). */
__device__ __forceinline__ bool hash_contains(const int32_t* table, uint32_t mask, int32_t x) {
  for (uint32_t h = hash_slot(x, mask);; h = (h + 1) & mask) {
    const int32_t y = table[h];
    if (y == x) return true;
    if (y < 0) return false;
  }
}

/* One wave per vertex u.  N+(u) goes into an LDS hash table of at least
 * twice its size (lists up to kStage ids; longer ones are binary-searched in
 * global memory, where they are sorted).  The wave then walks the lists
 * N+(v), v in N+(u), 64 of them laid end to end at a time (merge path), a lane
 * per arc (v, w): w in N+(u) closes the triangle (u, v, w), counted once at
 * its lowest-ranked corner. */
__global__ void __launch_bounds__(kBlock) intersect_kernel(const int64_t* dag_off, const int32_t* dag_cols, int64_t n,
                                                           unsigned long long* per_vertex, unsigned long long* total) {
  __shared__ int32_t table_all[kBlock / 64][2 * kStage];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  int32_t* table = table_all[wv];
  const int64_t gwave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  uint64_t acc_total = 0;
  for (int64_t u = gwave; u < n; u += nwaves) {
    const int64_t ub = dag_off[u], ue = dag_off[u + 1];
    const int32_t du = (int32_t)(ue - ub);
    if (du < 2) {
      if (per_vertex && lane == 0) per_vertex[u] = 0;
      continue;
    }
    const bool hashed = du <= kStage;
    uint32_t size = 64;
    while (size < 2u * (uint32_t)du) size <<= 1;
    const uint32_t mask = size - 1;
    if (hashed) {
      for (uint32_t i = lane; i < size; i += 64) table[i] = -1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      for (int32_t i = lane; i < du; i += 64) {
        const int32_t x = dag_cols[ub + i];
        for (uint32_t h = hash_slot(x, mask);; h = (h + 1) & mask)
          if (atomicCAS(&table[h], -1, x) == -1) break;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    const int32_t* list = dag_cols + ub;
    uint64_t cnt = 0;
    for (int32_t vbase = 0; vbase < du; vbase += 64) {
      int64_t vb = 0, vlen = 0;
      if (vbase + lane < du) {
        const int32_t v = list[vbase + lane];
        vb = dag_off[v];
        vlen = dag_off[v + 1] - vb;
      }
      const uint32_t incl = wave_inclusive_scan((uint32_t)vlen);
      const uint32_t tot = __shfl(incl, 63, 64);
      const uint32_t excl = incl - (uint32_t)vlen;
      for (uint32_t t0 = 0; t0 < tot; t0 += kTcUnroll * 64) {
        int32_t w[kTcUnroll];
        bool ok[kTcUnroll];
#pragma unroll
        for (int k = 0; k < kTcUnroll; ++k) {
          const uint32_t t = t0 + 64 * k + lane;
          const int owner = scan_owner(incl, t);
          const uint32_t before = __shfl(excl, owner, 64);
          const int64_t st = shfl_i64(vb, owner);
          ok[k] = t < tot;
          w[k] = ok[k] ? dag_cols[st + (t - before)] : 0;
        }
#pragma unroll
        for (int k = 0; k < kTcUnroll; ++k)
          if (ok[k]) cnt += (hashed ? hash_contains(table, mask, w[k]) : search(list, du, w[k])) ? 1u : 0u;
      }
    }
    cnt = wave_sum_u64(cnt);
    if (lane == 0 && per_vertex) per_vertex[u] = cnt;
    acc_total += cnt;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  if (lane == 0 && acc_total) atomicAdd(total, (unsigned long long)acc_total);
}

} // namespace

extern "C" int sg_tc(sg_graph_t g, uint64_t* count, uint64_t* per_vertex) {
  SG_CHECK(g && count, SG_ERR_INVALID, "sg_tc: null argument");
  SG_HIP(hipSetDevice(g->dev->ordinal));
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n;
  /* the oriented graph and the scan scratch are kept with the graph, so
   * repeated counts allocate nothing once the buffers have grown */
  uint32_t* outdeg = nullptr;
  int64_t* dag_off = nullptr;
  int32_t* dag_cols = nullptr;
  unsigned long long* total = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  SG_TRY(sg::graph_buffer(g, "tc_outdeg", 4 * (n + 1), &outdeg));
  SG_TRY(sg::graph_buffer(g, "tc_dag_off", 8 * (n + 1), &dag_off));
  SG_TRY(sg::graph_buffer(g, "tc_total", 8, &total));
  const int grid = dev->cus * 8;
  if (g->offset_bytes == 8)
    orient_kernel<int64_t, false><<<grid, kBlock, 0, s>>>((const int64_t*)g->offsets, g->cols, n, outdeg, nullptr, nullptr);
  else
    orient_kernel<uint32_t, false><<<grid, kBlock, 0, s>>>((const uint32_t*)g->offsets, g->cols, n, outdeg, nullptr,
                                                           nullptr);
  SG_HIP(hipGetLastError());
  /* dag_off = exclusive scan of outdeg over n+1 entries (outdeg[n] = 0) */
  SG_TRY(sg::zero_async(dev, outdeg + n, 4));
  if (rocprim::exclusive_scan(nullptr, tmp_bytes, outdeg, dag_off, (int64_t)0, (size_t)(n + 1), rocprim::plus<int64_t>(),
                              s) != hipSuccess)
    return sg::set_error("sg_tc: scan sizing failed"), SG_ERR_HIP;
  SG_TRY(sg::graph_buffer(g, "tc_scan_tmp", tmp_bytes, &tmp));
  if (rocprim::exclusive_scan(tmp, tmp_bytes, outdeg, dag_off, (int64_t)0, (size_t)(n + 1), rocprim::plus<int64_t>(),
                              s) != hipSuccess)
    return sg::set_error("sg_tc: scan failed"), SG_ERR_HIP;
  int64_t mdag = 0;
  SG_HIP(hipMemcpyAsync(&mdag, dag_off + n, 8, hipMemcpyDeviceToHost, s));
  SG_HIP(hipStreamSynchronize(s));
  SG_TRY(sg::graph_buffer(g, "tc_dag_cols", 4 * (mdag + 1), &dag_cols));
  if (g->offset_bytes == 8)
    orient_kernel<int64_t, true><<<grid, kBlock, 0, s>>>((const int64_t*)g->offsets, g->cols, n, nullptr, dag_off, dag_cols);
  else
    orient_kernel<uint32_t, true><<<grid, kBlock, 0, s>>>((const uint32_t*)g->offsets, g->cols, n, nullptr, dag_off,
                                                          dag_cols);
  SG_TRY(sg::zero_async(dev, total, 8));
  intersect_kernel<<<grid, kBlock, 0, s>>>(dag_off, dag_cols, n, (unsigned long long*)per_vertex, total);
  SG_HIP(hipGetLastError());
  unsigned long long h = 0;
  SG_HIP(hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, s));
  SG_HIP(hipStreamSynchronize(s));
  *count = h;
  return SG_OK;
}
