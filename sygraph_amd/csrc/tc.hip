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
constexpr int kTcUnroll = 4;  /* arcs (v, w) per lane in flight */

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

/* One wave per vertex u.  N+(u) goes into an open-addressing LDS hash table
 * (linear probing, empty slots hold -1) of at least twice its size (lists up
 * to kStage ids; longer ones are binary-searched in global memory, where they
 * are sorted).  The wave then looks up every w of the lists N+(v), v in
 * N+(u): w in N+(u) closes the triangle (u, v, w), counted once at its
 * lowest-ranked corner.  Lists of 64 or more arcs are walked one at a time by
 * the whole wave; the shorter ones, 64 at a time laid end to end (merge path).
 * Waves claim vertices kTcChunk at a time from a counter (next[0], zero at
 * launch): the work per vertex is skewed, and a fixed assignment leaves the
 * busiest wave with several times the mean. */
constexpr int kTcChunk = 1;

__global__ void __launch_bounds__(kBlock) intersect_kernel(const int64_t* dag_off, const int32_t* dag_cols, int64_t n,
                                                           unsigned long long* per_vertex, unsigned long long* total,
                                                           unsigned long long* next) {
  __shared__ int32_t table_all[kBlock / 64][2 * kStage];
  const int lane = lane_id();
  int32_t* table = table_all[threadIdx.x >> 6];
  uint64_t acc_total = 0;
  for (;;) {
    unsigned long long claim = 0;
    if (lane == 0) claim = atomicAdd(next, (unsigned long long)kTcChunk);
    const int64_t u0 = (int64_t)shfl_u64(claim, 0);
    if (u0 >= n) break;
    const int64_t u1 = u0 + kTcChunk < n ? u0 + kTcChunk : n;
    for (int64_t u = u0; u < u1; ++u) {
      const int64_t ub = dag_off[u], ue = dag_off[u + 1];
      const int32_t du = (int32_t)(ue - ub);
      if (du < 2) {
        if (per_vertex && lane == 0) per_vertex[u] = 0;
        continue;
      }
      const bool hashed = du <= kStage;
      /* at most 1/16 full where the table allows (2 * kStage slots), so most
       * lookups end at their first slot; at least 1/2 empty in every case */
      uint32_t size = 64;
      while (size < 16u * (uint32_t)du && size < 2u * kStage) size <<= 1;
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
      /* members of N+(u) among kTcUnroll ids per lane */
      auto count_hits = [&](const int32_t (&w)[kTcUnroll], const bool (&ok)[kTcUnroll]) {
        if (hashed) {
          /* probe rounds: the pending lookups of all kTcUnroll ids read their
           * next slot together (independent LDS reads), so the wave waits for
           * its longest probe sequence once instead of once per id */
          uint32_t h[kTcUnroll];
          uint32_t pend = 0;
#pragma unroll
          for (int k = 0; k < kTcUnroll; ++k) {
            h[k] = hash_slot(w[k], mask);
            pend |= ok[k] ? 1u << k : 0u;
          }
          while (pend) {
            int32_t y[kTcUnroll];
#pragma unroll
            for (int k = 0; k < kTcUnroll; ++k) y[k] = (pend >> k) & 1u ? table[h[k]] : -1;
#pragma unroll
            for (int k = 0; k < kTcUnroll; ++k) {
              if (!((pend >> k) & 1u)) continue;
              if (y[k] == w[k]) {
                ++cnt;
                pend &= ~(1u << k);
              } else if (y[k] < 0) {
                pend &= ~(1u << k);
              } else {
                h[k] = (h[k] + 1) & mask;
              }
            }
          }
        } else {
#pragma unroll
          for (int k = 0; k < kTcUnroll; ++k)
            if (ok[k]) cnt += search(list, du, w[k]) ? 1u : 0u;
        }
      };
      for (int32_t vbase = 0; vbase < du; vbase += 64) {
        int64_t vb = 0, vlen = 0;
        if (vbase + lane < du) {
          const int32_t v = list[vbase + lane];
          vb = dag_off[v];
          vlen = dag_off[v + 1] - vb;
        }
        int32_t w[kTcUnroll];
        bool ok[kTcUnroll];
        /* long lists: the whole wave walks each in turn, lanes side by side */
        const bool is_long = vlen >= 64;
        uint64_t longs = ballot(is_long);
        while (longs) {
          const int k2 = __ffsll((unsigned long long)longs) - 1;
          longs &= longs - 1;
          const int64_t lb = shfl_i64(vb, k2);
          const uint32_t ll = (uint32_t)__shfl((uint32_t)vlen, k2, 64);
          for (uint32_t j0 = 0; j0 < ll; j0 += kTcUnroll * 64) {
#pragma unroll
            for (int k = 0; k < kTcUnroll; ++k) {
              const uint32_t j = j0 + 64 * k + lane;
              ok[k] = j < ll;
              w[k] = ok[k] ? dag_cols[lb + j] : 0;
            }
            count_hits(w, ok);
          }
        }
        /* short lists: laid end to end, a lane per arc */
        const uint32_t slen = is_long ? 0u : (uint32_t)vlen;
        const uint32_t incl = wave_inclusive_scan(slen);
        const uint32_t tot = __shfl(incl, 63, 64);
        const uint32_t excl = incl - slen;
        for (uint32_t t0 = 0; t0 < tot; t0 += kTcUnroll * 64) {
#pragma unroll
          for (int k = 0; k < kTcUnroll; ++k) {
            const uint32_t t = t0 + 64 * k + lane;
            const int owner = scan_owner(incl, t);
            const uint32_t before = __shfl(excl, owner, 64);
            const int64_t st = shfl_i64(vb, owner);
            ok[k] = t < tot;
            w[k] = ok[k] ? dag_cols[st + (t - before)] : 0;
          }
          count_hits(w, ok);
        }
      }
      cnt = wave_sum_u64(cnt);
      if (lane == 0 && per_vertex) per_vertex[u] = cnt;
      acc_total += cnt;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
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
  SG_TRY(sg::graph_buffer(g, "tc_total", 16, &total)); /* the count, then the claim counter */
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
  SG_TRY(sg::zero_async(dev, total, 16));
  intersect_kernel<<<grid, kBlock, 0, s>>>(dag_off, dag_cols, n, (unsigned long long*)per_vertex, total, total + 1);
  SG_HIP(hipGetLastError());
  unsigned long long h = 0;
  SG_HIP(hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, s));
  SG_HIP(hipStreamSynchronize(s));
  *count = h;
  return SG_OK;
}
