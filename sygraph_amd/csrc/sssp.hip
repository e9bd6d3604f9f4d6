/*
 * sssp.hip -- single-source shortest paths for gfx950 (SSSP::run,
 * algorithms/sssp.hpp:152-210).
 *
 * Frontier Bellman-Ford, as the reference: every round relaxes all arcs of
 * the frontier with an atomic minimum on the tentative distance
 * (sssp.hpp:180-189, sync/atomics.hpp:76-81); vertices whose distance dropped
 * form the next frontier (the filter of sssp.hpp:192-196 only deduplicates
 * them, which the bitmap does by construction).  Each round reads the
 * frontier through the degree-class queues (thread / wave / workgroup bins),
 * relaxes with a plain-load pre-check before the atomic, and marks improved
 * vertices in the next-frontier bitmap with a test-then-OR.  Distances are
 * unique, so the result is independent of the relaxation order.
 *
 * Distances have the weights' 4-byte type: uint32 weights use integer min,
 * float weights (non-negative) compare through their bit patterns.  The
 * unreached value is n+1 in that type (sssp.hpp:43).
 */
#include <algorithm>
#include <cstring>

#include "traverse.hpp"

using namespace sg::trav;

namespace {

template<typename W>
struct DistOps;
template<>
struct DistOps<uint32_t> {
  __host__ __device__ static uint32_t bits(uint32_t x) { return x; }
  __host__ __device__ static uint32_t value(uint32_t b) { return b; }
};
template<>
struct DistOps<float> {
  __host__ __device__ static uint32_t bits(float x) { return __builtin_bit_cast(uint32_t, x); }
  __host__ __device__ static float value(uint32_t b) { return __builtin_bit_cast(float, b); }
};

template<typename OffT, typename W>
struct SsspArgs {
  const int32_t* cols;
  const W* w;
  uint32_t* dist; /* distance bits */
  uint64_t* next_bm;
  uint64_t* zero_bm; /* frontier bitmap of the previous round, cleared here */
  int64_t w0;
  QEntry* qin[3];
  const QueueHdr* qin_hdr;
};

template<typename OffT, typename W>
__global__ void __launch_bounds__(kBlock) relax_kernel(SsspArgs<OffT, W> a) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < a.w0; i += (int64_t)gridDim.x * kBlock)
    a.zero_bm[i] = 0ull;
  for_each_queued_arc<OffT>(a.qin, a.qin_hdr, a.cols, [&](uint32_t u, uint32_t v, int64_t e) {
    const W du = DistOps<W>::value(atomic_load(&a.dist[u]));
    const W nd = du + a.w[e];
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

template<typename OffT>
__global__ void reached_kernel(const OffT* off, const uint32_t* dist, uint32_t unreached, int64_t n,
                               unsigned long long* out) {
  uint64_t e = 0, c = 0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    if (dist[v] != unreached) {
      e += (uint64_t)(off[v + 1] - off[v]);
      ++c;
    }
  e = wave_sum_u64(e);
  c = wave_sum_u64(c);
  if (lane_id() == 0) {
    atomicAdd(&out[0], (unsigned long long)e);
    atomicAdd(&out[1], (unsigned long long)c);
  }
}

struct SsspWorkspace {
  uint64_t* bm[2] = {nullptr, nullptr};
  unsigned long long* acc = nullptr;
  QueueSet qs;
  int64_t w0 = 0;
};

void sssp_ws_free(SsspWorkspace* w) {
  if (!w) return;
  (void)hipFree(w->bm[0]);
  (void)hipFree(w->bm[1]);
  (void)hipFree(w->acc);
  w->qs.release();
  delete w;
}

template<typename OffT, typename W>
int run_sssp(sg_graph_t g, int32_t src, W* dist_out, sg_sssp_stats* stats) {
  sg_device_t dev = g->dev;
  hipStream_t s = dev->stream;
  const int64_t n = g->n;
  const Tuning tune = default_tuning();
  auto* ws = static_cast<SsspWorkspace*>(g->sssp_ws);
  if (!ws) {
    ws = new SsspWorkspace();
    ws->w0 = (n + 63) / 64;
    int st = SG_OK;
    if ((st = sg::alloc_device(dev, (void**)&ws->bm[0], 8 * (ws->w0 + 1))) ||
        (st = sg::alloc_device(dev, (void**)&ws->bm[1], 8 * (ws->w0 + 1))) ||
        (st = sg::alloc_device(dev, (void**)&ws->acc, 16)) || (st = ws->qs.init(dev, n, g->m, tune))) {
      sssp_ws_free(ws);
      return st;
    }
    g->sssp_ws = ws;
  }
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
  SG_TRY(sg::zero_async(dev, ws->bm[0], 8 * ws->w0));
  SG_TRY(sg::zero_async(dev, ws->bm[1], 8 * ws->w0));
  sssp_init_kernel<OffT><<<sg::grid_for(dev, n, 256, 4), 256, 0, s>>>(off, dist, unreached, n, src, Q.q[0][0],
                                                                       Q.q[0][1], Q.q[0][2], Q.hdr, tune);
  SG_HIP(hipGetLastError());
  QueueHdr* hh = (QueueHdr*)dev->host_scratch;
  int qs = 0;
  int32_t round = 0;
  int64_t fv = 1, fe = 0;
  {
    SG_HIP(hipMemcpyAsync(hh, Q.hdr, sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    fe = (int64_t)hh->deg;
  }
  int64_t relaxed = 0;
  for (;; ++round) {
    if (stats && round < SG_MAX_LEVELS_RECORDED) {
      stats->round_vertices[round] = fv;
      stats->round_edges[round] = fe;
    }
    relaxed += fe;
    SsspArgs<OffT, W> a{};
    a.cols = g->cols;
    a.w = (const W*)g->weights;
    a.dist = dist;
    a.next_bm = ws->bm[round & 1];
    a.zero_bm = ws->bm[(round + 1) & 1];
    a.w0 = ws->w0;
    for (int k = 0; k < 3; ++k) a.qin[k] = Q.q[qs][k];
    a.qin_hdr = Q.hdr + qs;
    relax_kernel<OffT, W><<<grid_main, kBlock, 0, s>>>(a);
    SG_HIP(hipGetLastError());
    SG_TRY((compact_to_queues<OffT, kFromBitmap>(s, Q, qs ^ 1, off, nullptr, nullptr, ws->bm[round & 1], tune)));
    SG_HIP(hipMemcpyAsync(hh, Q.hdr + (qs ^ 1), sizeof(QueueHdr), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    qs ^= 1;
    fv = (int64_t)hh->nvert;
    fe = (int64_t)hh->deg;
    if (fv == 0) {
      ++round;
      break;
    }
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
    stats->relaxations = relaxed;
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

extern "C" int sg_sssp(sg_graph_t g, int32_t source, int32_t load_balancer, void* dist, sg_sssp_stats* stats) {
  SG_CHECK(g && dist, SG_ERR_INVALID, "sg_sssp: null argument");
  SG_CHECK(source >= 0 && source < g->n, SG_ERR_INVALID, "sg_sssp: source %d out of range [0,%lld)", source,
           (long long)g->n);
  SG_CHECK(g->weights, SG_ERR_STATE, "sg_sssp: graph has no edge weights");
  SG_CHECK(load_balancer >= SG_LB_WORKITEM && load_balancer <= SG_LB_BUCKETING, SG_ERR_INVALID,
           "sg_sssp: unknown load balancer %d", load_balancer);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  SG_HIP(hipSetDevice(g->dev->ordinal));
  const bool fw = (g->flags & SG_GRAPH_WEIGHT_FLOAT) != 0;
  if (g->offset_bytes == 8)
    return fw ? run_sssp<int64_t, float>(g, source, (float*)dist, stats)
              : run_sssp<int64_t, uint32_t>(g, source, (uint32_t*)dist, stats);
  return fw ? run_sssp<uint32_t, float>(g, source, (float*)dist, stats)
            : run_sssp<uint32_t, uint32_t>(g, source, (uint32_t*)dist, stats);
}
