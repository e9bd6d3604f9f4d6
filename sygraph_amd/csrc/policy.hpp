/*
 * policy.hpp -- layout, features and one-wavefront inference of the learned
 * BFS policy (policy.hip trains it; bfs.hip runs it at every level decision).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace sg::policy {

constexpr int kIn = 8, kH = 32, kOut = 6;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kOut * kH;
constexpr int kParams = kOffB3 + kOut; /* 1,542 */
static_assert(kParams == 1542, "8-32-32-6 MLP");

/* Feature k (0..7) of the frontier entering a level, scaled to about [0, 1]:
 * log2 of frontier vertices, frontier arcs, unexplored arcs and the level,
 * mean and largest frontier degree, average degree and vertex count. */
__host__ __device__ inline float feature(int k, double frontier_vertices, double frontier_edges, double unexplored,
                                         int level, double max_degree, double n, double m) {
  switch (k) {
  case 0: return (float)(log2(1.0 + frontier_vertices) / 24.0);
  case 1: return (float)(log2(1.0 + frontier_edges) / 32.0);
  case 2: return (float)(log2(1.0 + unexplored) / 32.0);
  case 3: return (float)(log2(1.0 + (double)level) / 12.0);
  case 4: return (float)(log2(1.0 + frontier_edges / (frontier_vertices > 1.0 ? frontier_vertices : 1.0)) / 16.0);
  case 5: return (float)(log2(1.0 + max_degree) / 24.0);
  case 6: return (float)(log2(1.0 + m / (n > 1.0 ? n : 1.0)) / 8.0);
  default: return (float)(log2(n > 1.0 ? n : 1.0) / 32.0);
  }
}

__host__ __device__ inline void features(float* f, double frontier_vertices, double frontier_edges, double unexplored,
                                         int level, double max_degree, double n, double m) {
  for (int k = 0; k < kIn; ++k)
    f[k] = feature(k, frontier_vertices, frontier_edges, unexplored, level, max_degree, n, m);
}

/* The parameters as inference reads them from LDS.  Weight matrices are
 * transposed ([input][unit]) so that the lanes of one layer read consecutive
 * words; W3's rows are padded to 8 units. */
struct Staged {
  float w1t[kIn * kH];
  float b1[kH];
  float w2t[kH * kH];
  float b2[kH];
  float w3t[kH * 8];
  float b3[8];
};

/* Copy the parameters into LDS with threads tid = 0 .. nthreads-1 of the
 * workgroup; the caller synchronizes the workgroup before inference. */
__device__ inline void stage(Staged* s, const float* __restrict__ P, int tid, int nthreads) {
  for (int i = tid; i < kParams; i += nthreads) {
    const float v = P[i];
    if (i < kOffB1) s->w1t[(i % kIn) * kH + i / kIn] = v;
    else if (i < kOffW2) s->b1[i - kOffB1] = v;
    else if (i < kOffB2) s->w2t[((i - kOffW2) % kH) * kH + (i - kOffW2) / kH] = v;
    else if (i < kOffW3) s->b2[i - kOffB2] = v;
    else if (i < kOffB3) s->w3t[((i - kOffW3) % kH) * 8 + (i - kOffW3) / kH] = v;
    else s->b3[i - kOffB3] = v;
  }
}

/* Action = direction * 3 + load-balancer variant from staged parameters; all
 * 64 lanes of one wavefront call with the same level statistics.  Lane k < 8
 * evaluates feature k, lane j < 32 hidden unit j, lane o < 6 output o; every
 * unit accumulates its bias first and then its inputs in order.  Ties go to
 * the lowest action. */
__device__ inline int infer_staged(const Staged* __restrict__ s, double frontier_vertices, double frontier_edges,
                                   double unexplored, int level, double max_degree, double n, double m) {
  const int lane = (int)__lane_id();
  const float mine =
      feature(lane & 7, frontier_vertices, frontier_edges, unexplored, level, max_degree, n, m);
  float f[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) f[k] = __shfl(mine, k, 64);
  const int j = lane & (kH - 1);
  float a = s->b1[j];
#pragma unroll
  for (int k = 0; k < kIn; ++k) a += s->w1t[k * kH + j] * f[k];
  const float h1 = a > 0.f ? a : 0.f;
  float a2 = s->b2[j];
#pragma unroll
  for (int k = 0; k < kH; ++k) a2 += s->w2t[k * kH + j] * __shfl(h1, k, 64);
  const float h2 = a2 > 0.f ? a2 : 0.f;
  const int o = lane & 7;
  float sc = s->b3[o];
#pragma unroll
  for (int k = 0; k < kH; ++k) sc += s->w3t[k * 8 + o] * __shfl(h2, k, 64);
  int best = 0;
  float bs = __shfl(sc, 0, 64);
#pragma unroll
  for (int i = 1; i < kOut; ++i) {
    const float si = __shfl(sc, i, 64);
    if (si > bs) {
      bs = si;
      best = i;
    }
  }
  return best;
}

} // namespace sg::policy
