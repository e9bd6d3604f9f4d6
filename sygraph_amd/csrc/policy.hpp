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
 * mean and largest frontier degree, average degree and vertex count.  FP32
 * (one hardware log2 per feature: a decision sits on the level's critical
 * path); the training features are the same formulas in FP64. */
__host__ __device__ inline float feature(int k, double frontier_vertices, double frontier_edges, double unexplored,
                                         int level, double max_degree, double n, double m) {
  const float fv = (float)frontier_vertices, fe = (float)frontier_edges;
  switch (k) {
  case 0: return log2f(1.0f + fv) / 24.f;
  case 1: return log2f(1.0f + fe) / 32.f;
  case 2: return log2f(1.0f + (float)unexplored) / 32.f;
  case 3: return log2f(1.0f + (float)level) / 12.f;
  case 4: return log2f(1.0f + fe / (fv > 1.f ? fv : 1.f)) / 16.f;
  case 5: return log2f(1.0f + (float)max_degree) / 24.f;
  case 6: return log2f(1.0f + (float)(m / (n > 1.0 ? n : 1.0))) / 8.f;
  default: return log2f((float)(n > 1.0 ? n : 1.0)) / 32.f;
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

/* Value of lane k (a constant) in every lane: v_readlane, a few cycles,
 * against tens for an LDS permute. */
__device__ __forceinline__ float lane_value(float x, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}

/* Feature of lane k & 7 without divergence: one log2 per lane.  f6 and f7
 * (average degree and vertex count) are constant for a graph and come from
 * feature() on the host. */
__device__ __forceinline__ float feature_of_lane(int k, float fv, float fe, float un, float lvl, float md, float f6,
                                                 float f7) {
  const float x = k == 0 ? fv : (k == 1 ? fe : (k == 2 ? un : (k == 3 ? lvl : (k == 4 ? fe / (fv > 1.f ? fv : 1.f) : md))));
  const float sc = (k == 0 || k == 5) ? 24.f : (k == 3 ? 12.f : (k == 4 ? 16.f : 32.f));
  const float r = log2f(1.0f + x) / sc;
  return k == 6 ? f6 : (k == 7 ? f7 : r);
}

/* Action = direction * 3 + load-balancer variant from staged parameters; all
 * 64 lanes of one wavefront call with the same level statistics.  Lane k < 8
 * evaluates feature k, lane j < 32 hidden unit j, lane o < 6 output o; every
 * unit accumulates its bias first and then its inputs in order.  Ties go to
 * the lowest action. */
__device__ inline int infer_staged(const Staged* __restrict__ s, double frontier_vertices, double frontier_edges,
                                   double unexplored, int level, double max_degree, float f6, float f7) {
  const int lane = (int)__lane_id();
  const float mine = feature_of_lane(lane & 7, (float)frontier_vertices, (float)frontier_edges, (float)unexplored,
                                     (float)level, (float)max_degree, f6, f7);
  float f[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) f[k] = lane_value(mine, k);
  const int j = lane & (kH - 1);
  float a = s->b1[j];
#pragma unroll
  for (int k = 0; k < kIn; ++k) a += s->w1t[k * kH + j] * f[k];
  const float h1 = a > 0.f ? a : 0.f;
  float a2 = s->b2[j];
#pragma unroll
  for (int k = 0; k < kH; ++k) a2 += s->w2t[k * kH + j] * lane_value(h1, k);
  const float h2 = a2 > 0.f ? a2 : 0.f;
  const int o = lane & 7;
  float sc = s->b3[o];
#pragma unroll
  for (int k = 0; k < kH; ++k) sc += s->w3t[k * 8 + o] * lane_value(h2, k);
  int best = 0;
  float bs = lane_value(sc, 0);
#pragma unroll
  for (int i = 1; i < kOut; ++i) {
    const float si = lane_value(sc, i);
    if (si > bs) {
      bs = si;
      best = i;
    }
  }
  return best;
}

} // namespace sg::policy
