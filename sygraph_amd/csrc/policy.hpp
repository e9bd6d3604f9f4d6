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

/* Eight features of the frontier entering a level, each scaled to about
 * [0, 1]: log2 of frontier vertices, frontier arcs, unexplored arcs, mean and
 * largest frontier degree, average degree and vertex count, and the level. */
__host__ __device__ inline void features(float* f, double frontier_vertices, double frontier_edges, double unexplored,
                                         int level, double max_degree, double n, double m) {
  f[0] = (float)(log2(1.0 + frontier_vertices) / 24.0);
  f[1] = (float)(log2(1.0 + frontier_edges) / 32.0);
  f[2] = (float)(log2(1.0 + unexplored) / 32.0);
  f[3] = (float)level / 32.0f;
  f[4] = (float)(log2(1.0 + frontier_edges / (frontier_vertices > 1.0 ? frontier_vertices : 1.0)) / 16.0);
  f[5] = (float)(log2(1.0 + max_degree) / 24.0);
  f[6] = (float)(log2(1.0 + m / (n > 1.0 ? n : 1.0)) / 8.0);
  f[7] = (float)(log2(n > 1.0 ? n : 1.0) / 32.0);
}

/* Action = direction * 3 + load-balancer variant; every lane of one
 * wavefront must call (lane k < 32 owns hidden unit k, lane k < 6 output k).
 * Ties go to the lowest action. */
__device__ inline int infer_wave(const float* __restrict__ P, const float* f) {
  const int lane = (int)__lane_id();
  const float* W1 = P + kOffW1;
  const float* B1 = P + kOffB1;
  const float* W2 = P + kOffW2;
  const float* B2 = P + kOffB2;
  const float* W3 = P + kOffW3;
  const float* B3 = P + kOffB3;
  float h1 = 0.f;
  if (lane < kH) {
    float a = B1[lane];
    for (int k = 0; k < kIn; ++k) a += W1[lane * kIn + k] * f[k];
    h1 = a > 0.f ? a : 0.f;
  }
  float a2 = lane < kH ? B2[lane] : 0.f;
  for (int k = 0; k < kH; ++k) {
    const float hk = __shfl(h1, k, 64);
    if (lane < kH) a2 += W2[lane * kH + k] * hk;
  }
  const float h2 = a2 > 0.f ? a2 : 0.f;
  float s = lane < kOut ? B3[lane] : 0.f;
  for (int k = 0; k < kH; ++k) {
    const float hk = __shfl(h2, k, 64);
    if (lane < kOut) s += W3[lane * kH + k] * hk;
  }
  int best = 0;
  float bs = __shfl(s, 0, 64);
  for (int i = 1; i < kOut; ++i) {
    const float si = __shfl(s, i, 64);
    if (si > bs) {
      bs = si;
      best = i;
    }
  }
  return best;
}

} // namespace sg::policy
