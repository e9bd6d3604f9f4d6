/*
 * policy.hip -- the learned per-level BFS policy: an 8 -> 32 -> 32 -> 6 MLP
 * (ReLU, FP32, 1,542 parameters) that scores the six actions
 * {push, pull} x {three work mappings} from eight features of the frontier
 * entering a level.  No counterpart in the reference, whose
 * only policy is the alpha/beta rule of algorithms/bfs.hpp:325-350; that rule
 * remains the default and the fallback.
 *
 * Training runs on the device: a forward kernel (one thread per sample) keeps
 * the activations and the per-sample backward signals; a gradient kernel sums
 * each parameter over chunks of 2,048 samples in sample order and a second
 * pass adds the chunks in order (no atomics, so gradients are deterministic);
 * then an Adam kernel.  Inference inside the BFS level decisions (bfs.hip) is
 * sg::policy::infer_staged() on one wavefront (policy.hpp).
 *
 * Parameter layout (floats): W1[32][8] b1[32] W2[32][32] b2[32] W3[6][32] b3[6].
 */
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "policy.hpp"

using namespace sygraph::hipdev;

namespace {

constexpr int kB = 256;

/* Forward + per-sample backward signals of the mean softmax cross-entropy. */
__global__ void __launch_bounds__(kB) mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                         const int32_t* __restrict__ Y, int32_t batch,
                                                         float* __restrict__ H1, float* __restrict__ H2,
                                                         float* __restrict__ S, float* __restrict__ D1,
                                                         float* __restrict__ D2, float* __restrict__ DS,
                                                         float* __restrict__ loss) {
  using namespace sg::policy;
  const int b = blockIdx.x * kB + threadIdx.x;
  if (b >= batch) return;
  const float* W1 = P + kOffW1;
  const float* B1 = P + kOffB1;
  const float* W2 = P + kOffW2;
  const float* B2 = P + kOffB2;
  const float* W3 = P + kOffW3;
  const float* B3 = P + kOffB3;
  float x[kIn], h1[kH], h2[kH], s[kOut];
  for (int k = 0; k < kIn; ++k) x[k] = X[(int64_t)b * kIn + k];
  for (int i = 0; i < kH; ++i) {
    float a = B1[i];
    for (int k = 0; k < kIn; ++k) a += W1[i * kIn + k] * x[k];
    h1[i] = a > 0.f ? a : 0.f;
  }
  for (int i = 0; i < kH; ++i) {
    float a = B2[i];
    for (int k = 0; k < kH; ++k) a += W2[i * kH + k] * h1[k];
    h2[i] = a > 0.f ? a : 0.f;
  }
  float mx = -INFINITY;
  for (int i = 0; i < kOut; ++i) {
    float a = B3[i];
    for (int k = 0; k < kH; ++k) a += W3[i * kH + k] * h2[k];
    s[i] = a;
    mx = a > mx ? a : mx;
  }
  float z = 0.f;
  for (int i = 0; i < kOut; ++i) z += expf(s[i] - mx);
  const int y = Y ? Y[b] : 0;
  const float inv_b = 1.0f / (float)batch;
  float ds[kOut];
  for (int i = 0; i < kOut; ++i) {
    const float p = expf(s[i] - mx) / z;
    ds[i] = (p - (i == y ? 1.f : 0.f)) * inv_b;
  }
  if (loss) loss[b] = (logf(z) + mx - s[y]);
  float d2[kH], d1[kH];
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
    for (int i = 0; i < kOut; ++i) a += W3[i * kH + k] * ds[i];
    d2[k] = h2[k] > 0.f ? a : 0.f;
  }
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
    for (int i = 0; i < kH; ++i) a += W2[i * kH + k] * d2[i];
    d1[k] = h1[k] > 0.f ? a : 0.f;
  }
  for (int i = 0; i < kH; ++i) {
    H1[(int64_t)b * kH + i] = h1[i];
    H2[(int64_t)b * kH + i] = h2[i];
    D1[(int64_t)b * kH + i] = d1[i];
    D2[(int64_t)b * kH + i] = d2[i];
  }
  for (int i = 0; i < kOut; ++i) {
    S[(int64_t)b * kOut + i] = s[i];
    DS[(int64_t)b * kOut + i] = ds[i];
  }
}

/* dL/dp as fixed-order partial sums: thread (p, c) sums parameter p over
 * samples [c * kChunk, (c + 1) * kChunk) in sample order; mlp_reduce_kernel
 * adds the chunks in order, so the gradient is deterministic. */
constexpr int kChunk = 2048;

__global__ void __launch_bounds__(kB) mlp_grad_kernel(const float* __restrict__ X, int32_t batch,
                                                      const float* __restrict__ H1, const float* __restrict__ H2,
                                                      const float* __restrict__ D1, const float* __restrict__ D2,
                                                      const float* __restrict__ DS, float* __restrict__ part) {
  using namespace sg::policy;
  const int p = blockIdx.x * kB + threadIdx.x;
  const int c = blockIdx.y;
  if (p >= kParams) return;
  const float* a = nullptr; /* upstream signal (per sample, row of width wa) */
  const float* in = nullptr;  /* input activation (per sample, row of width wi), nullptr for a bias */
  int ia = 0, ii = 0, wa = 0, wi = 0;
  if (p < kOffB1) {
    ia = (p - kOffW1) / kIn, ii = (p - kOffW1) % kIn, a = D1, wa = kH, in = X, wi = kIn;
  } else if (p < kOffW2) {
    ia = p - kOffB1, a = D1, wa = kH;
  } else if (p < kOffB2) {
    ia = (p - kOffW2) / kH, ii = (p - kOffW2) % kH, a = D2, wa = kH, in = H1, wi = kH;
  } else if (p < kOffW3) {
    ia = p - kOffB2, a = D2, wa = kH;
  } else if (p < kOffB3) {
    ia = (p - kOffW3) / kH, ii = (p - kOffW3) % kH, a = DS, wa = kOut, in = H2, wi = kH;
  } else {
    ia = p - kOffB3, a = DS, wa = kOut;
  }
  const int b0 = c * kChunk, b1 = b0 + kChunk < batch ? b0 + kChunk : batch;
  float acc = 0.f;
  if (in)
    for (int b = b0; b < b1; ++b) acc += a[(int64_t)b * wa + ia] * in[(int64_t)b * wi + ii];
  else
    for (int b = b0; b < b1; ++b) acc += a[(int64_t)b * wa + ia];
  part[(int64_t)c * kParams + p] = acc;
}

__global__ void mlp_reduce_kernel(const float* __restrict__ part, int32_t chunks, float* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= sg::policy::kParams) return;
  float acc = 0.f;
  for (int c = 0; c < chunks; ++c) acc += part[(int64_t)c * sg::policy::kParams + p];
  G[p] = acc;
}

__global__ void mlp_adam_kernel(float* P, const float* G, float* M, float* V, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= sg::policy::kParams) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void mlp_sum_kernel(const float* x, int32_t n, float* out) {
  __shared__ float red[kB / 64];
  float acc = 0.f;
  for (int i = threadIdx.x; i < n; i += kB) acc += x[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane_id() == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kB / 64; ++w) t += red[w];
    *out = t / (float)n;
  }
}

} // namespace

struct sg_policy_s {
  sg::DevRef dev;
  float* params = nullptr; /* device */
  float *m = nullptr, *v = nullptr, *grad = nullptr;
  int64_t step = 0;
  /* training scratch, grown on demand */
  int32_t cap = 0;
  float *h1 = nullptr, *h2 = nullptr, *s = nullptr, *d1 = nullptr, *d2 = nullptr, *ds = nullptr, *loss = nullptr,
        *loss_mean = nullptr, *part = nullptr;
  void free_scratch() {
    for (float* p : {h1, h2, s, d1, d2, ds, loss, loss_mean, part}) (void)hipFree(p);
    h1 = h2 = s = d1 = d2 = ds = loss = loss_mean = part = nullptr;
    cap = 0;
  }
};

namespace {

int ensure_scratch(sg_policy_t p, int32_t batch) {
  using namespace sg::policy;
  if (batch <= p->cap) return SG_OK;
  p->free_scratch();
  const int64_t B = batch;
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->h1, 4 * B * kH));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->h2, 4 * B * kH));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->d1, 4 * B * kH));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->d2, 4 * B * kH));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->s, 4 * B * kOut));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->ds, 4 * B * kOut));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->loss, 4 * B));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->loss_mean, 16));
  SG_TRY(sg::alloc_device(p->dev, (void**)&p->part, 4 * (size_t)kParams * ((B + kChunk - 1) / kChunk)));
  p->cap = batch;
  return SG_OK;
}

int run_forward(sg_policy_t p, const float* X, const int32_t* Y, int32_t batch) {
  SG_TRY(ensure_scratch(p, batch));
  hipStream_t s = p->dev->stream;
  mlp_forward_kernel<<<(batch + kB - 1) / kB, kB, 0, s>>>(p->params, X, Y, batch, p->h1, p->h2, p->s, p->d1, p->d2,
                                                         p->ds, p->loss);
  SG_HIP(hipGetLastError());
  return SG_OK;
}

int run_grad(sg_policy_t p, const float* X, int32_t batch) {
  hipStream_t s = p->dev->stream;
  const int chunks = (batch + kChunk - 1) / kChunk;
  mlp_grad_kernel<<<dim3((sg::policy::kParams + kB - 1) / kB, chunks), kB, 0, s>>>(X, batch, p->h1, p->h2, p->d1,
                                                                                   p->d2, p->ds, p->part);
  mlp_reduce_kernel<<<(sg::policy::kParams + kB - 1) / kB, kB, 0, s>>>(p->part, chunks, p->grad);
  SG_HIP(hipGetLastError());
  return SG_OK;
}

} // namespace

extern "C" {

int sg_policy_create(sg_device_t dev, sg_policy_t* out) {
  SG_CHECK(dev && out, SG_ERR_INVALID, "sg_policy_create: null argument");
  SG_HIP(hipSetDevice(dev->ordinal));
  auto* p = new sg_policy_s();
  p->dev = dev;
  const size_t bytes = 4 * sg::policy::kParams;
  int st = SG_OK;
  if ((st = sg::alloc_device(dev, (void**)&p->params, bytes)) || (st = sg::alloc_device(dev, (void**)&p->m, bytes)) ||
      (st = sg::alloc_device(dev, (void**)&p->v, bytes)) || (st = sg::alloc_device(dev, (void**)&p->grad, bytes))) {
    sg_policy_destroy(p);
    return st;
  }
  if ((st = sg::zero_async(dev, p->params, bytes)) || (st = sg::zero_async(dev, p->m, bytes)) ||
      (st = sg::zero_async(dev, p->v, bytes))) {
    sg_policy_destroy(p);
    return st;
  }
  *out = p;
  return SG_OK;
}

int sg_policy_destroy(sg_policy_t p) {
  if (!p) return SG_OK;
  p->free_scratch();
  for (float* q : {p->params, p->m, p->v, p->grad}) (void)hipFree(q);
  delete p;
  return SG_OK;
}

int sg_policy_set_params(sg_policy_t p, const float* host_params) {
  SG_CHECK(p && host_params, SG_ERR_INVALID, "sg_policy_set_params: null argument");
  SG_HIP(hipMemcpyAsync(p->params, host_params, 4 * sg::policy::kParams, hipMemcpyHostToDevice, p->dev->stream));
  SG_TRY(sg::zero_async(p->dev, p->m, 4 * sg::policy::kParams));
  SG_TRY(sg::zero_async(p->dev, p->v, 4 * sg::policy::kParams));
  p->step = 0;
  SG_HIP(hipStreamSynchronize(p->dev->stream));
  return SG_OK;
}

int sg_policy_get_params(sg_policy_t p, float* host_params) {
  SG_CHECK(p && host_params, SG_ERR_INVALID, "sg_policy_get_params: null argument");
  SG_HIP(hipMemcpyAsync(host_params, p->params, 4 * sg::policy::kParams, hipMemcpyDeviceToHost, p->dev->stream));
  SG_HIP(hipStreamSynchronize(p->dev->stream));
  return SG_OK;
}

int sg_policy_forward(sg_policy_t p, const float* features, int32_t batch, float* scores) {
  SG_CHECK(p && features && scores && batch > 0, SG_ERR_INVALID, "sg_policy_forward: bad argument");
  SG_TRY(run_forward(p, features, nullptr, batch));
  SG_HIP(hipMemcpyAsync(scores, p->s, 4 * (size_t)batch * sg::policy::kOut, hipMemcpyDeviceToDevice, p->dev->stream));
  SG_HIP(hipStreamSynchronize(p->dev->stream));
  return SG_OK;
}

int sg_policy_gradient(sg_policy_t p, const float* features, const int32_t* labels, int32_t batch, float* grad,
                       float* loss) {
  SG_CHECK(p && features && labels && grad && batch > 0, SG_ERR_INVALID, "sg_policy_gradient: bad argument");
  SG_TRY(run_forward(p, features, labels, batch));
  SG_TRY(run_grad(p, features, batch));
  SG_HIP(hipMemcpyAsync(grad, p->grad, 4 * sg::policy::kParams, hipMemcpyDeviceToDevice, p->dev->stream));
  if (loss) {
    mlp_sum_kernel<<<1, kB, 0, p->dev->stream>>>(p->loss, batch, p->loss_mean);
    SG_HIP(hipMemcpyAsync(loss, p->loss_mean, 4, hipMemcpyDeviceToHost, p->dev->stream));
  }
  SG_HIP(hipStreamSynchronize(p->dev->stream));
  return SG_OK;
}

int sg_policy_train_step(sg_policy_t p, const float* features, const int32_t* labels, int32_t batch, float lr,
                         float* loss) {
  SG_CHECK(p && features && labels && batch > 0 && lr > 0.f, SG_ERR_INVALID, "sg_policy_train_step: bad argument");
  SG_TRY(run_forward(p, features, labels, batch));
  SG_TRY(run_grad(p, features, batch));
  ++p->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)p->step), c2 = 1.f - powf(b2, (float)p->step);
  mlp_adam_kernel<<<(sg::policy::kParams + kB - 1) / kB, kB, 0, p->dev->stream>>>(p->params, p->grad, p->m, p->v, lr,
                                                                                 b1, b2, 1e-8f, c1, c2);
  SG_HIP(hipGetLastError());
  if (loss) {
    mlp_sum_kernel<<<1, kB, 0, p->dev->stream>>>(p->loss, batch, p->loss_mean);
    SG_HIP(hipMemcpyAsync(loss, p->loss_mean, 4, hipMemcpyDeviceToHost, p->dev->stream));
  }
  SG_HIP(hipStreamSynchronize(p->dev->stream));
  return SG_OK;
}

} // extern "C"

namespace sg {
const float* policy_device_params(sg_policy_t p) { return p ? p->params : nullptr; }
} // namespace sg
