// gs_adam.hip -- one-launch Adam over all Gaussian parameter tensors.
//
// The optimiser of the training step (torch.optim.Adam with eps 1e-15 and one
// group per parameter, scene/gaussian_model_cf.py:256-268) updates ~59 floats
// per Gaussian; torch issues one multi-tensor launch per group.  Here every
// tensor is cut into 2048-element chunks and one grid walks all chunks,
// float4-vectorised (tensors are 16-byte aligned; ragged tails run scalar).
// Per element, in torch's order:
//   m = m + (g - m) * (1 - b1)            (lerp)
//   v = v * b2 + (1 - b2) * g * g
//   p = p - step * m / (sqrt(v) / sqrt(bc2) + eps),  step = lr / bc1
#include <cmath>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

constexpr int kAdamMax = 16;
constexpr int kAdamChunk = 2048;

struct AdamArgs {
  int ntensors;
  float* p[kAdamMax];
  const float* g[kAdamMax];
  float* m[kAdamMax];
  float* v[kAdamMax];
  long long n[kAdamMax];
  float step_size[kAdamMax];  // lr / (1 - b1^t)
  int chunk0[kAdamMax + 1];   // first chunk of each tensor
  float b1, b2, eps, inv_sqrt_bc2;
};

__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, float b1, float b2, float eps,
                                         float inv_sqrt_bc2, float step) {
  m = fmaf(g - m, 1.0f - b1, m);
  v = fmaf(v, b2, (1.0f - b2) * g * g);
  const float denom = sqrtf(v) * inv_sqrt_bc2 + eps;
  p = p - step * (m / denom);
}

__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
  const int c = blockIdx.x;
  int t = 0;
  while (t + 1 < a.ntensors && c >= a.chunk0[t + 1]) ++t;
  const long long base = (long long)(c - a.chunk0[t]) * kAdamChunk;
  const long long n = a.n[t];
  const long long end = base + kAdamChunk < n ? base + kAdamChunk : n;
  float* P = a.p[t];
  const float* G = a.g[t];
  float* M = a.m[t];
  float* V = a.v[t];
  const float st = a.step_size[t];
  // vector part
  const long long vend = base + ((end - base) & ~3LL);
  for (long long i = base + 4LL * threadIdx.x; i < vend; i += 4LL * 256) {
    float4 p = *reinterpret_cast<float4*>(P + i);
    const float4 g = *reinterpret_cast<const float4*>(G + i);
    float4 m = *reinterpret_cast<float4*>(M + i);
    float4 v = *reinterpret_cast<float4*>(V + i);
    adam_one(p.x, g.x, m.x, v.x, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, st);
    adam_one(p.y, g.y, m.y, v.y, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, st);
    adam_one(p.z, g.z, m.z, v.z, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, st);
    adam_one(p.w, g.w, m.w, v.w, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, st);
    *reinterpret_cast<float4*>(P + i) = p;
    *reinterpret_cast<float4*>(M + i) = m;
    *reinterpret_cast<float4*>(V + i) = v;
  }
  for (long long i = vend + threadIdx.x; i < end; i += 256) {
    float p = P[i], m = M[i], v = V[i];
    adam_one(p, G[i], m, v, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, st);
    P[i] = p;
    M[i] = m;
    V[i] = v;
  }
}

}  // namespace gs

using namespace gs;

extern "C" int gs_adam_step(int ntensors, float* const* params, const float* const* grads, float* const* exp_avg,
                            float* const* exp_avg_sq, const long long* numel, const float* lr, float beta1,
                            float beta2, float eps, int step, void* stream) {
  if (ntensors <= 0 || ntensors > kAdamMax || step < 1 || !params || !grads || !exp_avg || !exp_avg_sq || !numel ||
      !lr)
    return GS_E_ARG;
  AdamArgs a;
  a.ntensors = ntensors;
  long long chunks = 0;
  for (int k = 0; k < ntensors; ++k) {
    if (numel[k] < 0 || (numel[k] > 0 && (!params[k] || !grads[k] || !exp_avg[k] || !exp_avg_sq[k])))
      return GS_E_ARG;
    // 16-byte alignment for the vector path
    if (((uintptr_t)params[k] | (uintptr_t)grads[k] | (uintptr_t)exp_avg[k] | (uintptr_t)exp_avg_sq[k]) & 15)
      return GS_E_ARG;
    a.p[k] = params[k];
    a.g[k] = grads[k];
    a.m[k] = exp_avg[k];
    a.v[k] = exp_avg_sq[k];
    a.n[k] = numel[k];
    const double bc1 = 1.0 - std::pow((double)beta1, (double)step);
    a.step_size[k] = (float)(lr[k] / bc1);
    a.chunk0[k] = (int)chunks;
    chunks += (numel[k] + kAdamChunk - 1) / kAdamChunk;
  }
  a.chunk0[ntensors] = (int)chunks;
  if (chunks > 0x7fffffffLL) return GS_E_ARG;
  a.b1 = beta1;
  a.b2 = beta2;
  a.eps = eps;
  const double bc2 = 1.0 - std::pow((double)beta2, (double)step);
  a.inv_sqrt_bc2 = (float)(1.0 / std::sqrt(bc2));
  if (chunks == 0) return GS_OK;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)chunks), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
