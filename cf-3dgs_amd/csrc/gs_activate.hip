// gs_activate.hip -- the Gaussian model's parameter activations, forward and
// backward, one launch each.
//
// CFGaussianModel's getters (scene/gaussian_model_cf.py:48-56 setup_functions,
// :116-121 get_scaling / get_rotation, :168-169 get_opacity):
//   scales    = exp(raw_scale)                        [P,3]
//   rotations = raw_rot / max(||raw_rot||_2, 1e-12)   [P,4]  (F.normalize, dim 1)
//   opacities = sigmoid(raw_opacity)                  [P,1]
// In torch these are ~5 kernels forward and ~12 backward per training step
// (the backward of F.normalize alone is norm, clamp_min, expand and div
// backward, six kernels over [P,4]).  Here one thread owns one Gaussian and
// reads and writes each of its floats once.
//
// The backward follows torch's chain rule term by term:
//   exp:       d raw = g * scale
//   sigmoid:   d raw = g * (1 - o) * o
//   normalize: y = x / c, c = max(n, eps), n = ||x||
//              dc = -sum_i g_i x_i / (c c)
//              dn = dc where n >= eps, else 0       (clamp_min passes at n >= eps)
//              dx = g / c + x * (dn / n)            (the norm term is 0 at n = 0)
#include <cmath>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

constexpr float kNormalizeEps = 1e-12f;

__global__ __launch_bounds__(256) void k_activate_fwd(int P, const float* __restrict__ raw_scale,
                                                      const float4* __restrict__ raw_rot,
                                                      const float* __restrict__ raw_opac, float* __restrict__ scales,
                                                      float4* __restrict__ rots, float* __restrict__ opac) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  if (scales) {
    const float* s = raw_scale + 3LL * i;
    float* o = scales + 3LL * i;
    o[0] = expf(s[0]);
    o[1] = expf(s[1]);
    o[2] = expf(s[2]);
  }
  if (rots) {
    const float4 q = raw_rot[i];
    const float n = sqrtf(((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w);
    const float c = fmaxf(n, kNormalizeEps);
    rots[i] = make_float4(q.x / c, q.y / c, q.z / c, q.w / c);
  }
  if (opac) opac[i] = 1.0f / (1.0f + expf(-raw_opac[i]));
}

__global__ __launch_bounds__(256) void k_activate_bwd(int P, const float4* __restrict__ raw_rot,
                                                      const float* __restrict__ scales,
                                                      const float* __restrict__ opac, const float* __restrict__ g_s,
                                                      const float4* __restrict__ g_r, const float* __restrict__ g_o,
                                                      float* __restrict__ d_s, float4* __restrict__ d_r,
                                                      float* __restrict__ d_o) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  if (d_s) {
    float* o = d_s + 3LL * i;
    if (g_s) {
      const float* g = g_s + 3LL * i;
      const float* s = scales + 3LL * i;
      o[0] = g[0] * s[0];
      o[1] = g[1] * s[1];
      o[2] = g[2] * s[2];
    } else {
      o[0] = o[1] = o[2] = 0.0f;
    }
  }
  if (d_r) {
    if (g_r) {
      const float4 x = raw_rot[i];
      const float4 g = g_r[i];
      const float n = sqrtf(((x.x * x.x + x.y * x.y) + x.z * x.z) + x.w * x.w);
      const float c = fmaxf(n, kNormalizeEps);
      const float cc = c * c;
      const float dc = (((-g.x * x.x) / cc + (-g.y * x.y) / cc) + (-g.z * x.z) / cc) + (-g.w * x.w) / cc;
      const float dn = n >= kNormalizeEps ? dc : 0.0f;
      const float k = n > 0.0f ? dn / n : 0.0f;
      d_r[i] = make_float4(g.x / c + x.x * k, g.y / c + x.y * k, g.z / c + x.z * k, g.w / c + x.w * k);
    } else {
      d_r[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  if (d_o) {
    if (g_o) {
      const float o = opac[i];
      d_o[i] = (g_o[i] * (1.0f - o)) * o;
    } else {
      d_o[i] = 0.0f;
    }
  }
}

}  // namespace gs

using namespace gs;

extern "C" int gs_activate_forward(int P, const float* raw_scale, const float* raw_rot, const float* raw_opacity,
                                   float* scales, float* rotations, float* opacities, void* stream) {
  if (P < 0) return GS_E_ARG;
  if ((scales && !raw_scale) || (rotations && !raw_rot) || (opacities && !raw_opacity)) return GS_E_ARG;
  if (((uintptr_t)raw_rot | (uintptr_t)rotations) & 15) return GS_E_ARG;
  if (P == 0 || (!scales && !rotations && !opacities)) return GS_OK;
  hipLaunchKernelGGL(k_activate_fwd, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, (hipStream_t)stream, P,
                     raw_scale, reinterpret_cast<const float4*>(raw_rot), raw_opacity, scales,
                     reinterpret_cast<float4*>(rotations), opacities);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_activate_backward(int P, const float* raw_rot, const float* scales, const float* opacities,
                                    const float* dL_dscales, const float* dL_drotations, const float* dL_dopacities,
                                    float* dL_draw_scale, float* dL_draw_rot, float* dL_draw_opacity, void* stream) {
  if (P < 0) return GS_E_ARG;
  if ((dL_draw_scale && dL_dscales && !scales) || (dL_draw_rot && dL_drotations && !raw_rot) ||
      (dL_draw_opacity && dL_dopacities && !opacities))
    return GS_E_ARG;
  if (((uintptr_t)raw_rot | (uintptr_t)dL_drotations | (uintptr_t)dL_draw_rot) & 15) return GS_E_ARG;
  if (P == 0 || (!dL_draw_scale && !dL_draw_rot && !dL_draw_opacity)) return GS_OK;
  hipLaunchKernelGGL(k_activate_bwd, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, (hipStream_t)stream, P,
                     reinterpret_cast<const float4*>(raw_rot), scales, opacities, dL_dscales,
                     reinterpret_cast<const float4*>(dL_drotations), dL_dopacities, dL_draw_scale,
                     reinterpret_cast<float4*>(dL_draw_rot), dL_draw_opacity);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
