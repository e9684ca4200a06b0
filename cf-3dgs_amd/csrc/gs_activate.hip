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

// The three backward rules, shared by the plain backward and the one that
// applies Adam (so both round identically).
__device__ __forceinline__ float act_bwd_exp(float g, float s) { return g * s; }
__device__ __forceinline__ float act_bwd_sigmoid(float g, float o) { return (g * (1.0f - o)) * o; }
__device__ __forceinline__ float4 act_bwd_normalize(float4 x, float4 g) {
  const float n = sqrtf(((x.x * x.x + x.y * x.y) + x.z * x.z) + x.w * x.w);
  const float c = fmaxf(n, kNormalizeEps);
  const float cc = c * c;
  const float dc = (((-g.x * x.x) / cc + (-g.y * x.y) / cc) + (-g.z * x.z) / cc) + (-g.w * x.w) / cc;
  const float dn = n >= kNormalizeEps ? dc : 0.0f;
  const float k = n > 0.0f ? dn / n : 0.0f;
  return make_float4(g.x / c + x.x * k, g.y / c + x.y * k, g.z / c + x.z * k, g.w / c + x.w * k);
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
      o[0] = act_bwd_exp(g[0], s[0]);
      o[1] = act_bwd_exp(g[1], s[1]);
      o[2] = act_bwd_exp(g[2], s[2]);
    } else {
      o[0] = o[1] = o[2] = 0.0f;
    }
  }
  if (d_r) {
    if (g_r) {
      d_r[i] = act_bwd_normalize(raw_rot[i], g_r[i]);
    } else {
      d_r[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  if (d_o) {
    if (g_o) {
      d_o[i] = act_bwd_sigmoid(g_o[i], opac[i]);
    } else {
      d_o[i] = 0.0f;
    }
  }
}

struct ActAdam {
  float* raw_scale;  // parameters, updated in place
  float4* raw_rot;
  float* raw_opac;
  float* sm;
  float* sv;
  float4* rm;
  float4* rv;
  float* om;
  float* ov;
  float s_step, r_step, o_step, b1, b2, eps, inv_sqrt_bc2;
};

// The backward of the three activations followed by one Adam step of the raw
// parameters: the raw gradients are never written.  Every load is issued
// before the first store (the arrays may not alias, but the compiler cannot
// know that), so each thread waits for memory once.
__global__ __launch_bounds__(256) void k_activate_bwd_adam(int P, const float* __restrict__ scales,
                                                           const float* __restrict__ opac,
                                                           const float* __restrict__ g_s,
                                                           const float4* __restrict__ g_r,
                                                           const float* __restrict__ g_o, ActAdam a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  float ps[3], ms[3], vs[3], gs[3], sc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const long long e = 3LL * i + k;
    ps[k] = a.raw_scale[e];
    ms[k] = a.sm[e];
    vs[k] = a.sv[e];
    gs[k] = g_s[e];
    sc[k] = scales[e];
  }
  float4 pr = a.raw_rot[i], mr = a.rm[i], vr = a.rv[i];
  const float4 gr = g_r[i];
  float po = a.raw_opac[i], mo = a.om[i], vo = a.ov[i];
  const float go = g_o[i], oo = opac[i];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    adam_one(ps[k], act_bwd_exp(gs[k], sc[k]), ms[k], vs[k], a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.s_step);
  const float4 d = act_bwd_normalize(pr, gr);
  adam_one(pr.x, d.x, mr.x, vr.x, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.r_step);
  adam_one(pr.y, d.y, mr.y, vr.y, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.r_step);
  adam_one(pr.z, d.z, mr.z, vr.z, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.r_step);
  adam_one(pr.w, d.w, mr.w, vr.w, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.r_step);
  adam_one(po, act_bwd_sigmoid(go, oo), mo, vo, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.o_step);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const long long e = 3LL * i + k;
    a.raw_scale[e] = ps[k];
    a.sm[e] = ms[k];
    a.sv[e] = vs[k];
  }
  a.raw_rot[i] = pr;
  a.rm[i] = mr;
  a.rv[i] = vr;
  a.raw_opac[i] = po;
  a.om[i] = mo;
  a.ov[i] = vo;
}

}  // namespace gs

using namespace gs;

extern "C" int gs_activate_backward_adam(int P, float* raw_scale, float* raw_rot, float* raw_opacity,
                                         const float* scales, const float* opacities, const float* dL_dscales,
                                         const float* dL_drotations, const float* dL_dopacities,
                                         const gs_adam_act_t* adam, void* stream) {
  if (P < 0 || !adam || adam->step < 1) return GS_E_ARG;
  if (P == 0) return GS_OK;
  if (!raw_scale || !raw_rot || !raw_opacity || !scales || !opacities || !dL_dscales || !dL_drotations ||
      !dL_dopacities || !adam->scale_m || !adam->scale_v || !adam->rot_m || !adam->rot_v || !adam->opacity_m ||
      !adam->opacity_v)
    return GS_E_ARG;
  if (((uintptr_t)raw_rot | (uintptr_t)dL_drotations | (uintptr_t)adam->rot_m | (uintptr_t)adam->rot_v) & 15)
    return GS_E_ARG;
  // bias corrections as gs_adam_step_rows computes them
  const double bc1 = 1.0 - std::pow((double)adam->beta1, (double)adam->step);
  const double bc2 = 1.0 - std::pow((double)adam->beta2, (double)adam->step);
  ActAdam a;
  a.raw_scale = raw_scale;
  a.raw_rot = reinterpret_cast<float4*>(raw_rot);
  a.raw_opac = raw_opacity;
  a.sm = adam->scale_m;
  a.sv = adam->scale_v;
  a.rm = reinterpret_cast<float4*>(adam->rot_m);
  a.rv = reinterpret_cast<float4*>(adam->rot_v);
  a.om = adam->opacity_m;
  a.ov = adam->opacity_v;
  a.s_step = (float)(adam->scale_lr / bc1);
  a.r_step = (float)(adam->rot_lr / bc1);
  a.o_step = (float)(adam->opacity_lr / bc1);
  a.b1 = adam->beta1;
  a.b2 = adam->beta2;
  a.eps = adam->eps;
  a.inv_sqrt_bc2 = (float)(1.0 / std::sqrt(bc2));
  hipLaunchKernelGGL(k_activate_bwd_adam, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, (hipStream_t)stream, P,
                     scales, opacities, dL_dscales, reinterpret_cast<const float4*>(dL_drotations), dL_dopacities, a);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

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
