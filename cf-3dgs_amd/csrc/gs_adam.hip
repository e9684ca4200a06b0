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
// A tensor may carry a second learning rate for the leading `lead` elements of
// every `row` elements: the SH coefficients of all Gaussians live in one
// [P,16,3] tensor whose first 3 floats per Gaussian are the reference's f_dc
// group (lr feature_lr) and the other 45 its f_rest group (feature_lr / 20).
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
  float lead_step[kAdamMax];  // lead_lr / (1 - b1^t)
  int lead[kAdamMax];         // elements [0, lead) of every row take lead_step
  int row[kAdamMax];          // 0: no lead (row % 4 == 0, numel < 2^31 otherwise)
  int chunk0[kAdamMax + 1];   // first chunk of each tensor
  float b1, b2, eps, inv_sqrt_bc2;
};

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
  const float lst = a.lead_step[t];
  const unsigned row = (unsigned)a.row[t], lead = (unsigned)a.lead[t];
  // step size of element i (row > 0 implies numel < 2^31)
  auto step_of = [&](long long i) -> float {
    return (row != 0u && (unsigned)i % row < lead) ? lst : st;
  };
  // vector part: the chunk's two float4 per thread are all loaded before any
  // is updated (8 loads in flight), streamed past the caches (each byte is
  // touched once).  row % 4 == 0, so the 4 elements of a float4 sit in one row.
  const long long vend = base + ((end - base) & ~3LL);
  auto upd4 = [&](float4& p, const float4& g, float4& m, float4& v, long long i) {
    float s0 = st, s1 = st, s2 = st, s3 = st;
    if (row != 0u) {
      const unsigned r = (unsigned)i % row;
      s0 = r < lead ? lst : st;
      s1 = r + 1u < lead ? lst : st;
      s2 = r + 2u < lead ? lst : st;
      s3 = r + 3u < lead ? lst : st;
    }
    adam_one(p.x, g.x, m.x, v.x, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s0);
    adam_one(p.y, g.y, m.y, v.y, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s1);
    adam_one(p.z, g.z, m.z, v.z, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s2);
    adam_one(p.w, g.w, m.w, v.w, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s3);
  };
  const long long i0 = base + 4LL * threadIdx.x, i1 = i0 + 4LL * 256;
  if (i1 + 4 <= vend) {
    float4 p0 = ld_nt(P + i0), p1 = ld_nt(P + i1);
    const float4 g0 = ld_nt(G + i0), g1 = ld_nt(G + i1);
    float4 m0 = ld_nt(M + i0), m1 = ld_nt(M + i1);
    float4 v0 = ld_nt(V + i0), v1 = ld_nt(V + i1);
    upd4(p0, g0, m0, v0, i0);
    upd4(p1, g1, m1, v1, i1);
    st_nt(P + i0, p0);
    st_nt(P + i1, p1);
    st_nt(M + i0, m0);
    st_nt(M + i1, m1);
    st_nt(V + i0, v0);
    st_nt(V + i1, v1);
  } else {
    for (long long i = i0; i < vend; i += 4LL * 256) {
      float4 p = *reinterpret_cast<float4*>(P + i);
      const float4 g = *reinterpret_cast<const float4*>(G + i);
      float4 m = *reinterpret_cast<float4*>(M + i);
      float4 v = *reinterpret_cast<float4*>(V + i);
      upd4(p, g, m, v, i);
      *reinterpret_cast<float4*>(P + i) = p;
      *reinterpret_cast<float4*>(M + i) = m;
      *reinterpret_cast<float4*>(V + i) = v;
    }
  }
  for (long long i = vend + threadIdx.x; i < end; i += 256) {
    float p = P[i], m = M[i], v = V[i];
    adam_one(p, G[i], m, v, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, step_of(i));
    P[i] = p;
    M[i] = m;
    V[i] = v;
  }
}

// ---------------------------------------------------------------------------
// SE(3) tangent parameters (lietorch LieGroupParameter semantics): the pose is
// Exp(xi) * X0 with xi = [tau, phi].  The Jacobian d pose / d xi is evaluated
// by forward-mode dual numbers through the closed form, one pass per tangent
// direction, in double precision; near phi = 0 the trigonometric ratios use
// their Taylor series in th^2 so the derivative stays finite at the identity.
struct Dual {
  double v, d;
};
__device__ inline Dual mk(double v) { return {v, 0.0}; }
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator*(double s, Dual a) { return {s * a.v, s * a.d}; }
__device__ inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ inline Dual dsin(Dual a) { return {sin(a.v), cos(a.v) * a.d}; }
__device__ inline Dual dcos(Dual a) { return {cos(a.v), -sin(a.v) * a.d}; }
__device__ inline Dual dsqrt(Dual a) {
  const double s = sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}

// out[7] = Exp(xi) * X0
__device__ void se3_exp_compose(const Dual xi[6], const double X0[7], Dual out[7]) {
  const Dual tau[3] = {xi[0], xi[1], xi[2]};
  const Dual phi[3] = {xi[3], xi[4], xi[5]};
  const Dual th2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
  Dual qa_s, qa_w, c1, c2;  // sin(th/2)/th, cos(th/2), (1-cos th)/th^2, (th-sin th)/th^3
  if (th2.v < 1e-6) {
    const Dual th4 = th2 * th2;
    qa_s = mk(0.5) - (1.0 / 48.0) * th2 + (1.0 / 3840.0) * th4;
    qa_w = mk(1.0) - (1.0 / 8.0) * th2 + (1.0 / 384.0) * th4;
    c1 = mk(0.5) - (1.0 / 24.0) * th2 + (1.0 / 720.0) * th4;
    c2 = mk(1.0 / 6.0) - (1.0 / 120.0) * th2 + (1.0 / 5040.0) * th4;
  } else {
    const Dual th = dsqrt(th2);
    const Dual half = 0.5 * th;
    qa_s = dsin(half) / th;
    qa_w = dcos(half);
    c1 = (mk(1.0) - dcos(th)) / th2;
    c2 = (th - dsin(th)) / (th2 * th);
  }
  const Dual qa[3] = {qa_s * phi[0], qa_s * phi[1], qa_s * phi[2]};
  // t_a = tau + c1 (phi x tau) + c2 (phi x (phi x tau))
  const Dual k1[3] = {phi[1] * tau[2] - phi[2] * tau[1], phi[2] * tau[0] - phi[0] * tau[2],
                      phi[0] * tau[1] - phi[1] * tau[0]};
  const Dual k2[3] = {phi[1] * k1[2] - phi[2] * k1[1], phi[2] * k1[0] - phi[0] * k1[2],
                      phi[0] * k1[1] - phi[1] * k1[0]};
  Dual ta[3];
  for (int k = 0; k < 3; ++k) ta[k] = tau[k] + c1 * k1[k] + c2 * k2[k];
  // R(qa) t0 = t0 + 2 w (u x t0) + 2 u x (u x t0)
  const Dual t0[3] = {mk(X0[0]), mk(X0[1]), mk(X0[2])};
  const Dual uxt[3] = {qa[1] * t0[2] - qa[2] * t0[1], qa[2] * t0[0] - qa[0] * t0[2], qa[0] * t0[1] - qa[1] * t0[0]};
  const Dual uuxt[3] = {qa[1] * uxt[2] - qa[2] * uxt[1], qa[2] * uxt[0] - qa[0] * uxt[2],
                        qa[0] * uxt[1] - qa[1] * uxt[0]};
  for (int k = 0; k < 3; ++k) out[k] = ta[k] + t0[k] + 2.0 * (qa_w * uxt[k]) + 2.0 * uuxt[k];
  // q = qa (x) q0 (Hamilton product, [x, y, z, w])
  const Dual u0[3] = {mk(X0[3]), mk(X0[4]), mk(X0[5])};
  const Dual w0 = mk(X0[6]);
  const Dual cr[3] = {qa[1] * u0[2] - qa[2] * u0[1], qa[2] * u0[0] - qa[0] * u0[2], qa[0] * u0[1] - qa[1] * u0[0]};
  for (int k = 0; k < 3; ++k) out[3 + k] = qa_w * u0[k] + w0 * qa[k] + cr[k];
  out[6] = qa_w * w0 - (qa[0] * u0[0] + qa[1] * u0[1] + qa[2] * u0[2]);
}

__device__ void se3_load(const float* xi, const float* X0, double x[6], double x0[7]) {
  for (int k = 0; k < 6; ++k) x[k] = xi[k];
  if (X0) {
    for (int k = 0; k < 7; ++k) x0[k] = X0[k];
  } else {
    for (int k = 0; k < 7; ++k) x0[k] = k == 6 ? 1.0 : 0.0;
  }
}

__device__ void se3_pose(const double x[6], const double x0[7], float* pose) {
  Dual xi[6], out[7];
  for (int k = 0; k < 6; ++k) xi[k] = mk(x[k]);
  se3_exp_compose(xi, x0, out);
  for (int k = 0; k < 7; ++k) pose[k] = (float)out[k].v;
}

__device__ void se3_tangent_grad(const double x[6], const double x0[7], const float* gpose, double g[6]) {
  for (int j = 0; j < 6; ++j) {
    Dual xi[6], out[7];
    for (int k = 0; k < 6; ++k) xi[k] = {x[k], k == j ? 1.0 : 0.0};
    se3_exp_compose(xi, x0, out);
    double s = 0.0;
    for (int k = 0; k < 7; ++k) s += out[k].d * (double)gpose[k];
    g[j] = s;
  }
}

__global__ __launch_bounds__(64) void k_se3_retract(int n, const float* xi, const float* X0, float* pose) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  double x[6], x0[7];
  se3_load(xi + 6 * i, X0 ? X0 + 7 * i : nullptr, x, x0);
  se3_pose(x, x0, pose + 7 * i);
}

__global__ __launch_bounds__(64) void k_se3_grad(int n, const float* xi, const float* X0, const float* gpose,
                                                 float* gxi) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  double x[6], x0[7], g[6];
  se3_load(xi + 6 * i, X0 ? X0 + 7 * i : nullptr, x, x0);
  se3_tangent_grad(x, x0, gpose + 7 * i, g);
  for (int k = 0; k < 6; ++k) gxi[6 * i + k] = (float)g[k];
}

// One workgroup walks all n poses (one per concurrent solve, so n is small):
// the shared step count is read once before and incremented once after.
__global__ __launch_bounds__(64) void k_se3_adam(int n, float* xi, const float* X0, const float* gpose, float* M,
                                                 float* V, int* step, float lr, float b1, float b2, float eps,
                                                 float* pose) {
  __shared__ float s_step, s_inv;
  if (threadIdx.x == 0) {
    const double t = (double)(*step + 1);
    s_step = (float)(lr / (1.0 - pow((double)b1, t)));
    s_inv = (float)(1.0 / sqrt(1.0 - pow((double)b2, t)));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 64) {
    double x[6], x0[7], g[6];
    se3_load(xi + 6 * i, X0 ? X0 + 7 * i : nullptr, x, x0);
    se3_tangent_grad(x, x0, gpose + 7 * i, g);
    for (int k = 0; k < 6; ++k) {
      float p = xi[6 * i + k], m = M[6 * i + k], v = V[6 * i + k];
      adam_one(p, (float)g[k], m, v, b1, b2, eps, s_inv, s_step);
      xi[6 * i + k] = p;
      M[6 * i + k] = m;
      V[6 * i + k] = v;
      x[k] = p;
    }
    se3_pose(x, x0, pose + 7 * i);
  }
  __syncthreads();
  if (threadIdx.x == 0) *step += 1;
}

}  // namespace gs

using namespace gs;

extern "C" int gs_se3_retract(int n, const float* xi, const float* X0, float* pose_out, void* stream) {
  if (n < 0 || (n > 0 && (!xi || !pose_out))) return GS_E_ARG;
  if (n == 0) return GS_OK;
  hipLaunchKernelGGL(k_se3_retract, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, n, xi, X0, pose_out);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_se3_grad(int n, const float* xi, const float* X0, const float* grad_pose, float* grad_xi,
                           void* stream) {
  if (n < 0 || (n > 0 && (!xi || !grad_pose || !grad_xi))) return GS_E_ARG;
  if (n == 0) return GS_OK;
  hipLaunchKernelGGL(k_se3_grad, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, n, xi, X0, grad_pose,
                     grad_xi);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_se3_adam(int n, float* xi, const float* X0, const float* grad_pose, float* exp_avg,
                           float* exp_avg_sq, int* step, float lr, float beta1, float beta2, float eps,
                           float* pose_out, void* stream) {
  if (n < 0 || (n > 0 && (!xi || !grad_pose || !exp_avg || !exp_avg_sq || !step || !pose_out))) return GS_E_ARG;
  if (n == 0) return GS_OK;
  hipLaunchKernelGGL(k_se3_adam, dim3(1), dim3(64), 0, (hipStream_t)stream, n, xi, X0, grad_pose, exp_avg,
                     exp_avg_sq, step, lr, beta1, beta2, eps, pose_out);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_adam_step_rows(int ntensors, float* const* params, const float* const* grads,
                                 float* const* exp_avg, float* const* exp_avg_sq, const long long* numel,
                                 const float* lr, const float* lead_lr, const int* lead, const int* row, float beta1,
                                 float beta2, float eps, int step, void* stream) {
  if (ntensors <= 0 || ntensors > kAdamMax || step < 1 || !params || !grads || !exp_avg || !exp_avg_sq || !numel ||
      !lr)
    return GS_E_ARG;
  if ((lead_lr || lead || row) && !(lead_lr && lead && row)) return GS_E_ARG;
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
    a.row[k] = row ? row[k] : 0;
    a.lead[k] = row ? lead[k] : 0;
    a.lead_step[k] = row ? (float)(lead_lr[k] / bc1) : 0.0f;
    if (a.row[k] < 0 || a.row[k] % 4 || a.lead[k] < 0 || a.lead[k] > a.row[k] ||
        (a.row[k] > 0 && numel[k] >= (1LL << 31)))
      return GS_E_ARG;
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

extern "C" int gs_adam_step(int ntensors, float* const* params, const float* const* grads, float* const* exp_avg,
                            float* const* exp_avg_sq, const long long* numel, const float* lr, float beta1,
                            float beta2, float eps, int step, void* stream) {
  return gs_adam_step_rows(ntensors, params, grads, exp_avg, exp_avg_sq, numel, lr, nullptr, nullptr, nullptr, beta1,
                           beta2, eps, step, stream);
}
