// gs_loss.hip -- fused photometric loss of the training step:
//   loss = (1 - l) * mean|x - y| + l * (1 - mean SSIM(x, y)),  l = 0.2
// trainer/losses.py:98-136 with SSIM_V2 (:164-251): 11x11 Gaussian window
// (sigma 1.5, separable), zero padding, C1 = 0.01^2, C2 = 0.03^2.
//
//   k_ssim_fwd   per 32x16 output tile and channel: x, y with a 5-pixel halo
//                into LDS, horizontal then vertical 11-tap passes for
//                mu_x, mu_y, E[x^2], E[y^2], E[xy]; per pixel the SSIM term S
//                and the three partials dS/dmu_x, dS/dE[x^2], dS/dE[xy]
//                (written as maps for the backward); block sums of S, |x-y|
//   k_loss_final one workgroup: sums the block partials (fixed order) -> loss
//   k_ssim_bwd   window-correlates the three partial maps (the adjoint of the
//                zero-padded forward window) and combines them with x, y and
//                the L1 sign into dloss/dx, scaled by the upstream gradient.
#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

constexpr int kLTX = 32, kLTY = 16, kLR = 5;
constexpr int kLIW = kLTX + 2 * kLR, kLIH = kLTY + 2 * kLR;
constexpr float kC1 = 0.01f * 0.01f, kC2 = 0.03f * 0.03f;

struct Win {
  float g[11];
};

struct LossArgs {
  int C, H, W;
  float lambda;
  int clamp01;  // loss of clamp(pred, 0, 1), as the caller's image.clamp(0, 1)
  Win win;
  const float* x;  // prediction [C,H,W]
  const float* y;  // target [C,H,W]
  float* maps;     // [3][C][H][W]: dS/dmu_x, dS/dE[x^2], dS/dE[xy]
  float* partial;  // [2 * nblocks]: sum S, sum |x - y|
};

__device__ __forceinline__ float at_or_zero(const float* p, int H, int W, int r, int c) {
  return (r >= 0 && r < H && c >= 0 && c < W) ? p[r * W + c] : 0.0f;
}

__global__ __launch_bounds__(256) void k_ssim_fwd(LossArgs a) {
  __shared__ float sx[kLIH][kLIW];
  __shared__ float sy[kLIH][kLIW];
  __shared__ float sh[5][kLIH][kLTX];
  __shared__ float red[2][256];
  const int ch = blockIdx.z;
  const int r0 = blockIdx.y * kLTY - kLR, c0 = blockIdx.x * kLTX - kLR;
  const float* X = a.x + (size_t)ch * a.H * a.W;
  const float* Y = a.y + (size_t)ch * a.H * a.W;
  const int t = threadIdx.x;
  for (int k = t; k < kLIH * kLIW; k += 256) {
    const int r = k / kLIW, c = k - r * kLIW;
    const float xr = at_or_zero(X, a.H, a.W, r0 + r, c0 + c);
    sx[r][c] = a.clamp01 ? fminf(fmaxf(xr, 0.0f), 1.0f) : xr;
    sy[r][c] = at_or_zero(Y, a.H, a.W, r0 + r, c0 + c);
  }
  __syncthreads();
  for (int k = t; k < kLIH * kLTX; k += 256) {
    const int r = k / kLTX, c = k - r * kLTX;
    float h0 = 0.f, h1 = 0.f, h2 = 0.f, h3 = 0.f, h4 = 0.f;
#pragma unroll
    for (int j = 0; j < 11; ++j) {
      const float g = a.win.g[j];
      const float xv = sx[r][c + j], yv = sy[r][c + j];
      h0 = fmaf(g, xv, h0);
      h1 = fmaf(g, yv, h1);
      h2 = fmaf(g, xv * xv, h2);
      h3 = fmaf(g, yv * yv, h3);
      h4 = fmaf(g, xv * yv, h4);
    }
    sh[0][r][c] = h0;
    sh[1][r][c] = h1;
    sh[2][r][c] = h2;
    sh[3][r][c] = h3;
    sh[4][r][c] = h4;
  }
  __syncthreads();
  float sumS = 0.f, sumL1 = 0.f;
  const size_t HW = (size_t)a.H * a.W;
  for (int k = t; k < kLTY * kLTX; k += 256) {
    const int r = k / kLTX, c = k - r * kLTX;
    const int gr = blockIdx.y * kLTY + r, gc = blockIdx.x * kLTX + c;
    if (gr >= a.H || gc >= a.W) continue;
    float m1 = 0.f, m2 = 0.f, e11 = 0.f, e22 = 0.f, e12 = 0.f;
#pragma unroll
    for (int j = 0; j < 11; ++j) {
      const float g = a.win.g[j];
      m1 = fmaf(g, sh[0][r + j][c], m1);
      m2 = fmaf(g, sh[1][r + j][c], m2);
      e11 = fmaf(g, sh[2][r + j][c], e11);
      e22 = fmaf(g, sh[3][r + j][c], e22);
      e12 = fmaf(g, sh[4][r + j][c], e12);
    }
    const float s1 = e11 - m1 * m1, s2 = e22 - m2 * m2, s12 = e12 - m1 * m2;
    const float A1 = 2.f * m1 * m2 + kC1, A2 = 2.f * s12 + kC2;
    const float B1 = m1 * m1 + m2 * m2 + kC1, B2 = s1 + s2 + kC2;
    const float inv = 1.0f / (B1 * B2);
    const float S = A1 * A2 * inv;
    const size_t o = (size_t)ch * HW + (size_t)gr * a.W + gc;
    a.maps[o] = (2.f * m2 * (A2 - A1) - 2.f * m1 * S * (B2 - B1)) * inv;  // dS/dmu_x
    a.maps[(size_t)a.C * HW + o] = -S / B2;                             // dS/dE[x^2]
    a.maps[2 * (size_t)a.C * HW + o] = 2.f * A1 * inv;                  // dS/dE[xy]
    sumS += S;
    sumL1 += fabsf(sx[r + kLR][c + kLR] - sy[r + kLR][c + kLR]);
  }
  red[0][t] = sumS;
  red[1][t] = sumL1;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      red[0][t] += red[0][t + off];
      red[1][t] += red[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
    const int b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[2 * b] = red[0][0];
    a.partial[2 * b + 1] = red[1][0];
  }
}

// out[0] = loss, out[1] = mean SSIM, out[2] = mean |x - y|
__global__ __launch_bounds__(256) void k_loss_final(const float* partial, int nb, float inv_n, float lambda,
                                                    float* out) {
  __shared__ double red[2][256];
  double s = 0.0, l = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    s += (double)partial[2 * b];
    l += (double)partial[2 * b + 1];
  }
  red[0][threadIdx.x] = s;
  red[1][threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      red[0][threadIdx.x] += red[0][threadIdx.x + off];
      red[1][threadIdx.x] += red[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float ssim = (float)(red[0][0] * inv_n);
    const float l1 = (float)(red[1][0] * inv_n);
    out[0] = (1.0f - lambda) * l1 + lambda * (1.0f - ssim);
    out[1] = ssim;
    out[2] = l1;
  }
}

struct LossBwdArgs {
  int C, H, W;
  float lambda, inv_n;
  int clamp01;
  Win win;
  const float* x;
  const float* y;
  const float* maps;
  const float* grad_out;  // [1] upstream gradient of the loss (device)
  float* dx;
};

__global__ __launch_bounds__(256) void k_ssim_bwd(LossBwdArgs a) {
  __shared__ float sm[3][kLIH][kLIW];
  __shared__ float sh[3][kLIH][kLTX];
  const int ch = blockIdx.z;
  const size_t HW = (size_t)a.H * a.W;
  const int r0 = blockIdx.y * kLTY - kLR, c0 = blockIdx.x * kLTX - kLR;
  const int t = threadIdx.x;
  for (int q = 0; q < 3; ++q) {
    const float* Mq = a.maps + (size_t)q * a.C * HW + (size_t)ch * HW;
    for (int k = t; k < kLIH * kLIW; k += 256) {
      const int r = k / kLIW, c = k - r * kLIW;
      sm[q][r][c] = at_or_zero(Mq, a.H, a.W, r0 + r, c0 + c);
    }
  }
  __syncthreads();
  for (int k = t; k < kLIH * kLTX; k += 256) {
    const int r = k / kLTX, c = k - r * kLTX;
    float h0 = 0.f, h1 = 0.f, h2 = 0.f;
#pragma unroll
    for (int j = 0; j < 11; ++j) {
      const float g = a.win.g[j];
      h0 = fmaf(g, sm[0][r][c + j], h0);
      h1 = fmaf(g, sm[1][r][c + j], h1);
      h2 = fmaf(g, sm[2][r][c + j], h2);
    }
    sh[0][r][c] = h0;
    sh[1][r][c] = h1;
    sh[2][r][c] = h2;
  }
  __syncthreads();
  const float go = *a.grad_out;
  for (int k = t; k < kLTY * kLTX; k += 256) {
    const int r = k / kLTX, c = k - r * kLTX;
    const int gr = blockIdx.y * kLTY + r, gc = blockIdx.x * kLTX + c;
    if (gr >= a.H || gc >= a.W) continue;
    float ga = 0.f, gb = 0.f, gc2 = 0.f;
#pragma unroll
    for (int j = 0; j < 11; ++j) {
      const float g = a.win.g[j];
      ga = fmaf(g, sh[0][r + j][c], ga);
      gb = fmaf(g, sh[1][r + j][c], gb);
      gc2 = fmaf(g, sh[2][r + j][c], gc2);
    }
    const size_t o = (size_t)ch * HW + (size_t)gr * a.W + gc;
    const float xraw = a.x[o], yv = a.y[o];
    const float xv = a.clamp01 ? fminf(fmaxf(xraw, 0.0f), 1.0f) : xraw;
    // clamp's gradient passes where 0 <= x <= 1 (torch's clamp backward)
    const float pass = (!a.clamp01 || (xraw >= 0.0f && xraw <= 1.0f)) ? 1.0f : 0.0f;
    const float dssim = (ga + 2.f * xv * gb + yv * gc2) * a.inv_n;
    const float d = xv - yv;
    const float sgn = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
    a.dx[o] = pass * (go * ((1.0f - a.lambda) * sgn * a.inv_n - a.lambda * dssim));
  }
}

}  // namespace gs

using namespace gs;

static inline dim3 loss_grid(int C, int H, int W) {
  return dim3((unsigned)((W + kLTX - 1) / kLTX), (unsigned)((H + kLTY - 1) / kLTY), (unsigned)C);
}

extern "C" size_t gs_loss_workspace_bytes(int C, int H, int W) {
  if (C <= 0 || H <= 0 || W <= 0) return 0;
  const dim3 g = loss_grid(C, H, W);
  const size_t nb = (size_t)g.x * g.y * g.z;
  return align_up(3 * sizeof(float) * (size_t)C * H * W, 256) + align_up(2 * sizeof(float) * nb, 256);
}

extern "C" int gs_ssim_l1_forward(const float* pred, const float* target, int C, int H, int W,
                                  const float* window11, float lambda_dssim, int clamp01, float* out3, void* ws,
                                  size_t ws_bytes, void* stream) {
  if (!pred || !target || !window11 || !out3 || !ws || C <= 0 || H <= 0 || W <= 0) return GS_E_ARG;
  if (ws_bytes < gs_loss_workspace_bytes(C, H, W)) return GS_E_WORKSPACE;
  LossArgs a;
  a.C = C;
  a.H = H;
  a.W = W;
  a.lambda = lambda_dssim;
  a.clamp01 = clamp01;
  for (int k = 0; k < 11; ++k) a.win.g[k] = window11[k];
  a.x = pred;
  a.y = target;
  a.maps = (float*)ws;
  a.partial = (float*)((char*)ws + align_up(3 * sizeof(float) * (size_t)C * H * W, 256));
  const dim3 g = loss_grid(C, H, W);
  const int nb = (int)(g.x * g.y * g.z);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ssim_fwd, g, dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, s, a.partial, nb, 1.0f / ((float)C * H * W), lambda_dssim,
                     out3);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_ssim_l1_backward(const float* pred, const float* target, int C, int H, int W,
                                   const float* window11, float lambda_dssim, int clamp01, const float* grad_out,
                                   const void* ws, size_t ws_bytes, float* dpred, void* stream) {
  if (!pred || !target || !window11 || !grad_out || !ws || !dpred || C <= 0 || H <= 0 || W <= 0) return GS_E_ARG;
  if (ws_bytes < gs_loss_workspace_bytes(C, H, W)) return GS_E_WORKSPACE;
  LossBwdArgs a;
  a.C = C;
  a.H = H;
  a.W = W;
  a.lambda = lambda_dssim;
  a.clamp01 = clamp01;
  a.inv_n = 1.0f / ((float)C * H * W);
  for (int k = 0; k < 11; ++k) a.win.g[k] = window11[k];
  a.x = pred;
  a.y = target;
  a.maps = (const float*)ws;
  a.grad_out = grad_out;
  a.dx = dpred;
  hipLaunchKernelGGL(k_ssim_bwd, loss_grid(C, H, W), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
