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
//                (gs_ssim_l1_forward_backward has the backward's first
//                workgroup do it instead)
//   k_ssim_bwd   window-correlates the three partial maps (the adjoint of the
//                zero-padded forward window) and combines them with x, y and
//                the L1 sign into dloss/dx, scaled by the upstream gradient.
#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

constexpr int kLTX = 32, kLTY = 16, kLR = 5;
constexpr int kLIW = kLTX + 2 * kLR, kLIH = kLTY + 2 * kLR;
constexpr float kC1 = 0.01f * 0.01f, kC2 = 0.03f * 0.03f;
static_assert(kLTX * kLTY == 2 * 256, "the vertical passes give each of 256 threads two outputs");

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

// The zero-padded (kLIH x kLIW) window of a plane whose first pixel is (r0,
// c0), for 256 threads: every load of a thread is issued before its first
// LDS store (addresses clamped into the plane, out-of-plane values zeroed),
// instead of one load-to-store round trip per element.
template <int NQ>
__device__ __forceinline__ void stage_halo(float (*dst[NQ])[kLIW], const float* const src[NQ], int H, int W, int r0,
                                           int c0, int clamp_first) {
  constexpr int kPer = (kLIH * kLIW + 255) / 256;
  float v[NQ][kPer];
  int idx[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int k = (int)threadIdx.x + 256 * i;
    const int kk = k < kLIH * kLIW ? k : kLIH * kLIW - 1;
    const int r = kk / kLIW, c = kk - r * kLIW;
    const int gr = r0 + r, gc = c0 + c;
    const bool in = gr >= 0 && gr < H && gc >= 0 && gc < W;
    const int o = min(max(gr, 0), H - 1) * W + min(max(gc, 0), W - 1);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const float x = src[q][o];
      v[q][i] = in ? x : 0.0f;
    }
    idx[i] = kk;
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int r = idx[i] / kLIW, c = idx[i] - r * kLIW;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const float x = v[q][i];
      dst[q][r][c] = (q == 0 && clamp_first) ? fminf(fmaxf(x, 0.0f), 1.0f) : x;
    }
  }
}

// Tile (bx, by) of channel bz of a (gx x gy x C) grid of one image (the
// kernels pass their block index; a launch over many images, each image's
// block coordinates).
__device__ __forceinline__ void ssim_fwd_block(const LossArgs& a, int bx, int by, int bz, int gx, int gy) {
  __shared__ float sx[kLIH][kLIW];
  __shared__ float sy[kLIH][kLIW];
  __shared__ float sh[5][kLIH][kLTX];
  __shared__ float red[2][4];
  const int ch = bz;
  const int r0 = by * kLTY - kLR, c0 = bx * kLTX - kLR;
  const float* X = a.x + (size_t)ch * a.H * a.W;
  const float* Y = a.y + (size_t)ch * a.H * a.W;
  const int t = threadIdx.x;
  {
    float (*dst[2])[kLIW] = {sx, sy};
    const float* const src[2] = {X, Y};
    stage_halo<2>(dst, src, a.H, a.W, r0, c0, a.clamp01);
  }
  __syncthreads();
  // horizontal 11-tap pass, two neighbouring outputs per thread: they share
  // ten of their twelve inputs (the per-output sum order is unchanged)
  for (int k = t; k < kLIH * (kLTX / 2); k += 256) {
    const int r = k / (kLTX / 2), c = 2 * (k - r * (kLTX / 2));
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f;
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f, b4 = 0.f;
    float xp = sx[r][c], yp = sy[r][c];
#pragma unroll
    for (int j = 0; j < 11; ++j) {
      const float g = a.win.g[j];
      const float xn = sx[r][c + j + 1], yn = sy[r][c + j + 1];
      a0 = fmaf(g, xp, a0);
      a1 = fmaf(g, yp, a1);
      a2 = fmaf(g, xp * xp, a2);
      a3 = fmaf(g, yp * yp, a3);
      a4 = fmaf(g, xp * yp, a4);
      b0 = fmaf(g, xn, b0);
      b1 = fmaf(g, yn, b1);
      b2 = fmaf(g, xn * xn, b2);
      b3 = fmaf(g, yn * yn, b3);
      b4 = fmaf(g, xn * yn, b4);
      xp = xn;
      yp = yn;
    }
    sh[0][r][c] = a0;
    sh[1][r][c] = a1;
    sh[2][r][c] = a2;
    sh[3][r][c] = a3;
    sh[4][r][c] = a4;
    sh[0][r][c + 1] = b0;
    sh[1][r][c + 1] = b1;
    sh[2][r][c + 1] = b2;
    sh[3][r][c + 1] = b3;
    sh[4][r][c + 1] = b4;
  }
  __syncthreads();
  float sumS = 0.f, sumL1 = 0.f;
  const size_t HW = (size_t)a.H * a.W;
  // vertical pass, two vertically neighbouring outputs per thread (256 pairs)
  {
    const int r = 2 * (t / kLTX), c = t - (t / kLTX) * kLTX;
    float m1[2] = {0.f, 0.f}, m2[2] = {0.f, 0.f}, e11[2] = {0.f, 0.f}, e22[2] = {0.f, 0.f}, e12[2] = {0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      const float v0 = sh[0][r + j][c], v1 = sh[1][r + j][c], v2 = sh[2][r + j][c], v3 = sh[3][r + j][c],
                  v4 = sh[4][r + j][c];
      if (j < 11) {
        const float g = a.win.g[j];
        m1[0] = fmaf(g, v0, m1[0]);
        m2[0] = fmaf(g, v1, m2[0]);
        e11[0] = fmaf(g, v2, e11[0]);
        e22[0] = fmaf(g, v3, e22[0]);
        e12[0] = fmaf(g, v4, e12[0]);
      }
      if (j > 0) {
        const float g = a.win.g[j - 1];
        m1[1] = fmaf(g, v0, m1[1]);
        m2[1] = fmaf(g, v1, m2[1]);
        e11[1] = fmaf(g, v2, e11[1]);
        e22[1] = fmaf(g, v3, e22[1]);
        e12[1] = fmaf(g, v4, e12[1]);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int gr = by * kLTY + r + u, gc = bx * kLTX + c;
      if (gr >= a.H || gc >= a.W) continue;
      const float s1 = e11[u] - m1[u] * m1[u], s2 = e22[u] - m2[u] * m2[u], s12 = e12[u] - m1[u] * m2[u];
      const float A1 = 2.f * m1[u] * m2[u] + kC1, A2 = 2.f * s12 + kC2;
      const float B1 = m1[u] * m1[u] + m2[u] * m2[u] + kC1, B2 = s1 + s2 + kC2;
      const float inv = 1.0f / (B1 * B2);
      const float S = A1 * A2 * inv;
      const size_t o = (size_t)ch * HW + (size_t)gr * a.W + gc;
      a.maps[o] = (2.f * m2[u] * (A2 - A1) - 2.f * m1[u] * S * (B2 - B1)) * inv;  // dS/dmu_x
      a.maps[(size_t)a.C * HW + o] = -S / B2;                                      // dS/dE[x^2]
      a.maps[2 * (size_t)a.C * HW + o] = 2.f * A1 * inv;                           // dS/dE[xy]
      sumS += S;
      sumL1 += fabsf(sx[r + u + kLR][c + kLR] - sy[r + u + kLR][c + kLR]);
    }
  }
  // block sums: wave shuffles, then the four wave partials in order
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sumS += __shfl_down(sumS, off, 64);
    sumL1 += __shfl_down(sumL1, off, 64);
  }
  if ((t & 63) == 0) {
    red[0][t >> 6] = sumS;
    red[1][t >> 6] = sumL1;
  }
  __syncthreads();
  if (t == 0) {
    const int b = (bz * gy + by) * gx + bx;
    a.partial[2 * b] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    a.partial[2 * b + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
  }
}

__global__ __launch_bounds__(256) void k_ssim_fwd(LossArgs a) {
  ssim_fwd_block(a, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y);
}

// out[0] = loss, out[1] = mean SSIM, out[2] = mean |x - y|, by one
// 256-thread workgroup (red: its [2][256] doubles of LDS)
__device__ __forceinline__ void loss_final_block(const float* partial, int nb, float inv_n, float lambda, float* out,
                                                 double (*red)[256]) {
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

__global__ __launch_bounds__(256) void k_loss_final(const float* partial, int nb, float inv_n, float lambda,
                                                    float* out) {
  __shared__ double red[2][256];
  loss_final_block(partial, nb, inv_n, lambda, out, red);
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
  // gs_ssim_l1_forward_backward: the first workgroup also reduces the
  // forward's block partials into out3 (the gradient does not need the loss
  // value), which spares the separate k_loss_final launch
  const float* partial;
  int nb;
  float* out3;
};

__device__ __forceinline__ void ssim_bwd_block(const LossBwdArgs& a, int bx, int by, int bz) {
  __shared__ float sm[3][kLIH][kLIW];
  __shared__ float sh[3][kLIH][kLTX];
  if (a.out3 && bx == 0 && by == 0 && bz == 0) {
    __shared__ double red[2][256];
    loss_final_block(a.partial, a.nb, a.inv_n, a.lambda, a.out3, red);
  }
  const int ch = bz;
  const size_t HW = (size_t)a.H * a.W;
  const int r0 = by * kLTY - kLR, c0 = bx * kLTX - kLR;
  const int t = threadIdx.x;
  {
    float (*dst[3])[kLIW] = {sm[0], sm[1], sm[2]};
    const float* const src[3] = {a.maps + (size_t)ch * HW, a.maps + (size_t)a.C * HW + (size_t)ch * HW,
                                 a.maps + 2 * (size_t)a.C * HW + (size_t)ch * HW};
    stage_halo<3>(dst, src, a.H, a.W, r0, c0, 0);
  }
  __syncthreads();
  // horizontal pass of the three maps, two neighbouring outputs per thread
  for (int k = t; k < kLIH * (kLTX / 2); k += 256) {
    const int r = k / (kLTX / 2), c = 2 * (k - r * (kLTX / 2));
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, b0 = 0.f, b1 = 0.f, b2 = 0.f;
    float p0 = sm[0][r][c], p1 = sm[1][r][c], p2 = sm[2][r][c];
#pragma unroll
    for (int j = 0; j < 11; ++j) {
      const float g = a.win.g[j];
      const float n0 = sm[0][r][c + j + 1], n1 = sm[1][r][c + j + 1], n2 = sm[2][r][c + j + 1];
      a0 = fmaf(g, p0, a0);
      a1 = fmaf(g, p1, a1);
      a2 = fmaf(g, p2, a2);
      b0 = fmaf(g, n0, b0);
      b1 = fmaf(g, n1, b1);
      b2 = fmaf(g, n2, b2);
      p0 = n0;
      p1 = n1;
      p2 = n2;
    }
    sh[0][r][c] = a0;
    sh[1][r][c] = a1;
    sh[2][r][c] = a2;
    sh[0][r][c + 1] = b0;
    sh[1][r][c + 1] = b1;
    sh[2][r][c + 1] = b2;
  }
  __syncthreads();
  const float go = *a.grad_out;
  // vertical pass, two vertically neighbouring outputs per thread
  const int r = 2 * (t / kLTX), c = t - (t / kLTX) * kLTX;
  float ga[2] = {0.f, 0.f}, gb[2] = {0.f, 0.f}, gc2[2] = {0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    const float v0 = sh[0][r + j][c], v1 = sh[1][r + j][c], v2 = sh[2][r + j][c];
    if (j < 11) {
      const float g = a.win.g[j];
      ga[0] = fmaf(g, v0, ga[0]);
      gb[0] = fmaf(g, v1, gb[0]);
      gc2[0] = fmaf(g, v2, gc2[0]);
    }
    if (j > 0) {
      const float g = a.win.g[j - 1];
      ga[1] = fmaf(g, v0, ga[1]);
      gb[1] = fmaf(g, v1, gb[1]);
      gc2[1] = fmaf(g, v2, gc2[1]);
    }
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int gr = by * kLTY + r + u, gcol = bx * kLTX + c;
    if (gr >= a.H || gcol >= a.W) continue;
    const size_t o = (size_t)ch * HW + (size_t)gr * a.W + gcol;
    const float xraw = a.x[o], yv = a.y[o];
    const float xv = a.clamp01 ? fminf(fmaxf(xraw, 0.0f), 1.0f) : xraw;
    // clamp's gradient passes where 0 <= x <= 1 (torch's clamp backward)
    const float pass = (!a.clamp01 || (xraw >= 0.0f && xraw <= 1.0f)) ? 1.0f : 0.0f;
    const float dssim = (ga[u] + 2.f * xv * gb[u] + yv * gc2[u]) * a.inv_n;
    const float d = xv - yv;
    const float sgn = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
    a.dx[o] = pass * (go * ((1.0f - a.lambda) * sgn * a.inv_n - a.lambda * dssim));
  }
}

__global__ __launch_bounds__(256) void k_ssim_bwd(LossBwdArgs a) { ssim_bwd_block(a, blockIdx.x, blockIdx.y, blockIdx.z); }

// ---- the loss of many images in one launch of each kernel ------------------
// Image f owns blocks [prefix[f], prefix[f + 1]) of a 1-D grid, in the order
// of its own launch (x fastest, then y, then the channel).
__device__ __forceinline__ int image_of(const uint32_t* prefix, int n, uint32_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_ssim_fwd_multi(const uint32_t* prefix, const LossArgs* args, int n) {
  const uint32_t b = blockIdx.x;
  const int f = image_of(prefix, n, b);
  const LossArgs a = args[f];  // a copy: loaded once, not re-read after the stores to the maps
  const int gx = (a.W + kLTX - 1) / kLTX, gy = (a.H + kLTY - 1) / kLTY;
  const int l = (int)(b - prefix[f]);
  ssim_fwd_block(a, l % gx, (l / gx) % gy, l / (gx * gy), gx, gy);
}

__global__ __launch_bounds__(256) void k_ssim_bwd_multi(const uint32_t* prefix, const LossBwdArgs* args, int n) {
  const uint32_t b = blockIdx.x;
  const int f = image_of(prefix, n, b);
  const LossBwdArgs a = args[f];
  const int gx = (a.W + kLTX - 1) / kLTX, gy = (a.H + kLTY - 1) / kLTY;
  const int l = (int)(b - prefix[f]);
  ssim_bwd_block(a, l % gx, (l / gx) % gy, l / (gx * gy));
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
  a.partial = nullptr;
  a.nb = 0;
  a.out3 = nullptr;
  hipLaunchKernelGGL(k_ssim_bwd, loss_grid(C, H, W), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// The two launches' arguments of gs_ssim_l1_forward_backward.
static int fwd_bwd_args(const float* pred, const float* target, int C, int H, int W, const float* window11,
                        float lambda_dssim, int clamp01, float* out3, const float* grad_out, void* ws,
                        size_t ws_bytes, float* dpred, LossArgs& f, LossBwdArgs& b) {
  if (!pred || !target || !window11 || !out3 || !grad_out || !ws || !dpred || C <= 0 || H <= 0 || W <= 0)
    return GS_E_ARG;
  if (ws_bytes < gs_loss_workspace_bytes(C, H, W)) return GS_E_WORKSPACE;
  f.C = C;
  f.H = H;
  f.W = W;
  f.lambda = lambda_dssim;
  f.clamp01 = clamp01;
  for (int k = 0; k < 11; ++k) f.win.g[k] = window11[k];
  f.x = pred;
  f.y = target;
  f.maps = (float*)ws;
  f.partial = (float*)((char*)ws + align_up(3 * sizeof(float) * (size_t)C * H * W, 256));
  const dim3 g = loss_grid(C, H, W);
  b.C = C;
  b.H = H;
  b.W = W;
  b.lambda = lambda_dssim;
  b.clamp01 = clamp01;
  b.inv_n = 1.0f / ((float)C * H * W);
  b.win = f.win;
  b.x = pred;
  b.y = target;
  b.maps = f.maps;
  b.grad_out = grad_out;
  b.dx = dpred;
  b.partial = f.partial;
  b.nb = (int)(g.x * g.y * g.z);
  b.out3 = out3;
  return GS_OK;
}

extern "C" int gs_ssim_l1_forward_backward(const float* pred, const float* target, int C, int H, int W,
                                           const float* window11, float lambda_dssim, int clamp01, float* out3,
                                           const float* grad_out, void* ws, size_t ws_bytes, float* dpred,
                                           void* stream) {
  LossArgs f;
  LossBwdArgs b;
  const int rc = fwd_bwd_args(pred, target, C, H, W, window11, lambda_dssim, clamp01, out3, grad_out, ws, ws_bytes,
                              dpred, f, b);
  if (rc != GS_OK) return rc;
  const dim3 g = loss_grid(C, H, W);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ssim_fwd, g, dim3(256), 0, s, f);
  hipLaunchKernelGGL(k_ssim_bwd, g, dim3(256), 0, s, b);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// Set layout: uint32 prefix[n + 1], then LossArgs[n], then LossBwdArgs[n],
// each at a 256-byte boundary.
static size_t multi_loss_fwd_offset(int n) { return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256); }
static size_t multi_loss_bwd_offset(int n) {
  return multi_loss_fwd_offset(n) + align_up(sizeof(LossArgs) * (size_t)n, 256);
}

extern "C" size_t gs_multi_loss_bytes(int n) {
  return n <= 0 ? 0 : multi_loss_bwd_offset(n) + sizeof(LossBwdArgs) * (size_t)n;
}

extern "C" int gs_multi_loss_build(int n, const float* const* pred, const float* const* target, const int* C,
                                   const int* H, const int* W, const float* window11, float lambda_dssim,
                                   int clamp01, float* const* out3, const float* grad_out, void* const* ws,
                                   const size_t* ws_bytes, float* const* dpred, void* host_set, size_t set_bytes,
                                   unsigned* blocks) {
  if (n <= 0 || !pred || !target || !C || !H || !W || !out3 || !ws || !ws_bytes || !dpred || !host_set || !blocks)
    return GS_E_ARG;
  if (set_bytes < gs_multi_loss_bytes(n)) return GS_E_WORKSPACE;
  uint32_t* prefix = (uint32_t*)host_set;
  LossArgs* fa = (LossArgs*)((char*)host_set + multi_loss_fwd_offset(n));
  LossBwdArgs* ba = (LossBwdArgs*)((char*)host_set + multi_loss_bwd_offset(n));
  unsigned long long total = 0;
  for (int i = 0; i < n; ++i) {
    const int rc = fwd_bwd_args(pred[i], target[i], C[i], H[i], W[i], window11, lambda_dssim, clamp01, out3[i],
                                grad_out, ws[i], ws_bytes[i], dpred[i], fa[i], ba[i]);
    if (rc != GS_OK) return rc;
    const dim3 g = loss_grid(C[i], H[i], W[i]);
    prefix[i] = (uint32_t)total;
    total += (unsigned long long)g.x * g.y * g.z;
    if (total > 0x7fffffffULL) return GS_E_ARG;
  }
  prefix[n] = (uint32_t)total;
  *blocks = (unsigned)total;
  return GS_OK;
}

extern "C" int gs_multi_loss(int n, const void* dev_set, unsigned blocks, void* stream) {
  if (n <= 0 || !dev_set || blocks == 0) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t* prefix = (const uint32_t*)dev_set;
  const LossArgs* fa = (const LossArgs*)((const char*)dev_set + multi_loss_fwd_offset(n));
  const LossBwdArgs* ba = (const LossBwdArgs*)((const char*)dev_set + multi_loss_bwd_offset(n));
  hipLaunchKernelGGL(k_ssim_fwd_multi, dim3(blocks), dim3(256), 0, s, prefix, fa, n);
  hipLaunchKernelGGL(k_ssim_bwd_multi, dim3(blocks), dim3(256), 0, s, prefix, ba, n);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
