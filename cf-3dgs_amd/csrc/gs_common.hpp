// gs_common.hpp -- device-side building blocks shared by the rasteriser kernels.
//
// Rounding contract: every kernel is compiled with -ffp-contract=off, keeps
// HIP's correctly rounded fp32 division and square root, and spells out each
// fused multiply-add it wants.  Together with the exponential below this
// makes radii, tile rects, keys, ranges, contributor counts and the forward
// image bit-identical to the fp32 CPU restatement kept in oracle/.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gs_raster.h"

namespace gs {

constexpr int kBX = 16;
constexpr int kBY = 16;
constexpr int kBlock = kBX * kBY;  // one 16x16 tile per 256-thread workgroup
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Exponential of the blend: Cody & Waite reduction (Cody, Waite, "Software
// Manual for the Elementary Functions", 1980) + degree-7 Taylor polynomial,
// all steps explicit fma / rint / ldexp, which gfx950 and the host round
// identically.  Domain used by the rasteriser: x <= 0.
__device__ __forceinline__ float spec_expf(float x) {
  x = fmaxf(x, -80.0f);
  const float k = rintf(x * 1.44269502f);
  float r = fmaf(k, -0.693115234375f, x);
  r = fmaf(k, -3.19461833e-05f, r);
  float p = 1.98412701e-04f;
  p = fmaf(p, r, 1.38888892e-03f);
  p = fmaf(p, r, 8.33333377e-03f);
  p = fmaf(p, r, 4.16666679e-02f);
  p = fmaf(p, r, 1.66666672e-01f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// -0.5 (a dx^2 + c dy^2) - b dx dy with the oracle's fma placement.
__device__ __forceinline__ float blend_power(float a, float b, float c, float dx, float dy) {
  const float q = fmaf(a * dx, dx, (c * dy) * dy);
  return fmaf(-0.5f, q, -((b * dx) * dy));
}

// Exponent below which o * exp(power) < 1/255 for certain: ln(1/(255 o)) less
// a 1e-3 margin that covers the rounding of logf, of spec_expf (a few ulp)
// and of the product.  A Gaussian skipped by `power < threshold` is one the
// exact test (alpha < 1/255) skips too, so results are unchanged; the test
// only spares the exponential.  o = 0 gives +inf (always skipped, as the
// exact test does); o < 0 or NaN give NaN (never skipped here, left to the
// exact test).
__device__ __forceinline__ float skip_threshold(float o) {
  return logf(1.0f / (255.0f * o)) - 1e-3f;
}

// Saturating float -> int (NaN -> 0).
__device__ __forceinline__ int sat_i(float f) {
  if (f != f) return 0;
  if (f >= 2147483647.0f) return 2147483647;
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  return (int)f;
}

// ((v + 1) * S - 1) * 0.5 evaluated in double, rounded once to float.
__device__ __forceinline__ float ndc2pix(float v, int S) {
  return (float)((((double)v + 1.0) * (double)S - 1.0) * 0.5);
}

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 xf43(const float* m, f3 p) {
  return {m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
          m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]};
}
__device__ __forceinline__ float xf44w(const float* m, f3 p) {
  return m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15];
}

// Row-major rotation of quaternion (r, x, y, z); utils/general_utils.py:88-96.
__device__ __forceinline__ void quat_to_R(float r, float x, float y, float z, float R[9]) {
  R[0] = 1.0f - 2.0f * (y * y + z * z);
  R[1] = 2.0f * (x * y - r * z);
  R[2] = 2.0f * (x * z + r * y);
  R[3] = 2.0f * (x * y + r * z);
  R[4] = 1.0f - 2.0f * (x * x + z * z);
  R[5] = 2.0f * (y * z - r * x);
  R[6] = 2.0f * (x * z - r * y);
  R[7] = 2.0f * (y * z + r * x);
  R[8] = 1.0f - 2.0f * (x * x + y * y);
}

// d(loss)/d(r,x,y,z) given d(loss)/dR (row-major) for the polynomial above.
__device__ __forceinline__ void quat_R_grad(float r, float x, float y, float z, const float G[9], float dq[4]) {
  dq[0] = 2.0f * (-z * G[1] + y * G[2] + z * G[3] - x * G[5] - y * G[6] + x * G[7]);
  dq[1] = 2.0f * (y * G[1] + z * G[2] + y * G[3] - 2.0f * x * G[4] - r * G[5] + z * G[6] + r * G[7] -
                  2.0f * x * G[8]);
  dq[2] = 2.0f * (-2.0f * y * G[0] + x * G[1] + r * G[2] + x * G[3] + z * G[5] - r * G[6] + z * G[7] -
                  2.0f * y * G[8]);
  dq[3] = 2.0f * (-2.0f * z * G[0] - r * G[1] + x * G[2] + r * G[3] - 2.0f * z * G[4] + y * G[5] +
                  x * G[6] + y * G[7]);
}

// Pose [tx, ty, tz, qx, qy, qz, qw] -> normalised quaternion + rotation.
struct Pose {
  float R[9];
  float t[3];
  float q[4];  // (w, x, y, z), normalised
  float inv_norm;
};
__device__ __forceinline__ void make_pose(const float* p7, Pose& P) {
  const float qx = p7[3], qy = p7[4], qz = p7[5], qw = p7[6];
  const float n2 = ((qx * qx + qy * qy) + qz * qz) + qw * qw;
  P.inv_norm = 1.0f / sqrtf(n2);
  P.q[0] = qw * P.inv_norm;
  P.q[1] = qx * P.inv_norm;
  P.q[2] = qy * P.inv_norm;
  P.q[3] = qz * P.inv_norm;
  quat_to_R(P.q[0], P.q[1], P.q[2], P.q[3], P.R);
  P.t[0] = p7[0];
  P.t[1] = p7[1];
  P.t[2] = p7[2];
}
__device__ __forceinline__ f3 pose_apply(const Pose& P, f3 p) {
  return {((P.R[0] * p.x + P.R[1] * p.y) + P.R[2] * p.z) + P.t[0],
          ((P.R[3] * p.x + P.R[4] * p.y) + P.R[5] * p.z) + P.t[1],
          ((P.R[6] * p.x + P.R[7] * p.y) + P.R[8] * p.z) + P.t[2]};
}

// Real SH basis, constants of utils/sh_utils.py:26-45, polynomial of :74-100.
#define GS_SH_C0 0.28209479177387814f
#define GS_SH_C1 0.4886025119029199f
#define GS_SH_C2_0 1.0925484305920792f
#define GS_SH_C2_1 -1.0925484305920792f
#define GS_SH_C2_2 0.31539156525252005f
#define GS_SH_C2_3 -1.0925484305920792f
#define GS_SH_C2_4 0.5462742152960396f
#define GS_SH_C3_0 -0.5900435899266435f
#define GS_SH_C3_1 2.890611442640554f
#define GS_SH_C3_2 -0.4570457994644658f
#define GS_SH_C3_3 0.3731763325901154f
#define GS_SH_C3_4 -0.4570457994644658f
#define GS_SH_C3_5 1.445305721320277f
#define GS_SH_C3_6 -0.5900435899266435f

__device__ __forceinline__ void sh_basis(int deg, float x, float y, float z, float b[16]) {
  b[0] = GS_SH_C0;
  if (deg < 1) return;
  b[1] = -GS_SH_C1 * y;
  b[2] = GS_SH_C1 * z;
  b[3] = -GS_SH_C1 * x;
  if (deg < 2) return;
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  b[4] = GS_SH_C2_0 * xy;
  b[5] = GS_SH_C2_1 * yz;
  b[6] = GS_SH_C2_2 * ((2.0f * zz - xx) - yy);
  b[7] = GS_SH_C2_3 * xz;
  b[8] = GS_SH_C2_4 * (xx - yy);
  if (deg < 3) return;
  b[9] = (GS_SH_C3_0 * y) * (3.0f * xx - yy);
  b[10] = (GS_SH_C3_1 * xy) * z;
  b[11] = (GS_SH_C3_2 * y) * ((4.0f * zz - xx) - yy);
  b[12] = (GS_SH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
  b[13] = (GS_SH_C3_4 * x) * ((4.0f * zz - xx) - yy);
  b[14] = (GS_SH_C3_5 * z) * (xx - yy);
  b[15] = (GS_SH_C3_6 * x) * (xx - 3.0f * yy);
}

// d(basis_k)/d(x,y,z).
__device__ __forceinline__ void sh_basis_grad(int deg, float x, float y, float z, float g[16][3]) {
#pragma unroll
  for (int k = 0; k < 16; ++k) g[k][0] = g[k][1] = g[k][2] = 0.0f;
  if (deg < 1) return;
  g[1][1] = -GS_SH_C1;
  g[2][2] = GS_SH_C1;
  g[3][0] = -GS_SH_C1;
  if (deg < 2) return;
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  g[4][0] = GS_SH_C2_0 * y;
  g[4][1] = GS_SH_C2_0 * x;
  g[5][1] = GS_SH_C2_1 * z;
  g[5][2] = GS_SH_C2_1 * y;
  g[6][0] = -2.0f * GS_SH_C2_2 * x;
  g[6][1] = -2.0f * GS_SH_C2_2 * y;
  g[6][2] = 4.0f * GS_SH_C2_2 * z;
  g[7][0] = GS_SH_C2_3 * z;
  g[7][2] = GS_SH_C2_3 * x;
  g[8][0] = 2.0f * GS_SH_C2_4 * x;
  g[8][1] = -2.0f * GS_SH_C2_4 * y;
  if (deg < 3) return;
  g[9][0] = GS_SH_C3_0 * 6.0f * xy;
  g[9][1] = GS_SH_C3_0 * (3.0f * xx - 3.0f * yy);
  g[10][0] = GS_SH_C3_1 * yz;
  g[10][1] = GS_SH_C3_1 * xz;
  g[10][2] = GS_SH_C3_1 * xy;
  g[11][0] = GS_SH_C3_2 * (-2.0f * xy);
  g[11][1] = GS_SH_C3_2 * (4.0f * zz - xx - 3.0f * yy);
  g[11][2] = GS_SH_C3_2 * (8.0f * yz);
  g[12][0] = GS_SH_C3_3 * (-6.0f * xz);
  g[12][1] = GS_SH_C3_3 * (-6.0f * yz);
  g[12][2] = GS_SH_C3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
  g[13][0] = GS_SH_C3_4 * (4.0f * zz - 3.0f * xx - yy);
  g[13][1] = GS_SH_C3_4 * (-2.0f * xy);
  g[13][2] = GS_SH_C3_4 * (8.0f * xz);
  g[14][0] = GS_SH_C3_5 * (2.0f * xz);
  g[14][1] = GS_SH_C3_5 * (-2.0f * yz);
  g[14][2] = GS_SH_C3_5 * (xx - yy);
  g[15][0] = GS_SH_C3_6 * (3.0f * xx - 3.0f * yy);
  g[15][1] = GS_SH_C3_6 * (-6.0f * xy);
}

// ---------------------------------------------------------------------------
// Wave64 helpers.
// A DPP lane permutation within 16-lane rows (quad_perm, row_mirror,
// row_half_mirror: every lane has a source lane, so no lane keeps an old
// value and none needs to be zero-filled first).
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

// Sum over all 64 lanes; every lane must be active.  Result is wave-uniform.
// Maximum of v over the 64 lanes of the wave (all lanes active).
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Tile rectangle [x0, x1) x [y0, y1) of a splat at (px, py) with radius r
// (getRect of the reference rasteriser: ((p + r) + 16 - 1) / 16 for the end).
__device__ __forceinline__ void tile_rect(float px, float py, int r, int gx, int gy, int& x0, int& y0, int& x1,
                                          int& y1) {
  const float rr = (float)r;
  x0 = min(gx, max(0, sat_i((px - rr) / (float)kBX)));
  y0 = min(gy, max(0, sat_i((py - rr) / (float)kBY)));
  x1 = min(gx, max(0, sat_i((((px + rr) + (float)kBX) - 1.0f) / (float)kBX)));
  y1 = min(gy, max(0, sat_i((((py + rr) + (float)kBY) - 1.0f) / (float)kBY)));
}

__device__ __forceinline__ float wave_sum(float v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
  return (r0 + r1) + (r2 + r3);
}

// Block copies between global memory and LDS for 256-thread workgroups,
// n16 16-byte vectors.  Every thread issues all of its (up to 12) loads
// before its first store, so a workgroup has its whole block in flight at
// once instead of one load-to-store round trip per vector (the SH rows of
// 256 Gaussians at M = 16 are exactly 12 vectors per thread).
// (The loads are unconditional, clamped to the last vector, so the staging
// array is fully written and stays in registers.)  dst and src must not overlap.
__device__ __forceinline__ void block_copy16(float4* __restrict__ dst, const float4* __restrict__ src, int n16) {
  constexpr int kMax = 12;
  if (n16 <= 0) return;
  float4 v[kMax];
#pragma unroll
  for (int r = 0; r < kMax; ++r) v[r] = src[min((int)threadIdx.x + r * 256, n16 - 1)];
  // stores clamped the same way: the duplicates write the last vector's own
  // value to it again, and no store waits inside a branch for its load
#pragma unroll
  for (int r = 0; r < kMax; ++r) dst[min((int)threadIdx.x + r * 256, n16 - 1)] = v[r];
  for (int k = (int)threadIdx.x + kMax * 256; k < n16; k += 256) dst[k] = src[k];
}

// Rows of `per4` 16-byte vectors, contiguous in global memory, into LDS rows
// of `prow` floats.  A thread that then works on its own row reads
// LDS at a stride of prow words across the wave: an odd prow puts the 64
// lanes on 64 different banks (a stride of 48 puts them on 4).  The row /
// column of vector k = threadIdx.x + 256 r advance by constant steps, so no
// division per vector.  Loads come first, clamped to the last vector, as in
// block_copy16.  256 threads.
__device__ __forceinline__ void rows_to_lds(float* __restrict__ dst, const float4* __restrict__ src, int nrows,
                                            int per4, int prow) {
  constexpr int kMax = 12;
  const int n16 = nrows * per4;
  if (n16 <= 0) return;
  float4 v[kMax];
#pragma unroll
  for (int r = 0; r < kMax; ++r) v[r] = src[min((int)threadIdx.x + r * 256, n16 - 1)];
  const int q = 256 / per4, rem = 256 - q * per4;
  const int lg = (n16 - 1) / per4, lj = (n16 - 1) - lg * per4;  // the last vector
  int k = (int)threadIdx.x, g = k / per4, j = k - g * per4;
#pragma unroll
  for (int r = 0; r < kMax; ++r) {
    const bool in = k < n16;
    float* d = dst + (in ? g : lg) * prow + 4 * (in ? j : lj);
    d[0] = v[r].x;
    d[1] = v[r].y;
    d[2] = v[r].z;
    d[3] = v[r].w;
    k += 256;
    g += q;
    j += rem;
    if (j >= per4) {
      j -= per4;
      ++g;
    }
  }
  for (; k < n16; k += 256) {
    const float4 x = src[k];
    float* d = dst + g * prow + 4 * j;
    d[0] = x.x;
    d[1] = x.y;
    d[2] = x.z;
    d[3] = x.w;
    g += q;
    j += rem;
    if (j >= per4) {
      j -= per4;
      ++g;
    }
  }
}
// Inclusive prefix sum over a workgroup of NT threads (a multiple of 64, at
// most 1024); `total` receives the workgroup sum.  Each wave scans by
// shuffles and one barrier combines the wave totals (a Hillis-Steele scan
// needs two barriers per step).  s_w (NT / 64 words of LDS) may be rewritten
// only after a later barrier.
template <int NT>
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t* s_w, uint32_t& total) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const uint32_t sw = s_w[i];
    before += i < w ? sw : 0u;
    tot += sw;
  }
  total = tot;
  return x + before;
}

// One Adam element update in torch.optim.Adam's order (shared by the Adam
// kernel and the backward's fused update, so both round identically):
//   m = m + (g - m) * (1 - b1);  v = v * b2 + (1 - b2) * g * g;
//   p = p - step * m / (sqrt(v) / sqrt(bc2) + eps),  step = lr / bc1.
// The moments' update and the parameter's decrement step * m / denom, which
// does not depend on the parameter (with contraction off, p - adam_delta(..)
// is adam_one's result bit for bit).
__device__ __forceinline__ float adam_delta(float g, float& m, float& v, float b1, float b2, float eps,
                                            float inv_sqrt_bc2, float step) {
  m = fmaf(g - m, 1.0f - b1, m);
  v = fmaf(v, b2, (1.0f - b2) * g * g);
  const float denom = sqrtf(v) * inv_sqrt_bc2 + eps;
  return step * (m / denom);
}

__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, float b1, float b2, float eps,
                                         float inv_sqrt_bc2, float step) {
  p = p - adam_delta(g, m, v, b1, b2, eps, inv_sqrt_bc2, step);
}

// 16-byte loads and stores that bypass cache retention (each byte of an
// optimiser pass is touched once).
typedef float nt4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt(const float* p) {
  const nt4 v = __builtin_nontemporal_load(reinterpret_cast<const nt4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt(float* p, const float4& v) {
  const nt4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<nt4*>(p));
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Pixel-independent form of the same test: an axis-aligned box around the
// splat centre outside which power < thr - 0.01 everywhere.  The set
// {power >= thr - 0.01} is the ellipse {Q <= L}, Q = a dx^2 + 2 b dx dy + c dy^2,
// L = -2 (thr - 0.01), whose bounding box has half-extents sqrt(L c / det) and
// sqrt(L a / det), det = a c - b^2.  The 0.01 in the exponent, the relative
// 1e-4 and the absolute 1e-3 pixel on the extents cover the rounding of the
// per-pixel power and of this computation by orders of magnitude, so a pixel
// outside the box is one the per-pixel test skips as well.  Returned as
// (x0, x1, y0, y1); thr NaN or a non-positive-definite conic give an infinite
// box (never skipped), thr >= 0.01 an empty one (power <= 0 < thr anyway).
__device__ __forceinline__ float4 splat_box(float x, float y, float a, float b, float c, float thr) {
  const float inf = __builtin_inff();
  if (thr != thr) return make_float4(-inf, inf, -inf, inf);
  const float L = -2.0f * (thr - 0.01f);
  if (!(L > 0.0f)) return make_float4(inf, -inf, inf, -inf);
  const float det = a * c - b * b;
  if (!(det > 0.0f)) return make_float4(-inf, inf, -inf, inf);
  const float ex = sqrtf(L * c / det) * 1.0001f + 1e-3f;
  const float ey = sqrtf(L * a / det) * 1.0001f + 1e-3f;
  if (!(ex == ex) || !(ey == ey)) return make_float4(-inf, inf, -inf, inf);
  return make_float4(x - ex, x + ex, y - ey, y + ey);
}

// Pixel (dx, dy) of thread t inside a 16 x 16 tile: wave w = t / 64 takes the
// 8 x 8 quadrant (w & 1, w >> 1), lane l the pixel (l & 7, l >> 3) in it.  A
// small splat covers more lanes of an 8 x 8 block than of a 16 x 4 strip.
__device__ __forceinline__ void tile_pixel(int t, int& dx, int& dy) {
  const int w = t >> 6, l = t & 63;
  dx = ((w & 1) << 3) + (l & 7);
  dy = ((w >> 1) << 3) + (l >> 3);
}

// Bit w set when the box meets the 8 x 8 pixels of quadrant w of the tile whose
// first pixel is (x0, y0).
__device__ __forceinline__ uint32_t wave_quads_hit(float4 bx, float x0, float y0) {
  uint32_t m = 0u;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const float qx = x0 + (float)((w & 1) << 3), qy = y0 + (float)((w >> 1) << 3);
    if (!(bx.y < qx || bx.x > qx + 7.0f || bx.w < qy || bx.z > qy + 7.0f)) m |= 1u << w;
  }
  return m;
}

// Quadrants of the tile at (x0, y0) that a splat at (x, y) with conic (a, b, c)
// and pre-test threshold thr can reach: the box test of wave_quads_hit, then,
// per quadrant, the exact minimum over its pixel rectangle of
// Q = a dx^2 + 2 b dx dy + c dy^2 (dx = x - px, dy = y - py): 0 when the centre
// lies inside, else the least of the four edge minima (a 1-D quadratic each,
// its vertex clamped to the edge).  A quadrant whose minimum exceeds
// L = -2 (thr - 0.01), padded by 1e-3 relative + 1e-3, has power below the
// pre-test threshold at every pixel and is dropped.
__device__ __forceinline__ float edge_min_q(float a, float b, float c, float u, float vlo, float vhi) {
  // min over v in [vlo, vhi] of a u^2 + 2 b u v + c v^2 (c > 0)
  const float v = fminf(vhi, fmaxf(vlo, -(b * u) / c));
  return a * u * u + 2.0f * b * u * v + c * v * v;
}
__device__ __forceinline__ uint32_t splat_quads(float x, float y, float a, float b, float c, float thr, float x0,
                                                float y0) {
  uint32_t m = wave_quads_hit(splat_box(x, y, a, b, c, thr), x0, y0);
  if (m == 0u || thr != thr) return m;
  const float L = -2.0f * (thr - 0.01f);
  if (!(a > 0.0f) || !(c > 0.0f) || !(a * c - b * b > 0.0f)) return m;
  const float lim = L * 1.001f + 1e-3f;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (!((m >> w) & 1u)) continue;
    const float qx = x0 + (float)((w & 1) << 3), qy = y0 + (float)((w >> 1) << 3);
    const float dxlo = x - (qx + 7.0f), dxhi = x - qx;
    const float dylo = y - (qy + 7.0f), dyhi = y - qy;
    if (dxlo <= 0.0f && dxhi >= 0.0f && dylo <= 0.0f && dyhi >= 0.0f) continue;  // centre inside
    float q = edge_min_q(a, b, c, dxlo, dylo, dyhi);
    q = fminf(q, edge_min_q(a, b, c, dxhi, dylo, dyhi));
    q = fminf(q, edge_min_q(c, b, a, dylo, dxlo, dxhi));
    q = fminf(q, edge_min_q(c, b, a, dyhi, dxlo, dxhi));
    if (q > lim) m &= ~(1u << w);
  }
  return m;
}

// The wave's own list of the staged splats that meet its pixels: positions j
// < limit (ascending) with bit `wv` set in s_mask[j]; returns the count
// (wave-uniform).  Needs all 64 lanes active.
template <int NC = 4>  // staged entries: NC * 64
__device__ __forceinline__ int compact_wave_list(const uint32_t* s_mask, uint8_t* s_list, int wv, int lane,
                                                 int limit = NC * 64) {
  int n = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int idx = c * 64 + lane;
    const bool hit = idx < limit && ((s_mask[idx] >> wv) & 1u);
    const uint64_t bal = __ballot(hit);
    if (hit) s_list[n + __popcll(bal & lanemask_lt())] = (uint8_t)idx;
    n += __popcll(bal);
  }
  return n;
}

// Block id -> logical id so that each of the 8 XCDs (dispatch is round-robin
// over XCDs) receives a contiguous run of logical work items: neighbouring
// tiles share Gaussians, so this keeps their gathers in one L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t full = nb & ~7u;
  if (b >= full) return b;
  const uint32_t per = full >> 3;
  return (b & 7u) * per + (b >> 3);
}

}  // namespace gs
