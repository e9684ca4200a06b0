// gs_densify.hip -- adaptive density control of the Gaussian set on the
// device: the per-step densification statistics, and clone / split / prune
// as one stable compaction of every parameter and its two Adam moments.
//
// The reference does this in torch ops (scene/gaussian_model_cf.py:523-719,
// the per-step statistics at trainer/cf3dgs_trainer.py:144-152): boolean
// masks, index_select of every parameter and moment, torch.cat of the new
// points and zero moments, once for the clones, once for the split copies and
// once per prune -- each a full read and write of ~59 floats x 3 per Gaussian.
// Here:
//   k_dens_stats  per Gaussian with radius > 0: grad_accum += |dL/dmean2D.xy|,
//                 denom += 1, max_radii2D = max(max_radii2D, radius)
//   k_dens_mark   per Gaussian: the clone / split selection and the prune
//                 decision of the original, of its clone and of its two split
//                 copies; per-1024-block counts of the four output segments
//   k_dens_scan   one workgroup: exclusive prefix of the block counts and the
//                 segment totals (the host reads the totals to size the new
//                 arrays, the reference's own synchronisation point)
//   k_dens_emit   per Gaussian: its surviving rows written straight to their
//                 final positions in the new arrays (one read of every source
//                 row, one write of every surviving row)
//
// Final order, as the reference's sequence of cat and boolean indexing
// leaves it: [originals kept] [clones kept] [split copy 1 kept]
// [split copy 2 kept], each segment in original index order.  A split
// original is removed; clones and split copies get zero moments.  The
// arithmetic that decides a mask is torch's: expf for get_scaling,
// 1 / (1 + expf(-x)) for get_opacity, the split scale log(exp(s) * 0.625)
// (torch divides by the host scalar 1.6 as a multiply by its reciprocal,
// 0.625 exactly), and |g| = sqrt(g * g) for the norm of the 1-vector grads.
#include <cmath>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

constexpr int kDensBlock = 1024;
// flag bits per Gaussian
constexpr uint8_t kKeepOrig = 1, kKeepClone = 2, kSplitSel = 4, kKeepSplit = 8;

struct DensLayout {
  size_t counts, excl, total;
};

static DensLayout dens_layout(int P) {
  DensLayout L;
  const size_t nb = (size_t)(P + kDensBlock - 1) / kDensBlock;
  size_t o = 0;
  L.counts = o;  // [nb][4]
  o = align_up(o + 16 * (nb > 0 ? nb : 1), 256);
  L.excl = o;  // [nb][4]
  o = align_up(o + 16 * (nb > 0 ? nb : 1), 256);
  L.total = o;
  return L;
}

__global__ __launch_bounds__(256) void k_dens_stats(int P, const float* __restrict__ grad2d,
                                                    const int* __restrict__ radii, float* __restrict__ accum,
                                                    float* __restrict__ denom, float* __restrict__ max_radii) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const int r = radii[i];
  if (r <= 0) return;
  const float gx = grad2d[3 * i], gy = grad2d[3 * i + 1];
  accum[i] = accum[i] + sqrtf(gx * gx + gy * gy);
  denom[i] = denom[i] + 1.0f;
  max_radii[i] = fmaxf(max_radii[i], (float)r);
}

struct DensMark {
  int P;
  const float* accum;
  const float* denom;
  const float* scaling;  // raw [P,3]
  const float* opacity;  // raw [P]
  const float* max_radii;
  float grad_thr, dense_ext, min_opacity, big_ext, max_screen;
  int densify;  // clone + split selection on
  int prune;    // prune pass on
  uint8_t* flags;
  uint32_t* counts;  // [nb][4]
};

__device__ __forceinline__ float max_scale(float s0, float s1, float s2) {
  return fmaxf(fmaxf(expf(s0), expf(s1)), expf(s2));
}

__global__ __launch_bounds__(256) void k_dens_mark(DensMark a) {
  __shared__ uint32_t s_c[4][4];
  const int i = blockIdx.x * kDensBlock + threadIdx.x * 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t c[4] = {0u, 0u, 0u, 0u};
  uint32_t fl4 = 0u;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int k = i + e;
    uint8_t f = 0;
    if (k < a.P) {
      const float s0 = a.scaling[3 * k], s1 = a.scaling[3 * k + 1], s2 = a.scaling[3 * k + 2];
      const float smax = max_scale(s0, s1, s2);
      const float op = 1.0f / (1.0f + expf(-a.opacity[k]));
      bool clone = false, split = false;
      if (a.densify) {
        float g = a.accum[k] / a.denom[k];
        if (isnan(g)) g = 0.0f;
        clone = sqrtf(g * g) >= a.grad_thr && smax <= a.dense_ext;
        split = g >= a.grad_thr && smax > a.dense_ext;
      }
      bool pr_o = false, pr_s = false;
      if (a.prune) {
        pr_o = op < a.min_opacity;
        if (a.max_screen > 0.0f) pr_o = pr_o || a.max_radii[k] > a.max_screen;
        if (a.big_ext > 0.0f) pr_o = pr_o || smax > a.big_ext;
        // a split copy: same opacity, scale log(exp(s) * 0.625); its radius
        // statistic is zero (the reference resets max_radii2D when it
        // appends points)
        pr_s = op < a.min_opacity;
        if (a.big_ext > 0.0f)
          pr_s = pr_s || max_scale(logf(expf(s0) * 0.625f), logf(expf(s1) * 0.625f), logf(expf(s2) * 0.625f)) >
                             a.big_ext;
      }
      if (!split && !pr_o) f |= kKeepOrig;
      if (clone && !pr_o) f |= kKeepClone;
      if (split) f |= kSplitSel;
      if (split && !pr_s) f |= kKeepSplit;
      c[0] += (f & kKeepOrig) ? 1u : 0u;
      c[1] += (f & kKeepClone) ? 1u : 0u;
      c[2] += (f & kKeepSplit) ? 1u : 0u;
      c[3] += (f & kSplitSel) ? 1u : 0u;
    }
    fl4 |= (uint32_t)f << (8 * e);
  }
  if (i < a.P) {
    if (i + 3 < a.P) {
      *reinterpret_cast<uint32_t*>(a.flags + i) = fl4;  // i is a multiple of 4 and flags 4-byte aligned
    } else {
      for (int e = 0; e < 4 && i + e < a.P; ++e) a.flags[i + e] = (uint8_t)(fl4 >> (8 * e));
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c[q] += __shfl_xor(c[q], off, 64);
  }
  if (lane == 0)
    for (int q = 0; q < 4; ++q) s_c[w][q] = c[q];
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    a.counts[4 * blockIdx.x + q] = (s_c[0][q] + s_c[1][q]) + (s_c[2][q] + s_c[3][q]);
  }
}

// One workgroup: exclusive prefix of the per-block counts of the four
// segments; totals[0..3] = (originals kept, clones kept, split copies kept
// (per copy), split-selected).
__global__ __launch_bounds__(1024) void k_dens_scan(const uint32_t* counts, uint32_t* excl, int nb, int* totals) {
  __shared__ uint32_t s_w[16][4];
  __shared__ uint32_t s_carry[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t < 4) s_carry[t] = 0u;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int b = b0 + t;
    uint32_t v[4], x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] = b < nb ? counts[4 * b + q] : 0u;
      x[q] = v[q];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x[q], off, 64);
        if (lane >= off) x[q] += y;
      }
      if (lane == 63) s_w[w][q] = x[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t before = s_carry[q];
      for (int k = 0; k < w; ++k) before += s_w[k][q];
      if (b < nb) excl[4 * b + q] = before + x[q] - v[q];
    }
    __syncthreads();
    if (t < 4) {
      uint32_t sum = 0;
      for (int k = 0; k < 16; ++k) sum += s_w[k][t];
      s_carry[t] += sum;
    }
    __syncthreads();
  }
  if (t < 4) totals[t] = (int)s_carry[t];
}

struct DensEmit {
  int P, M;
  const uint8_t* flags;
  const uint32_t* excl;  // [nb][4]
  const int* totals;     // [4] on the device
  const float* src[15];  // xyz, features, opacity, scaling, rotation; then their exp_avg, then exp_avg_sq
  float* dst[15];
  const float* samples;  // [2 * n_split, 3] N(0, std) draws of the split copies (copy 1 rows, then copy 2)
};

__device__ __forceinline__ void copy_row(const float* s, float* d, int w) {
  for (int k = 0; k < w; ++k) d[k] = s[k];
}
__device__ __forceinline__ void zero_row(float* d, int w) {
  for (int k = 0; k < w; ++k) d[k] = 0.0f;
}

__global__ __launch_bounds__(256) void k_dens_emit(DensEmit a) {
  __shared__ uint32_t s_w[4][4];
  const int blk = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i0 = blk * kDensBlock + t * 4;
  // this thread's 4 flags and its exclusive position in each segment within
  // the block (thread-local counts, then wave and workgroup prefixes)
  uint8_t f[4];
  uint32_t c[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f[e] = i0 + e < a.P ? a.flags[i0 + e] : (uint8_t)0;
    c[0] += (f[e] & kKeepOrig) ? 1u : 0u;
    c[1] += (f[e] & kKeepClone) ? 1u : 0u;
    c[2] += (f[e] & kKeepSplit) ? 1u : 0u;
    c[3] += (f[e] & kSplitSel) ? 1u : 0u;
  }
  uint32_t x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = c[q];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x[q], off, 64);
      if (lane >= off) x[q] += y;
    }
    if (lane == 63) s_w[w][q] = x[q];
  }
  __syncthreads();
  uint32_t pos[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t before = a.excl[4 * blk + q];
    for (int k = 0; k < w; ++k) before += s_w[k][q];
    pos[q] = before + x[q] - c[q];
  }
  const uint32_t K_o = (uint32_t)a.totals[0], K_c = (uint32_t)a.totals[1], K_s = (uint32_t)a.totals[2];
  const uint32_t K_sel = (uint32_t)a.totals[3];
  const int width[5] = {3, 3 * a.M, 1, 3, 4};
  const bool moments = a.src[5] != nullptr;
#pragma unroll 1
  for (int e = 0; e < 4; ++e) {
    const int i = i0 + e;
    if (i >= a.P) break;
    const uint8_t fl = f[e];
    if (fl & kKeepOrig) {
      const size_t o = pos[0]++;
      for (int p = 0; p < (moments ? 15 : 5); ++p) {
        const int wd = width[p % 5];
        copy_row(a.src[p] + (size_t)i * wd, a.dst[p] + o * wd, wd);
      }
    }
    if (fl & kKeepClone) {
      const size_t o = K_o + pos[1]++;
      for (int p = 0; p < 5; ++p) copy_row(a.src[p] + (size_t)i * width[p], a.dst[p] + o * width[p], width[p]);
      if (moments)
        for (int p = 5; p < 15; ++p) zero_row(a.dst[p] + o * width[p % 5], width[p % 5]);
    }
    if (fl & kSplitSel) {
      const uint32_t r = pos[3]++;  // rank among the split-selected: its rows of samples
      if (fl & kKeepSplit) {
        const uint32_t slot = pos[2]++;
        // rotation of the raw quaternion (w, x, y, z), normalised
        const float* q = a.src[4] + (size_t)i * 4;
        const float qn = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        const float qr = q[0] / qn, qx = q[1] / qn, qy = q[2] / qn, qz = q[3] / qn;
        const float R[9] = {1.f - 2.f * (qy * qy + qz * qz), 2.f * (qx * qy - qr * qz), 2.f * (qx * qz + qr * qy),
                            2.f * (qx * qy + qr * qz), 1.f - 2.f * (qx * qx + qz * qz), 2.f * (qy * qz - qr * qx),
                            2.f * (qx * qz - qr * qy), 2.f * (qy * qz + qr * qx), 1.f - 2.f * (qx * qx + qy * qy)};
        const float* xyz = a.src[0] + (size_t)i * 3;
        const float* sc = a.src[3] + (size_t)i * 3;
        float ns[3];
        for (int k = 0; k < 3; ++k) ns[k] = logf(expf(sc[k]) * 0.625f);
        for (int copy = 0; copy < 2; ++copy) {
          const size_t o = (size_t)K_o + K_c + (size_t)copy * K_s + slot;
          const float* smp = a.samples + ((size_t)copy * K_sel + r) * 3;
          float* dx = a.dst[0] + o * 3;
          for (int k = 0; k < 3; ++k)
            dx[k] = ((R[3 * k] * smp[0] + R[3 * k + 1] * smp[1]) + R[3 * k + 2] * smp[2]) + xyz[k];
          for (int p = 1; p < 5; ++p)
            if (p != 3) copy_row(a.src[p] + (size_t)i * width[p], a.dst[p] + o * width[p], width[p]);
          float* dsc = a.dst[3] + o * 3;
          for (int k = 0; k < 3; ++k) dsc[k] = ns[k];
          if (moments)
            for (int p = 5; p < 15; ++p) zero_row(a.dst[p] + o * width[p % 5], width[p % 5]);
        }
      }
    }
  }
}

}  // namespace gs

using namespace gs;

extern "C" int gs_densify_stats(int P, const float* grad_means2D, const int* radii, float* grad_accum, float* denom,
                                float* max_radii2D, void* stream) {
  if (P < 0 || (P > 0 && (!grad_means2D || !radii || !grad_accum || !denom || !max_radii2D))) return GS_E_ARG;
  if (P == 0) return GS_OK;
  hipLaunchKernelGGL(k_dens_stats, dim3((P + 255) / 256), dim3(256), 0, (hipStream_t)stream, P, grad_means2D, radii,
                     grad_accum, denom, max_radii2D);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" size_t gs_densify_workspace_bytes(int P) { return dens_layout(P < 0 ? 0 : P).total; }

extern "C" int gs_densify_mark(int P, const float* grad_accum, const float* denom, const float* scaling,
                               const float* opacity, const float* max_radii2D, float grad_threshold,
                               float dense_extent, float min_opacity, float big_extent, float max_screen,
                               int densify, int prune, uint8_t* flags, int* totals, void* ws, size_t ws_bytes,
                               void* stream) {
  if (P < 0 || !totals || (P > 0 && (!scaling || !opacity || !flags || !ws))) return GS_E_ARG;
  if (P > 0 && densify && (!grad_accum || !denom)) return GS_E_ARG;
  if (P > 0 && prune && max_screen > 0.0f && !max_radii2D) return GS_E_ARG;
  if (((uintptr_t)flags & 3) || P > 0x3fffffff) return GS_E_ARG;
  const DensLayout L = dens_layout(P);
  if (ws_bytes < L.total) return GS_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int nb = (P + kDensBlock - 1) / kDensBlock;
  char* b = (char*)ws;
  DensMark m;
  m.P = P;
  m.accum = grad_accum;
  m.denom = denom;
  m.scaling = scaling;
  m.opacity = opacity;
  m.max_radii = max_radii2D;
  m.grad_thr = grad_threshold;
  m.dense_ext = dense_extent;
  m.min_opacity = min_opacity;
  m.big_ext = big_extent;
  m.max_screen = max_screen;
  m.densify = densify != 0;
  m.prune = prune != 0;
  m.flags = flags;
  m.counts = (uint32_t*)(b + L.counts);
  if (nb > 0) hipLaunchKernelGGL(k_dens_mark, dim3(nb), dim3(256), 0, s, m);
  hipLaunchKernelGGL(k_dens_scan, dim3(1), dim3(1024), 0, s, (const uint32_t*)(b + L.counts),
                     (uint32_t*)(b + L.excl), nb, totals);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_densify_emit(int P, int M, const float* const* src, float* const* dst, const float* split_samples,
                               const uint8_t* flags, const int* totals, const void* ws, size_t ws_bytes,
                               void* stream) {
  if (P < 0 || M < 1 || M > 16 || !src || !dst || !totals) return GS_E_ARG;
  if (P == 0) return GS_OK;
  if (!flags || !ws || ((uintptr_t)flags & 3)) return GS_E_ARG;
  const DensLayout L = dens_layout(P);
  if (ws_bytes < L.total) return GS_E_WORKSPACE;
  DensEmit e;
  e.P = P;
  e.M = M;
  e.flags = flags;
  e.excl = (const uint32_t*)((const char*)ws + L.excl);
  e.totals = totals;
  const bool moments = src[5] != nullptr;
  for (int k = 0; k < 15; ++k) {
    if (k < 5 || moments) {
      if (!src[k] || !dst[k]) return GS_E_ARG;
    } else if (src[k]) {
      return GS_E_ARG;  // the moments come all or none
    }
    e.src[k] = src[k];
    e.dst[k] = dst[k];
  }
  e.samples = split_samples;
  const int nb = (P + kDensBlock - 1) / kDensBlock;
  hipLaunchKernelGGL(k_dens_emit, dim3(nb), dim3(256), 0, (hipStream_t)stream, e);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
