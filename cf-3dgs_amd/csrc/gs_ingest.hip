// gs_ingest.hip -- frame ingest: a depth map and its image become the local
// Gaussians of a new frame, without leaving the device.
//
// The reference lifts the depth with kornia's depth_to_3d (float32, on the
// device), copies the cloud to the host, builds an open3d PointCloud,
// voxel-downsamples it (voxel 0.01) and uploads it again for create_from_pcd
// (trainer/trainer.py:389-423 and :476-512; scene/gaussian_model_cf.py:178-212).
//
//   k_lift          pixel (u, v) with depth d -> ((u - cx) / fx * d,
//                   (v - cy) / fy * d, d) in float32, the operation order of
//                   kornia's unproject_points; optionally the colours as
//                   float32 (an 8-bit value v becomes v / 255 in double, as
//                   numpy's np.asarray(image) / 255.0 does, then float32)
//   k_bounds_part   per-workgroup min / max of the cloud
//   k_bounds_final  one workgroup: the cloud's min / max (the host reads them
//                   back to size the voxel key, as open3d's GetMinBound /
//                   GetMaxBound feed its grid)
//   k_vox_keys      voxel index floor((p - vmin) / s) per axis in double,
//                   vmin = min - s / 2 (open3d's VoxelDownSample), packed as
//                   ix << (by + bz) | iy << bz | iz
//   radix sort      (key, point index), stable: each voxel's points stay in
//                   index order
//   k_vox_count     voxel heads (first sorted point of each voxel) per
//                   workgroup of 1024 sorted points
//   k_vox_scan      one workgroup: exclusive prefix of the head counts, and
//                   the number of voxels
//   k_vox_emit      each head sums its voxel's points and colours in double,
//                   in point-index order (open3d's AccumulatedPoint::AddPoint),
//                   and writes sum / count as float32; voxels come out in
//                   ascending key order (open3d's order is its hash map's)
//   k_from_pcd      create_from_pcd: xyz, SH [P, K, 3] with DC = RGB2SH and
//                   the rest 0, scale log(sqrt(clamp_min(dist2, 1e-7))) x 3,
//                   rotation (1, 0, 0, 0), opacity inverse_sigmoid(0.1)
#include <float.h>
#include <limits.h>
#include <math.h>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

constexpr int kBoundsParts = 1024;  // partial min/max records at most
constexpr int kVoxItems = 4;        // sorted keys per thread in count / emit
constexpr int kVoxTile = 256 * kVoxItems;

struct ColorSrc {
  const void* p;
  int u8;          // 1: uint8 values v -> v / 255; 0: float32 values
  long long pix;   // element stride between pixels
  long long chan;  // element stride between channels
};

__device__ __forceinline__ double color_at(const ColorSrc& c, long long i, int ch) {
  const long long o = i * c.pix + (long long)ch * c.chan;
  if (c.u8) return (double)((const uint8_t*)c.p)[o] / 255.0;
  return (double)((const float*)c.p)[o];
}

__global__ __launch_bounds__(256) void k_lift(int W, int H, const float* __restrict__ depth, float fx, float fy,
                                              float cx, float cy, float* __restrict__ pts, ColorSrc col,
                                              float* __restrict__ colors) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)W * H) return;
  const int v = (int)(i / W), u = (int)(i - (long long)v * W);
  const float d = depth[i];
  // kornia: x = (u - cx) / fx, y = (v - cy) / fy, then [x, y, 1] * depth
  const float x = ((float)u - cx) / fx;
  const float y = ((float)v - cy) / fy;
  pts[3 * i + 0] = x * d;
  pts[3 * i + 1] = y * d;
  pts[3 * i + 2] = d;
  if (colors) {
#pragma unroll
    for (int c = 0; c < 3; ++c) colors[3 * i + c] = (float)color_at(col, i, c);
  }
}

// Records of 8 floats: min xyz, max xyz, a non-finite flag, padding.
__global__ __launch_bounds__(256) void k_bounds_part(const float* __restrict__ pts, long long N, float* part) {
  __shared__ float s[7][4];
  float r[7] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, 0.0f};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)gridDim.x * 256) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float p = pts[3 * i + c];
      r[c] = fminf(r[c], p);
      r[3 + c] = fmaxf(r[3 + c], p);
      r[6] = isfinite(p) ? r[6] : 1.0f;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[c] = fminf(r[c], __shfl_xor(r[c], off, 64));
      r[3 + c] = fmaxf(r[3 + c], __shfl_xor(r[3 + c], off, 64));
    }
    r[6] = fmaxf(r[6], __shfl_xor(r[6], off, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int c = 0; c < 7; ++c) s[c][w] = r[c];
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int c = threadIdx.x;
    float v = 0.0f;
    if (c < 7) {
      v = s[c][0];
      for (int k = 1; k < 4; ++k) v = c < 3 ? fminf(v, s[c][k]) : fmaxf(v, s[c][k]);
    }
    part[8 * blockIdx.x + c] = v;
  }
}

// bounds6 = min xyz, max xyz; all NaN when any coordinate is not finite.
__global__ __launch_bounds__(256) void k_bounds_final(const float* part, int nparts, float* bounds6) {
  __shared__ float s[7][4];
  float r[7] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, 0.0f};
  for (int k = threadIdx.x; k < nparts; k += 256) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[c] = fminf(r[c], part[8 * k + c]);
      r[3 + c] = fmaxf(r[3 + c], part[8 * k + 3 + c]);
    }
    r[6] = fmaxf(r[6], part[8 * k + 6]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[c] = fminf(r[c], __shfl_xor(r[c], off, 64));
      r[3 + c] = fmaxf(r[3 + c], __shfl_xor(r[3 + c], off, 64));
    }
    r[6] = fmaxf(r[6], __shfl_xor(r[6], off, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int c = 0; c < 7; ++c) s[c][w] = r[c];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    float v = s[c][0];
    for (int k = 1; k < 4; ++k) v = c < 3 ? fminf(v, s[c][k]) : fmaxf(v, s[c][k]);
    const bool bad = fmaxf(fmaxf(s[6][0], s[6][1]), fmaxf(s[6][2], s[6][3])) > 0.0f;
    bounds6[c] = bad ? __builtin_nanf("") : v;
  }
}

struct VoxGrid {
  double vx, vy, vz, size;
  int by, bz;
};

template <typename K>
__global__ __launch_bounds__(256) void k_vox_keys(const float* __restrict__ pts, long long N, VoxGrid g,
                                                  K* __restrict__ keys, uint32_t* __restrict__ vals,
                                                  uint32_t* n_dev) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i == 0) n_dev[0] = (uint32_t)N;
  if (i >= N) return;
  // open3d: ref = (p - voxel_min_bound) / voxel_size; index = int(floor(ref))
  const double rx = ((double)pts[3 * i + 0] - g.vx) / g.size;
  const double ry = ((double)pts[3 * i + 1] - g.vy) / g.size;
  const double rz = ((double)pts[3 * i + 2] - g.vz) / g.size;
  const uint64_t ix = (uint64_t)(int)floor(rx), iy = (uint64_t)(int)floor(ry), iz = (uint64_t)(int)floor(rz);
  keys[i] = (K)((ix << (g.by + g.bz)) | (iy << g.bz) | iz);
  vals[i] = (uint32_t)i;
}

template <typename K>
__device__ __forceinline__ bool is_head(const K* keys, long long i) {
  return i == 0 || keys[i] != keys[i - 1];
}

template <typename K>
__global__ __launch_bounds__(256) void k_vox_count(const K* __restrict__ keys, long long N, uint32_t* blk) {
  __shared__ uint32_t s_w[4];
  const long long i0 = (long long)blockIdx.x * kVoxTile + (long long)threadIdx.x * kVoxItems;
  uint32_t c = 0;
#pragma unroll
  for (int r = 0; r < kVoxItems; ++r) {
    const long long i = i0 + r;
    c += (i < N && is_head(keys, i)) ? 1u : 0u;
  }
  uint32_t tot;
  block_scan_incl<256>(c, s_w, tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// One workgroup: blk[0..nb) becomes its exclusive prefix; the total goes to
// *n_out.
__global__ __launch_bounds__(1024) void k_vox_scan(uint32_t* blk, int nb, uint32_t* n_out) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const int t = threadIdx.x;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int b = b0 + t;
    const uint32_t v = b < nb ? blk[b] : 0u;
    uint32_t tot;
    const uint32_t incl = block_scan_incl<1024>(v, s_w, tot);
    const uint32_t carry = s_carry;
    if (b < nb) blk[b] = carry + incl - v;
    __syncthreads();
    if (t == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (t == 0) n_out[0] = s_carry;
}

template <typename K>
__global__ __launch_bounds__(256) void k_vox_emit(const K* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                  long long N, const uint32_t* __restrict__ blk,
                                                  const float* __restrict__ pts, ColorSrc col,
                                                  float* __restrict__ out_pts, float* __restrict__ out_col) {
  __shared__ uint32_t s_w[4];
  const long long i0 = (long long)blockIdx.x * kVoxTile + (long long)threadIdx.x * kVoxItems;
  bool head[kVoxItems];
  uint32_t c = 0;
#pragma unroll
  for (int r = 0; r < kVoxItems; ++r) {
    const long long i = i0 + r;
    head[r] = i < N && is_head(keys, i);
    c += head[r] ? 1u : 0u;
  }
  uint32_t tot;
  uint32_t seg = blk[blockIdx.x] + block_scan_incl<256>(c, s_w, tot) - c;
#pragma unroll
  for (int r = 0; r < kVoxItems; ++r) {
    if (!head[r]) continue;
    const long long i = i0 + r;
    const K k = keys[i];
    double sp[3] = {0.0, 0.0, 0.0}, sc[3] = {0.0, 0.0, 0.0};
    long long n = 0;
    for (long long j = i; j < N && keys[j] == k; ++j, ++n) {
      const long long p = vals[j];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        sp[a] += (double)pts[3 * p + a];
        sc[a] += color_at(col, p, a);
      }
    }
    const double dn = (double)n;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      out_pts[3 * (long long)seg + a] = (float)(sp[a] / dn);
      out_col[3 * (long long)seg + a] = (float)(sc[a] / dn);
    }
    ++seg;
  }
}

// create_from_pcd (scene/gaussian_model_cf.py:178-212) on device-resident
// points: the arithmetic of the torch ops it runs on the GPU.  RGB2SH's
// division by the host scalar C0 is what torch's true division by a host
// scalar computes: a multiply by the reciprocal, taken in double and rounded
// to float32.  The opacity is log(x / (1 - x)) of x = 0.1f, evaluated here so
// it uses the device logf.
__global__ __launch_bounds__(256) void k_from_pcd(int P, int K, const float* __restrict__ pts,
                                                  const float* __restrict__ col, const float* __restrict__ dist2,
                                                  float inv_c0, float opac_x, float* __restrict__ xyz,
                                                  float* __restrict__ feat, float* __restrict__ scaling,
                                                  float* __restrict__ rot, float* __restrict__ opac) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) xyz[3 * i + a] = pts[3 * i + a];
  float* f = feat + (long long)i * K * 3;
#pragma unroll
  for (int a = 0; a < 3; ++a) f[a] = (col[3 * i + a] - 0.5f) * inv_c0;
  for (int k = 3; k < 3 * K; ++k) f[k] = 0.0f;
  const float d2 = dist2[i];
  const float dc = d2 < 1e-7f ? 1e-7f : d2;  // clamp_min (NaN passes through)
  const float s = logf(sqrtf(dc));
#pragma unroll
  for (int a = 0; a < 3; ++a) scaling[3 * i + a] = s;
  rot[4 * i + 0] = 1.0f;
  rot[4 * i + 1] = 0.0f;
  rot[4 * i + 2] = 0.0f;
  rot[4 * i + 3] = 0.0f;
  opac[i] = logf(opac_x / (1.0f - opac_x));
}

struct VoxLayout {
  size_t keys_a, keys_b, vals_a, vals_b, n_dev, blk, sort_tmp, total;
};

static VoxLayout vox_layout(long long N) {
  VoxLayout L;
  const long long n = N > 0 ? N : 1;
  const long long nb = (n + kVoxTile - 1) / kVoxTile;
  size_t o = 0;
  L.keys_a = o;
  o = align_up(o + 8 * (size_t)n, 256);
  L.keys_b = o;
  o = align_up(o + 8 * (size_t)n, 256);
  L.vals_a = o;
  o = align_up(o + 4 * (size_t)n, 256);
  L.vals_b = o;
  o = align_up(o + 4 * (size_t)n, 256);
  L.n_dev = o;
  o = align_up(o + 16, 256);
  L.blk = o;
  o = align_up(o + 4 * (size_t)(nb + 1), 256);
  L.sort_tmp = o;
  o = align_up(o + sort_tmp_bytes(n), 256);
  L.total = o;
  return L;
}

template <typename K>
static hipError_t vox_run(const float* pts, long long N, const VoxGrid& g, int end_bit, const ColorSrc& col,
                          float* out_pts, float* out_col, uint32_t* n_out, char* b, const VoxLayout& L,
                          hipStream_t s) {
  K* ka = (K*)(b + L.keys_a);
  K* kb = (K*)(b + L.keys_b);
  uint32_t* va = (uint32_t*)(b + L.vals_a);
  uint32_t* vb = (uint32_t*)(b + L.vals_b);
  uint32_t* n_dev = (uint32_t*)(b + L.n_dev);
  uint32_t* blk = (uint32_t*)(b + L.blk);
  const unsigned nb256 = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL((k_vox_keys<K>), dim3(nb256), dim3(256), 0, s, pts, N, g, ka, va, n_dev);
  int in_b = 0;
  hipError_t e;
  if constexpr (sizeof(K) == 8)
    e = sort_pairs_u64(ka, va, kb, vb, n_dev, N, end_bit, b + L.sort_tmp, s, &in_b);
  else
    e = sort_pairs_u32(ka, va, kb, vb, n_dev, N, end_bit, b + L.sort_tmp, s, &in_b);
  if (e != hipSuccess) return e;
  const K* keys = in_b ? kb : ka;
  const uint32_t* vals = in_b ? vb : va;
  const int nb = (int)((N + kVoxTile - 1) / kVoxTile);
  hipLaunchKernelGGL((k_vox_count<K>), dim3(nb), dim3(256), 0, s, keys, N, blk);
  hipLaunchKernelGGL(k_vox_scan, dim3(1), dim3(1024), 0, s, blk, nb, n_out);
  hipLaunchKernelGGL((k_vox_emit<K>), dim3(nb), dim3(256), 0, s, keys, vals, N, (const uint32_t*)blk, pts, col,
                     out_pts, out_col);
  return hipGetLastError();
}

}  // namespace gs

using namespace gs;

extern "C" int gs_depth_to_points(int W, int H, const float* depth, float fx, float fy, float cx, float cy,
                                  float* points, const void* colors_src, int colors_u8, long long color_pix_stride,
                                  long long color_chan_stride, float* colors, void* stream) {
  if (W < 0 || H < 0 || (long long)W * H > 0x7fffffffLL) return GS_E_ARG;
  if ((long long)W * H == 0) return GS_OK;
  if (!depth || !points || (colors && !colors_src)) return GS_E_ARG;
  if (!(fx != 0.0f) || !(fy != 0.0f)) return GS_E_ARG;
  const ColorSrc col{colors_src, colors_u8 ? 1 : 0, color_pix_stride, color_chan_stride};
  const long long n = (long long)W * H;
  hipLaunchKernelGGL(k_lift, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, W, H, depth, fx,
                     fy, cx, cy, points, col, colors);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" size_t gs_bounds_workspace_bytes(void) { return 8 * sizeof(float) * kBoundsParts; }

extern "C" int gs_points_bounds(const float* points, long long N, float* bounds6, void* ws, size_t ws_bytes,
                                void* stream) {
  if (N <= 0 || N > 0x7fffffffLL || !points || !bounds6 || !ws) return GS_E_ARG;
  if (ws_bytes < gs_bounds_workspace_bytes()) return GS_E_WORKSPACE;
  const long long nb = (N + 255) / 256;
  const int parts = (int)(nb < kBoundsParts ? nb : kBoundsParts);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bounds_part, dim3(parts), dim3(256), 0, s, points, N, (float*)ws);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, s, (const float*)ws, parts, bounds6);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" size_t gs_voxel_workspace_bytes(long long N) { return vox_layout(N).total; }

extern "C" int gs_voxel_down_sample(const float* points, long long N, const void* colors_src, int colors_u8,
                                    long long color_pix_stride, long long color_chan_stride, const double* voxel_min3,
                                    double voxel_size, int bits_x, int bits_y, int bits_z, float* out_points,
                                    float* out_colors, uint32_t* n_out, void* ws, size_t ws_bytes, void* stream) {
  if (N <= 0 || N > 0x7fffffffLL || !points || !colors_src || !voxel_min3 || !out_points || !out_colors || !n_out ||
      !ws)
    return GS_E_ARG;
  if (!(voxel_size > 0.0) || bits_x < 0 || bits_y < 0 || bits_z < 0 || bits_x > 31 || bits_y > 31 || bits_z > 31)
    return GS_E_ARG;
  const int end_bit = bits_x + bits_y + bits_z;
  if (end_bit > 64) return GS_E_ARG;
  const VoxLayout L = vox_layout(N);
  if (ws_bytes < L.total) return GS_E_WORKSPACE;
  const VoxGrid g{voxel_min3[0], voxel_min3[1], voxel_min3[2], voxel_size, bits_y, bits_z};
  const ColorSrc col{colors_src, colors_u8 ? 1 : 0, color_pix_stride, color_chan_stride};
  const int eb = end_bit > 0 ? end_bit : 1;
  hipError_t e = eb <= 32 ? vox_run<uint32_t>(points, N, g, eb, col, out_points, out_colors, n_out, (char*)ws, L,
                                              (hipStream_t)stream)
                          : vox_run<uint64_t>(points, N, g, eb, col, out_points, out_colors, n_out, (char*)ws, L,
                                              (hipStream_t)stream);
  return e == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_create_from_pcd(int P, int n_coef, const float* points, const float* colors, const float* dist2,
                                  float* xyz, float* features, float* scaling, float* rotation, float* opacity,
                                  void* stream) {
  if (P < 0 || n_coef < 1 || n_coef > 16) return GS_E_ARG;
  if (P == 0) return GS_OK;
  if (!points || !colors || !dist2 || !xyz || !features || !scaling || !rotation || !opacity) return GS_E_ARG;
  const float inv_c0 = (float)(1.0 / 0.28209479177387814);
  hipLaunchKernelGGL(k_from_pcd, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, (hipStream_t)stream, P, n_coef,
                     points, colors, dist2, inv_c0, 0.1f, xyz, features, scaling, rotation, opacity);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
