// gs_knn.hip -- exact mean squared distance to the 3 nearest other points
// (the quantity simple_knn's distCUDA2 returns; used for the initial scales
// at scene/gaussian_model_cf.py:194-196).
//
// Uniform grid built entirely on the device (the host never sees the
// bounding box):
//   k_knn_bbox      per-workgroup min/max
//   k_knn_setup     one thread: bbox, cell size for ~2 points per cell, dims
//   k_knn_count     cell of every point, per-cell counts
//   scan            exclusive prefix of the counts
//   k_knn_scatter   points grouped by cell (float4: x, y, z, index bits)
//   k_knn_query     one point per thread in cell order; rings of cells at
//                   Chebyshev distance r = 0, 1, ... until the 3rd best is
//                   closer than (r - 0.01) cells -- every unvisited point is
//                   then farther, so the result equals a brute-force search.
// The three best squared distances are kept sorted and summed smallest
// first, ((b0 + b1) + b2) / 3, so the value is independent of visit order.
#include <float.h>

#include <mutex>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

struct KnnLayout {
  size_t partial, params, cell_of, counts, starts, scan_tmp, sorted, total;
  long long ncell_max;
};

constexpr KnnLayout knn_layout(int P) {
  KnnLayout L{};
  const long long nb = (P + 255) / 256;
  L.ncell_max = (long long)P / 2 + 64;
  const long long nscan = (L.ncell_max + 1023) / 1024;
  size_t o = 0;
  L.partial = o;
  o = align_up(o + 24 * (size_t)(nb > 0 ? nb : 1), 256);
  L.params = o;
  o = align_up(o + 64, 256);
  L.cell_of = o;
  o = align_up(o + 4 * (size_t)P, 256);
  L.counts = o;
  o = align_up(o + 4 * (size_t)L.ncell_max, 256);
  L.starts = o;
  o = align_up(o + 4 * (size_t)L.ncell_max, 256);
  L.scan_tmp = o;
  o = align_up(o + 4 * (size_t)(nscan + 1), 256);
  L.sorted = o;
  o = align_up(o + 16 * (size_t)P, 256);
  L.total = o;
  return L;
}

struct KnnParams {
  float ox, oy, oz, cs, inv_cs;
  int dx, dy, dz;
  int ncell;
};

__global__ __launch_bounds__(256) void k_knn_bbox(const float* pts, int P, float* partial) {
  __shared__ float s[6][256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  if (i < P) {
    for (int k = 0; k < 3; ++k) {
      v[k] = pts[3 * i + k];
      v[3 + k] = v[k];
    }
  }
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
      for (int k = 3; k < 6; ++k) s[k][threadIdx.x] = fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) partial[6 * blockIdx.x + k] = s[k][0];
}

__global__ __launch_bounds__(256) void k_knn_setup(const float* partial, int nb, int P, long long ncell_max,
                                                   KnnParams* prm) {
  __shared__ float s[6][256];
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int b = threadIdx.x; b < nb; b += 256) {
    for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], partial[6 * b + k]);
    for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], partial[6 * b + k]);
  }
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
      for (int k = 3; k < 6; ++k) s[k][threadIdx.x] = fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  KnnParams q;
  q.ox = s[0][0];
  q.oy = s[1][0];
  q.oz = s[2][0];
  const float ex = s[3][0] - s[0][0], ey = s[4][0] - s[1][0], ez = s[5][0] - s[2][0];
  const float me = fmaxf(ex, fmaxf(ey, ez));
  if (!(me > 0.0f) || !isfinite(me) || P < 2) {
    q.cs = 1.0f;
    q.dx = q.dy = q.dz = 1;
  } else {
    const double e0 = fmax((double)ex, me * 1e-4), e1 = fmax((double)ey, me * 1e-4), e2 = fmax((double)ez, me * 1e-4);
    const double target = fmax(1.0, P * 0.5);
    double cs = cbrt(e0 * e1 * e2 / target);
    for (int it = 0; it < 200; ++it) {
      const long long nx = (long long)fmin(1024.0, fmax(1.0, ceil(e0 / cs)));
      const long long ny = (long long)fmin(1024.0, fmax(1.0, ceil(e1 / cs)));
      const long long nz = (long long)fmin(1024.0, fmax(1.0, ceil(e2 / cs)));
      q.dx = (int)nx;
      q.dy = (int)ny;
      q.dz = (int)nz;
      if (nx * ny * nz <= ncell_max) break;
      cs *= 1.25;
    }
    if ((long long)q.dx * q.dy * q.dz > ncell_max) {  // never expected; keeps every cell index in bounds
      q.dx = q.dy = q.dz = 1;
      cs = fmax(e0, fmax(e1, e2));
    }
    q.cs = (float)cs;
  }
  q.inv_cs = 1.0f / q.cs;
  q.ncell = q.dx * q.dy * q.dz;
  *prm = q;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ void cell_coords(const KnnParams& q, float x, float y, float z, int& cx, int& cy, int& cz) {
  cx = clampi(sat_i((x - q.ox) * q.inv_cs), 0, q.dx - 1);
  cy = clampi(sat_i((y - q.oy) * q.inv_cs), 0, q.dy - 1);
  cz = clampi(sat_i((z - q.oz) * q.inv_cs), 0, q.dz - 1);
}

__global__ __launch_bounds__(256) void k_knn_zero(uint32_t* p, long long n) {
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) p[k] = 0u;
}

__global__ __launch_bounds__(256) void k_knn_count(const float* pts, int P, const KnnParams* prm, uint32_t* cell_of,
                                                   uint32_t* counts) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const KnnParams q = *prm;
  int cx, cy, cz;
  cell_coords(q, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], cx, cy, cz);
  const uint32_t c = (uint32_t)((cz * q.dy + cy) * q.dx + cx);
  cell_of[i] = c;
  atomicAdd(&counts[c], 1u);
}

// 3-phase exclusive scan over n = prm->ncell entries (1024 per workgroup).
__global__ __launch_bounds__(1024) void k_knn_scan_local(const uint32_t* in, uint32_t* out, const KnnParams* prm,
                                                         uint32_t* block_sums) {
  __shared__ uint32_t s[1024];
  const int n = prm->ncell;
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if ((int)blockIdx.x * 1024 >= n) return;
  const uint32_t v = i < n ? in[i] : 0u;
  uint32_t tot;
  const uint32_t incl = block_scan_incl<1024>(v, s, tot);
  if (i < n) out[i] = incl - v;
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_knn_scan_top(uint32_t* block_sums, const KnnParams* prm) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t s_carry;
  const int nb = (prm->ncell + 1023) / 1024;
  const int t = threadIdx.x;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int b = b0 + t;
    const uint32_t v = b < nb ? block_sums[b] : 0u;
    uint32_t tot;
    const uint32_t incl = block_scan_incl<1024>(v, s, tot);
    const uint32_t carry = s_carry;
    if (b < nb) block_sums[b] = carry + incl - v;
    __syncthreads();
    if (t == 0) s_carry = carry + tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_knn_scan_add(uint32_t* out, const KnnParams* prm,
                                                       const uint32_t* block_sums) {
  const int n = prm->ncell;
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += block_sums[blockIdx.x];
}

__global__ __launch_bounds__(256) void k_knn_scatter(const float* pts, int P, const uint32_t* cell_of,
                                                     const uint32_t* starts, uint32_t* counts, float4* sorted) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const uint32_t c = cell_of[i];
  const uint32_t k = atomicSub(&counts[c], 1u) - 1u;
  sorted[starts[c] + k] = make_float4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], __int_as_float(i));
}

// Sorted insert into b0 <= b1 <= b2 (the three smallest so far), written as
// min / max selects: the nested-branch form is lowered to a stack array with
// a computed index, which puts the three values in scratch memory.  d >= b2
// (and NaN) leaves them unchanged; ties give the same three values either way.
__device__ __forceinline__ void best3_insert(float d, float& b0, float& b1, float& b2) {
  if (d < b2) {
    b2 = fminf(b2, fmaxf(b1, d));
    b1 = fminf(b1, fmaxf(b0, d));
    b0 = fminf(b0, d);
  }
}

__global__ __launch_bounds__(256) void k_knn_query(int P, const KnnParams* prm, const uint32_t* starts,
                                                   const uint32_t* cell_of_sorted_unused, const float4* sorted,
                                                   float* out) {
  (void)cell_of_sorted_unused;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= P) return;
  const KnnParams q = *prm;
  const float4 me = sorted[k];
  const int myidx = __float_as_int(me.w);
  int cx, cy, cz;
  cell_coords(q, me.x, me.y, me.z, cx, cy, cz);
  float b0 = FLT_MAX, b1 = FLT_MAX, b2 = FLT_MAX;
  const int rmax = max(q.dx, max(q.dy, q.dz));
  for (int r = 0; r <= rmax; ++r) {
    for (int dz = -r; dz <= r; ++dz) {
      const int z = cz + dz;
      if (z < 0 || z >= q.dz) continue;
      for (int dy = -r; dy <= r; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= q.dy) continue;
        const bool shell = (dz == -r || dz == r || dy == -r || dy == r);
        const int step = shell ? 1 : (r > 0 ? 2 * r : 1);
        for (int dx = -r; dx <= r; dx += step) {
          const int x = cx + dx;
          if (x < 0 || x >= q.dx) continue;
          const int c = (z * q.dy + y) * q.dx + x;
          const uint32_t s0 = starts[c];
          const uint32_t s1 = (c + 1 < q.ncell) ? starts[c + 1] : (uint32_t)P;
          for (uint32_t j = s0; j < s1; ++j) {
            const float4 o = sorted[j];
            if (__float_as_int(o.w) == myidx) continue;
            const float ddx = me.x - o.x, ddy = me.y - o.y, ddz = me.z - o.z;
            const float d = (ddx * ddx + ddy * ddy) + ddz * ddz;
            best3_insert(d, b0, b1, b2);
          }
        }
      }
    }
    const float reach = ((float)r - 0.01f) * q.cs;
    if (r > 0 && b2 < reach * reach) break;
  }
  out[myidx] = ((b0 + b1) + b2) / 3.0f;
}

// The launch sequence over a workspace laid out by knn_layout(P).
static int knn_launch(const float* points, int P, float* out, char* b, const KnnLayout& L, hipStream_t s) {
  float* partial = (float*)(b + L.partial);
  KnnParams* prm = (KnnParams*)(b + L.params);
  uint32_t* cell_of = (uint32_t*)(b + L.cell_of);
  uint32_t* counts = (uint32_t*)(b + L.counts);
  uint32_t* starts = (uint32_t*)(b + L.starts);
  uint32_t* scan_tmp = (uint32_t*)(b + L.scan_tmp);
  float4* sorted = (float4*)(b + L.sorted);
  const int nb = (P + 255) / 256;
  const int nscan = (int)((L.ncell_max + 1023) / 1024);
  timing::Scope ts(timing::kKnn, s);
  hipLaunchKernelGGL(k_knn_bbox, dim3(nb), dim3(256), 0, s, points, P, partial);
  hipLaunchKernelGGL(k_knn_setup, dim3(1), dim3(256), 0, s, partial, nb, P, L.ncell_max, prm);
  const long long zb = (L.ncell_max + 255) / 256;
  hipLaunchKernelGGL(k_knn_zero, dim3((unsigned)(zb < 2048 ? zb : 2048)), dim3(256), 0, s, counts, L.ncell_max);
  hipLaunchKernelGGL(k_knn_count, dim3(nb), dim3(256), 0, s, points, P, prm, cell_of, counts);
  hipLaunchKernelGGL(k_knn_scan_local, dim3(nscan), dim3(1024), 0, s, counts, starts, prm, scan_tmp);
  hipLaunchKernelGGL(k_knn_scan_top, dim3(1), dim3(1024), 0, s, scan_tmp, prm);
  hipLaunchKernelGGL(k_knn_scan_add, dim3(nscan), dim3(1024), 0, s, starts, prm, scan_tmp);
  hipLaunchKernelGGL(k_knn_scatter, dim3(nb), dim3(256), 0, s, points, P, cell_of, starts, counts, sorted);
  hipLaunchKernelGGL(k_knn_query, dim3(nb), dim3(256), 0, s, P, prm, starts, cell_of, sorted, out);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

}  // namespace gs

using namespace gs;

// Scratch of gs_knn3_meandist_static: the whole workspace of
// gs_knn3_meandist for GS_KNN_STATIC_MAX_POINTS points (about 24 MiB).  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code never passes this name to a kernel: it is a host-side stand-in, and
// the device address comes from hipGetSymbolAddress (knn_static_scratch).
static __device__ __align__(256) unsigned char scratch[knn_layout(GS_KNN_STATIC_MAX_POINTS).total];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device and kept.  nullptr when the lookup fails.
void* knn_static_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" size_t gs_knn_workspace_bytes(int P) { return knn_layout(P < 0 ? 0 : P).total; }

extern "C" int gs_knn3_meandist(const float* points, int P, float* out, void* ws, size_t ws_bytes, void* stream) {
  if (P < 0 || (P > 0 && (!points || !out || !ws))) return GS_E_ARG;
  if (P == 0) return GS_OK;
  const KnnLayout L = knn_layout(P);
  if (ws_bytes < L.total) return GS_E_WORKSPACE;
  return knn_launch(points, P, out, (char*)ws, L, (hipStream_t)stream);
}

extern "C" int gs_knn3_meandist_static(const float* points, int P, float* out, void* stream) {
  if (P < 0 || (P > 0 && (!points || !out))) return GS_E_ARG;
  if (P > GS_KNN_STATIC_MAX_POINTS) return GS_E_STATIC_LIMIT;
  if (P == 0) return GS_OK;
  void* ws = knn_static_scratch();
  if (!ws) return GS_E_SCRATCH;
  return knn_launch(points, P, out, (char*)ws, knn_layout(P), (hipStream_t)stream);
}
