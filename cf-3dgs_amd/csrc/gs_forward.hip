// gs_forward.hip -- forward rasteriser for gfx950.
//
// Pipeline (one frame):
//   k_preprocess   per Gaussian: optional rigid pose, near cull, 3-D and EWA
//                  2-D covariance, conic, radius, tile rect, SH -> RGB; writes a
//                  48-byte splat record and a workgroup-local prefix sum of the
//                  tiles touched (SH coefficients staged through LDS so the
//                  [P,M,3] array is read with coalesced 16-byte loads)
//   k_scan_top     one workgroup: exclusive prefix of the per-workgroup sums,
//                  total -> counters[0], status -> counters[1]
//   k_duplicate    per Gaussian: (tile << 32 | depth bits, id) pairs
//   radix sort     gs_sort.hip, 8-bit LSD passes over bits [0, 32 + bits(T))
//   k_ranges       [start, end) per tile in the sorted list
//   (the depth-first schedule instead sorts the Gaussians by depth, expands
//   them into tile ids in that order (k_df_*), sorts on the tile bits alone,
//   and records the tile ranges in that sort's last pass)
//   k_blend        one 16x16 tile per 256-thread workgroup; splats staged in
//                  LDS 256 at a time, each wave keeping its own list of the
//                  staged splats whose alpha can reach 1/255 on its 8 x 8
//                  pixels (splat_box); block-wide vote ends the tile early
//                  once every pixel is saturated; XCD-aware tile order.
#include "gs_common.hpp"
#include "gs_internal.hpp"

#include <vector>

namespace gs {

struct PreArgs {
  int P, M, D, W, H, gx, gy, prefiltered;
  float tanfovx, tanfovy, fx, fy, scale_mod;
  const float* view;
  const float* proj;
  const float* campos;
  const float* pose;
  const float* means;
  const float* shs;
  const float* colors;
  const float* opac;
  const float* scales;
  const float* rots;
  const float* cov_pre;
  int* radii;
  float4* splat;
  float* cov3D;
  uint32_t* tiles;
  uint32_t* local_incl;  // workgroup-local inclusive prefix of tiles
  uint32_t* block_sum;   // [nblocks] sums, then [nblocks] status flags
};

// Inclusive prefix sum over the 256 threads of a workgroup (uint32).
__device__ __forceinline__ uint32_t block_incl_scan_256(uint32_t v, uint32_t* s_tmp /*[>= 4]*/, uint32_t& total) {
  return block_scan_incl<256>(v, s_tmp, total);
}

// Workgroup bid of the nblocks of one frame's preprocess.  s_sh: [256][prow]
// staged SH coefficients (3 nc used per row).
__device__ __forceinline__ void preprocess_block(const PreArgs& a, int bid, int nblocks, float* s_sh) {
  __shared__ uint32_t s_scan[256];
  __shared__ uint32_t s_flag;
  const int base = bid * 256;
  const int i = base + threadIdx.x;
  const int nc = (a.D + 1) * (a.D + 1);
  const int stride = 3 * nc;
  const int prow = stride | 1;  // odd LDS row stride: no bank conflicts on the per-thread row reads
  if (threadIdx.x == 0) s_flag = 0;
  if (a.shs) {
    const int nloc = min(256, a.P - base);
    if (nc == a.M && (a.M & 3) == 0) {
      // contiguous rows of 3M floats, a multiple of 16 bytes when M % 4 == 0
      rows_to_lds(s_sh, reinterpret_cast<const float4*>(a.shs + (size_t)base * 3 * a.M), nloc, stride >> 2, prow);
    } else {
      for (int k = threadIdx.x; k < nloc * stride; k += 256) {
        const int g = k / stride, e = k - g * stride;
        s_sh[g * prow + e] = a.shs[(size_t)(base + g) * 3 * a.M + e];
      }
    }
  }
  __syncthreads();

  uint32_t my_tiles = 0;
  if (i < a.P) {
    a.radii[i] = 0;
    f3 p = {a.means[3 * i], a.means[3 * i + 1], a.means[3 * i + 2]};
    if (a.pose) {
      Pose ps;
      make_pose(a.pose, ps);
      p = pose_apply(ps, p);
    }
    const f3 pv = xf43(a.view, p);
    bool ok = true;
    if (pv.z <= 0.2f) {
      if (a.prefiltered) s_flag = GS_STATUS_PREFILTERED;
      ok = false;
    }
    float c3[6];
    float px = 0.f, py = 0.f, con_a = 0.f, con_b = 0.f, con_c = 0.f;
    int ir = 0;
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    if (ok) {
      const f3 ph = xf43(a.proj, p);
      const float pw = 1.0f / (xf44w(a.proj, p) + 0.0000001f);
      const float ndc_x = ph.x * pw, ndc_y = ph.y * pw;
      if (a.cov_pre) {
#pragma unroll
        for (int k = 0; k < 6; ++k) c3[k] = a.cov_pre[6 * i + k];
      } else {
        float R[9];
        const float4 q = reinterpret_cast<const float4*>(a.rots)[i];
        quat_to_R(q.x, q.y, q.z, q.w, R);
        const float s0 = a.scale_mod * a.scales[3 * i], s1 = a.scale_mod * a.scales[3 * i + 1],
                    s2 = a.scale_mod * a.scales[3 * i + 2];
        float L[9];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          L[3 * r] = R[3 * r] * s0;
          L[3 * r + 1] = R[3 * r + 1] * s1;
          L[3 * r + 2] = R[3 * r + 2] * s2;
        }
#define GS_SIG(u, v) ((L[3 * (u)] * L[3 * (v)] + L[3 * (u) + 1] * L[3 * (v) + 1]) + L[3 * (u) + 2] * L[3 * (v) + 2])
        c3[0] = GS_SIG(0, 0);
        c3[1] = GS_SIG(0, 1);
        c3[2] = GS_SIG(0, 2);
        c3[3] = GS_SIG(1, 1);
        c3[4] = GS_SIG(1, 2);
        c3[5] = GS_SIG(2, 2);
#undef GS_SIG
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) a.cov3D[6 * i + k] = c3[k];
      const float limx = 1.3f * a.tanfovx, limy = 1.3f * a.tanfovy;
      const float txtz = pv.x / pv.z, tytz = pv.y / pv.z;
      const float tx = fminf(limx, fmaxf(-limx, txtz)) * pv.z;
      const float ty = fminf(limy, fmaxf(-limy, tytz)) * pv.z;
      const float tz = pv.z;
      const float J00 = a.fx / tz, J02 = -(a.fx * tx) / (tz * tz);
      const float J11 = a.fy / tz, J12 = -(a.fy * ty) / (tz * tz);
      float Tm0[3], Tm1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Tm0[c] = J00 * a.view[c * 4 + 0] + J02 * a.view[c * 4 + 2];
        Tm1[c] = J11 * a.view[c * 4 + 1] + J12 * a.view[c * 4 + 2];
      }
      const float S[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
      float MS0[3], MS1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        MS0[c] = (Tm0[0] * S[0][c] + Tm0[1] * S[1][c]) + Tm0[2] * S[2][c];
        MS1[c] = (Tm1[0] * S[0][c] + Tm1[1] * S[1][c]) + Tm1[2] * S[2][c];
      }
      const float ca = ((MS0[0] * Tm0[0] + MS0[1] * Tm0[1]) + MS0[2] * Tm0[2]) + 0.3f;
      const float cb = (MS0[0] * Tm1[0] + MS0[1] * Tm1[1]) + MS0[2] * Tm1[2];
      const float cc = ((MS1[0] * Tm1[0] + MS1[1] * Tm1[1]) + MS1[2] * Tm1[2]) + 0.3f;
      const float det = ca * cc - cb * cb;
      if (det == 0.0f) ok = false;
      if (ok) {
        const float det_inv = 1.0f / det;
        con_a = cc * det_inv;
        con_b = -cb * det_inv;
        con_c = ca * det_inv;
        const float mid = 0.5f * (ca + cc);
        const float disc = sqrtf(fmaxf(0.1f, mid * mid - det));
        const float l1 = mid + disc, l2 = mid - disc;
        const float radius = ceilf(3.0f * sqrtf(fmaxf(l1, l2)));
        px = ndc2pix(ndc_x, a.W);
        py = ndc2pix(ndc_y, a.H);
        if (!isfinite(px) || !isfinite(py) || !isfinite(radius) || !isfinite(con_a) || !isfinite(con_b) ||
            !isfinite(con_c) || !isfinite(pv.z))
          ok = false;
        if (ok) {
          ir = sat_i(radius);
          tile_rect(px, py, ir, a.gx, a.gy, x0, y0, x1, y1);
          my_tiles = (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0);
          if (my_tiles == 0) ok = false;
        }
      }
    }
    if (ok) {
      float rgb[3];
      uint32_t cbits = 0;
      if (a.shs) {
        const f3 dv = {p.x - a.campos[0], p.y - a.campos[1], p.z - a.campos[2]};
        const float len = sqrtf((dv.x * dv.x + dv.y * dv.y) + dv.z * dv.z);
        const float dx = dv.x / len, dy = dv.y / len, dz = dv.z / len;
        float b[16];
        sh_basis(a.D, dx, dy, dz, b);
        const float* sh = s_sh + threadIdx.x * prow;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          float acc = b[0] * sh[ch];
          for (int k = 1; k < nc; ++k) acc = acc + b[k] * sh[3 * k + ch];
          acc = acc + 0.5f;
          if (acc < 0.0f) cbits |= 1u << ch;
          rgb[ch] = fmaxf(acc, 0.0f);
        }
      } else {
        rgb[0] = a.colors[3 * i];
        rgb[1] = a.colors[3 * i + 1];
        rgb[2] = a.colors[3 * i + 2];
      }
      a.radii[i] = ir;
      a.splat[3 * i + 0] = make_float4(px, py, con_a, con_b);
      a.splat[3 * i + 1] = make_float4(con_c, a.opac[i], pv.z, __uint_as_float(cbits));
      a.splat[3 * i + 2] = make_float4(rgb[0], rgb[1], rgb[2], __int_as_float(ir));
    } else {
      my_tiles = 0;
    }
    a.tiles[i] = my_tiles;
  }
  uint32_t total;
  const uint32_t incl = block_incl_scan_256(my_tiles, s_scan, total);
  if (i < a.P) a.local_incl[i] = incl;
  if (threadIdx.x == 0) {
    a.block_sum[bid] = total;
    a.block_sum[nblocks + bid] = s_flag;
  }
}

__global__ __launch_bounds__(256) void k_preprocess(PreArgs a) {
  extern __shared__ float s_sh[];
  preprocess_block(a, blockIdx.x, gridDim.x, s_sh);
}

// The block of many frames' launch that frame f owns: the last f with
// prefix[f] <= b (frames owning no block share their start with the next).
__device__ __forceinline__ int frame_of_block(const uint32_t* prefix, int n, uint32_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Many frames' preprocess in one launch: frame f owns workgroups [prefix[f],
// prefix[f + 1]); the dynamic LDS covers the largest frame's SH rows.
__global__ __launch_bounds__(256) void k_preprocess_multi(const uint32_t* prefix, const PreArgs* args, int n) {
  extern __shared__ float s_sh[];
  const uint32_t b = blockIdx.x;
  const int f = frame_of_block(prefix, n, b);
  const PreArgs a = args[f];
  preprocess_block(a, (int)(b - prefix[f]), (int)(prefix[f + 1] - prefix[f]), s_sh);
}

// Single workgroup: exclusive prefix of the per-workgroup tile sums (in place),
// num_rendered and the OR of the status flags.
__device__ __forceinline__ void scan_top_block(uint32_t* block_sum, int nblocks, uint32_t* counters) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t s_carry, s_flags;
  const int t = threadIdx.x;
  if (t == 0) {
    s_carry = 0;
    s_flags = 0;
  }
  __syncthreads();
  uint32_t myflags = 0;
  for (int b0 = 0; b0 < nblocks; b0 += 1024) {
    const int b = b0 + t;
    const uint32_t v = b < nblocks ? block_sum[b] : 0u;
    if (b < nblocks) myflags |= block_sum[nblocks + b];
    uint32_t tot;
    const uint32_t incl = block_scan_incl<1024>(v, s, tot);
    const uint32_t carry = s_carry;
    if (b < nblocks) block_sum[b] = carry + incl - v;
    __syncthreads();
    if (t == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (myflags) atomicOr(&s_flags, myflags);
  __syncthreads();
  if (t == 0) {
    counters[0] = s_carry;
    counters[1] = s_flags;
  }
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* block_sum, int nblocks, uint32_t* counters) {
  scan_top_block(block_sum, nblocks, counters);
}

struct ScanTop {
  uint32_t* block_sum;
  uint32_t* counters;
  int nblocks;
};

// Workgroup f: the scan of frame f.
__global__ __launch_bounds__(1024) void k_scan_top_multi(const ScanTop* st) {
  const ScanTop t = st[blockIdx.x];
  scan_top_block(t.block_sum, t.nblocks, t.counters);
}

struct DupArgs {
  int P, gx, gy;
  long long capacity;
  const float4* splat;
  const uint32_t* tiles;
  const uint32_t* local_incl;
  const uint32_t* block_excl;
  uint32_t* counters;
  uint64_t* keys;
  uint32_t* vals;
};

__global__ __launch_bounds__(256) void k_duplicate(DupArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t n = a.counters[0];
  if ((long long)n > a.capacity) {
    if (i == 0) {
      atomicOr(&a.counters[1], GS_STATUS_CAPACITY);
      atomicOr(&a.counters[3], GS_STATUS_CAPACITY);  // sticky copy
    }
    return;
  }
  if (i >= a.P) return;
  const uint32_t nt = a.tiles[i];
  if (nt == 0) return;
  uint32_t off = a.block_excl[i >> 8] + a.local_incl[i] - nt;
  const float4 s0 = a.splat[3 * i];
  const float4 s1 = a.splat[3 * i + 1];
  const float4 s2 = a.splat[3 * i + 2];
  int x0, y0, x1, y1;
  tile_rect(s0.x, s0.y, __float_as_int(s2.w), a.gx, a.gy, x0, y0, x1, y1);
  const uint64_t dbits = (uint64_t)__float_as_uint(s1.z);
  for (int y = y0; y < y1; ++y)
    for (int x = x0; x < x1; ++x) {
      a.keys[off] = ((uint64_t)(uint32_t)(y * a.gx + x) << 32) | dbits;
      a.vals[off] = (uint32_t)i;
      ++off;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, long long n, uint32_t v) {
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) p[k] = v;
}

__global__ __launch_bounds__(256) void k_ranges(const uint64_t* keys, const uint32_t* counters, long long capacity,
                                                uint32_t* ranges) {
  const uint32_t n0 = counters[0];
  const uint32_t n = ((long long)n0 > capacity) ? 0u : n0;
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const uint32_t t = (uint32_t)(keys[k] >> 32);
  if (k == 0) {
    ranges[2 * t] = 0;
  } else {
    const uint32_t tp = (uint32_t)(keys[k - 1] >> 32);
    if (t != tp) {
      ranges[2 * tp + 1] = k;
      ranges[2 * t] = k;
    }
  }
  if (k == n - 1) ranges[2 * t + 1] = n;
}

// ---- depth-first binning -------------------------------------------------
// Keys for the depth sort: depth bits of visible Gaussians (positive floats
// order as their bit patterns), 0xffffffff for the rest (sorted last, and they
// touch no tile).  counters[2] = P is the sort's element count.  The tile
// rect of each visible Gaussian is packed into 8 bytes here, while the splat
// records are read in order, so the duplicate (which visits Gaussians in
// depth order, i.e. at random) reads one 8-byte record instead of the tile
// count and two 16-byte splat vectors.
// Also sets the n_ranges words of the tile ranges to [UINT_MAX, 0) per tile,
// the start the list sort's atomic min / max needs (the grid covers
// max(P, n_ranges) threads).
__device__ __forceinline__ void df_keys_block(int P, const uint32_t* tiles, const float4* splat, int gx, int gy,
                                              uint32_t* keys, uint32_t* vals, uint2* rect, uint32_t* counters,
                                              uint32_t* ranges, int n_ranges, int bid) {
  const int i = bid * 256 + threadIdx.x;
  if (i == 0) counters[2] = (uint32_t)P;
  if (i < n_ranges) ranges[i] = (i & 1) ? 0u : 0xffffffffu;
  if (i >= P) return;
  const bool vis = tiles[i] != 0u;
  keys[i] = vis ? __float_as_uint(splat[3 * i + 1].z) : 0xffffffffu;
  vals[i] = (uint32_t)i;
  if (vis) {
    const float4 s0 = splat[3 * i];
    const float4 s2 = splat[3 * i + 2];
    int x0, y0, x1, y1;
    tile_rect(s0.x, s0.y, __float_as_int(s2.w), gx, gy, x0, y0, x1, y1);
    rect[i] = make_uint2((uint32_t)x0 | ((uint32_t)y0 << 16), (uint32_t)x1 | ((uint32_t)y1 << 16));
  }
}

__global__ __launch_bounds__(256) void k_df_keys(int P, const uint32_t* tiles, const float4* splat, int gx, int gy,
                                                 uint32_t* keys, uint32_t* vals, uint2* rect, uint32_t* counters,
                                                 uint32_t* ranges, int n_ranges) {
  df_keys_block(P, tiles, splat, gx, gy, keys, vals, rect, counters, ranges, n_ranges, blockIdx.x);
}

struct DfKeysArgs {
  int P, gx, gy, n_ranges;
  const uint32_t* tiles;
  const float4* splat;
  uint32_t* keys;
  uint32_t* vals;
  uint2* rect;
  uint32_t* counters;
  uint32_t* ranges;
};

// Many frames' depth keys in one launch: frame f owns workgroups [prefix[f],
// prefix[f + 1]).
__global__ __launch_bounds__(256) void k_df_keys_multi(const uint32_t* prefix, const DfKeysArgs* args, int n) {
  const uint32_t b = blockIdx.x;
  const int f = frame_of_block(prefix, n, b);
  const DfKeysArgs a = args[f];
  df_keys_block(a.P, a.tiles, a.splat, a.gx, a.gy, a.keys, a.vals, a.rect, a.counters, a.ranges, a.n_ranges,
                (int)(b - prefix[f]));
}

// Workgroup-local inclusive prefix of the tile counts in depth order.
__device__ __forceinline__ void df_scan_block(int P, const uint32_t* dvals, const uint32_t* tiles, uint32_t* dscan,
                                              uint32_t* dblocks, int bid) {
  __shared__ uint32_t s_scan[256];
  const int k = bid * 256 + threadIdx.x;
  const uint32_t v = k < P ? tiles[dvals[k]] : 0u;
  uint32_t total;
  const uint32_t incl = block_incl_scan_256(v, s_scan, total);
  if (k < P) dscan[k] = incl;
  if (threadIdx.x == 0) dblocks[bid] = total;
}

__global__ __launch_bounds__(256) void k_df_scan(int P, const uint32_t* dvals, const uint32_t* tiles, uint32_t* dscan,
                                                 uint32_t* dblocks) {
  df_scan_block(P, dvals, tiles, dscan, dblocks, blockIdx.x);
}

// One workgroup: in-place exclusive prefix of n values.
__device__ __forceinline__ void excl_scan_block(uint32_t* data, int n) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t s_carry;
  const int t = threadIdx.x;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < n; b0 += 1024) {
    const int b = b0 + t;
    const uint32_t v = b < n ? data[b] : 0u;
    uint32_t tot;
    const uint32_t incl = block_scan_incl<1024>(v, s, tot);
    const uint32_t carry = s_carry;
    if (b < n) data[b] = carry + incl - v;
    __syncthreads();
    if (t == 0) s_carry = carry + tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_excl_scan_1wg(uint32_t* data, int n) { excl_scan_block(data, n); }

struct DfDupArgs {
  int P, gx, gy;
  long long capacity;
  const uint2* rect;  // packed tile rect per Gaussian (k_df_keys)
  const uint32_t* tiles;
  const uint32_t* dvals;
  const uint32_t* dscan;
  const uint32_t* dblocks;
  uint32_t* counters;
  uint32_t* keys;  // tile ids, or tile << id_bits | id when id_bits > 0
  uint32_t* vals;  // ids (unused when id_bits > 0)
  int id_bits;
};

// Duplicate tile ids in depth order (sorted position k -> Gaussian dvals[k]).
// Each Gaussian's entries form one run at its prefix offset.  The wave
// expands its 64 runs cooperatively: output j of the wave's concatenated runs
// goes to lane j mod 64, which finds its run by binary search over the
// wave's exclusive prefix, so consecutive lanes write consecutive entries of
// a run (a run leaves as one contiguous segment, not as one word from each of
// 64 places per store).  Entry order and values are those of a per-Gaussian
// loop over the rect, row by row.
__device__ __forceinline__ void df_duplicate_block(const DfDupArgs& a, int bid) {
  __shared__ uint32_t s_excl[4][64], s_off[4][64], s_g[4][64], s_xy[4][64], s_w[4][64];
  const int k = bid * 256 + threadIdx.x;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t n = a.counters[0];
  if ((long long)n > a.capacity) {
    if (k == 0) {
      atomicOr(&a.counters[1], GS_STATUS_CAPACITY);
      atomicOr(&a.counters[3], GS_STATUS_CAPACITY);  // sticky copy
    }
    return;
  }
  uint32_t nt = 0, off = 0, g = 0, x0 = 0, y0 = 0, rw = 1;
  if (k < a.P) {
    g = a.dvals[k];
    nt = a.tiles[g];
    if (nt) {
      off = a.dblocks[k >> 8] + a.dscan[k] - nt;
      const uint2 r = a.rect[g];
      x0 = r.x & 0xffffu;
      y0 = r.x >> 16;
      rw = (r.y & 0xffffu) - x0;
    }
  }
  // wave exclusive prefix of the run lengths
  uint32_t incl = nt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  s_excl[w][lane] = incl - nt;
  s_off[w][lane] = off;
  s_g[w][lane] = g;
  s_xy[w][lane] = x0 | (y0 << 16);
  s_w[w][lane] = rw;
  __builtin_amdgcn_wave_barrier();
  for (uint32_t o = (uint32_t)lane; o < total; o += 64) {
    // run j = last lane with excl <= o (empty runs share their successor's
    // excl and lose to it)
    int j = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
      if (s_excl[w][j + step] <= o) j += step;
    const uint32_t r = o - s_excl[w][j];
    const uint32_t wj = s_w[w][j], xy = s_xy[w][j];
    const uint32_t ry = r / wj, rx = r - ry * wj;
    const uint32_t dst = s_off[w][j] + r;
    const uint32_t tile = ((xy >> 16) + ry) * (uint32_t)a.gx + (xy & 0xffffu) + rx;
    if (a.id_bits) {
      a.keys[dst] = (tile << a.id_bits) | s_g[w][j];
    } else {
      a.keys[dst] = tile;
      a.vals[dst] = s_g[w][j];
    }
  }
}

__global__ __launch_bounds__(256) void k_df_duplicate(DfDupArgs a) { df_duplicate_block(a, blockIdx.x); }

// The depth-first binning's steps between its two sorts for one frame of a
// multi-frame launch.
struct DfMid {
  DfDupArgs dup;
  uint32_t* dscan;
  uint32_t* dblocks;
  int nblocks;
};

// Many frames' workgroup-local scans of the tile counts in depth order: frame
// f owns workgroups [prefix[f], prefix[f + 1]).
__global__ __launch_bounds__(256) void k_df_scan_multi(const uint32_t* prefix, const DfMid* mid, int n) {
  const uint32_t b = blockIdx.x;
  const int f = frame_of_block(prefix, n, b);
  const DfMid m = mid[f];
  df_scan_block(m.dup.P, m.dup.dvals, m.dup.tiles, m.dscan, m.dblocks, (int)(b - prefix[f]));
}

// Workgroup f: the exclusive prefix of frame f's block totals.
__global__ __launch_bounds__(1024) void k_excl_scan_multi(const DfMid* mid) {
  const DfMid m = mid[blockIdx.x];
  excl_scan_block(m.dblocks, m.nblocks);
}

// Many frames' duplicates, on the same workgroup ranges as k_df_scan_multi.
__global__ __launch_bounds__(256) void k_df_duplicate_multi(const uint32_t* prefix, const DfMid* mid, int n) {
  const uint32_t b = blockIdx.x;
  const int f = frame_of_block(prefix, n, b);
  const DfDupArgs a = mid[f].dup;
  df_duplicate_block(a, (int)(b - prefix[f]));
}

struct BlendArgs {
  int W, H, gx;
  const uint32_t* ranges;
  const uint32_t* list;  // sorted entries; the Gaussian id is entry & id_mask
  uint32_t id_mask;
  const float4* splat;
  const float* bg;
  float* out_color;
  float* out_depth;
  float* out_alpha;
  float* final_T;
  uint32_t* n_contrib;
  // backward support: per-pixel state at every 256-entry batch boundary of
  // the tile's list ([ckpt][T, C0, C1, C2, D][256]), the tile's largest
  // contributor count and the pre-background sums ([C0, C1, C2, D][H*W])
  const uint32_t* ckpt_base;
  float* ckpt;
  uint32_t* tile_maxc;
  const uint32_t* tile_order;  // the tiles, longest list first (k_tile_ckpt_scan)
  float* accum;
  // per (tile, batch) at ckpt_base[tile] + tile + batch: the longest
  // per-wave list of the batch, the backward's cost estimate for that chunk
  uint32_t* bweight;
  // the backward's work items, appended as tiles finish: [0] count, then
  // (tile, chunk << 6 | cost class) -- cleared by k_tile_ckpt_scan
  uint32_t* item_log;
};

__device__ __forceinline__ void blend_tile(const BlendArgs& a, uint32_t tile) {
  __shared__ float4 s_a[256];  // x, y, conic a, conic b
  __shared__ float4 s_b[256];  // conic c, opacity, depth, -
  __shared__ float4 s_c[256];  // r, g, b, -
  __shared__ uint32_t s_mask[256];    // wave rows each staged splat can reach (splat_box)
  __shared__ __align__(4) uint8_t s_list[4][256];  // per wave: the staged splats that reach its 8 x 8 pixels
  __shared__ uint32_t s_maxc;
  __shared__ uint32_t s_bw;  // longest per-wave list of the current batch
  const int tx = tile % a.gx, ty = tile / a.gx;
  int pdx, pdy;
  tile_pixel(threadIdx.x, pdx, pdy);
  const int px = tx * kBX + pdx, py = ty * kBY + pdy;
  const bool inside = px < a.W && py < a.H;
  const float pfx = (float)px, pfy = (float)py;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t r0 = a.ranges[2 * tile], r1 = a.ranges[2 * tile + 1];
  bool done = !inside;
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, D = 0.0f;
  uint32_t last = 0;
  const float a_min = 1.0f / 255.0f;
  if (threadIdx.x == 0) s_maxc = 0;
  uint32_t batch = 0;
  for (uint32_t base = r0; base < r1; base += 256, ++batch) {
    if (batch > 0 && inside) {
      float* ck = a.ckpt + (size_t)(a.ckpt_base[tile] + batch - 1) * 5 * 256 + threadIdx.x;
      ck[0] = T;
      ck[256] = C0;
      ck[512] = C1;
      ck[768] = C2;
      ck[1024] = D;
    }
    if (__syncthreads_count(done) == 256) break;
    if (threadIdx.x == 0) {
      if (batch > 0) a.bweight[a.ckpt_base[tile] + tile + batch - 1] = s_bw;
      s_bw = 0;
    }
    const uint32_t k = base + threadIdx.x;
    uint32_t m = 0u;
    if (k < r1) {
      const uint32_t g = a.list[k] & a.id_mask;
      const float4 sa = a.splat[3 * g];
      s_a[threadIdx.x] = sa;
      float4 b = a.splat[3 * g + 1];
      b.w = skip_threshold(b.y);  // (the clamp bits in .w are not needed here)
      s_b[threadIdx.x] = b;
      s_c[threadIdx.x] = a.splat[3 * g + 2];
      m = splat_quads(sa.x, sa.y, sa.z, sa.w, b.x, b.w, (float)(tx * kBX), (float)(ty * kBY));
    }
    s_mask[threadIdx.x] = m;
    __syncthreads();
    // a splat that misses the wave's pixels is one every lane would skip
    const int nw = compact_wave_list(s_mask, s_list[wv], wv, lane);
    if (lane == 0) atomicMax(&s_bw, (uint32_t)nw);
    const uint32_t first = base - r0 + 1;  // contributor index of staged position 0
    // alpha of staged splat j at this pixel, 0 when it is skipped.  Branch
    // free (the exponential is evaluated and then discarded for skipped
    // pairs), so the loads and exponentials of the entries of one step
    // overlap instead of each waiting behind its own branch.
    auto alpha_at = [&](int j) -> float {
      const float4 ga = s_a[j];
      const float4 gb = s_b[j];
      const float dx = ga.x - pfx, dy = ga.y - pfy;
      const float power = blend_power(ga.z, ga.w, gb.x, dx, dy);
      const float alpha = fminf(0.99f, gb.y * spec_expf(power));
      // gb.w: alpha < 1/255 for sure below it
      const bool skip = power > 0.0f || power < gb.w || alpha < a_min;
      return skip ? 0.0f : alpha;
    };
    // in list order; selects instead of branches (the same arithmetic: a
    // skipped or saturating pair leaves the state as it was)
    auto apply = [&](int j, float alpha) {
      const float test_T = T * (1.0f - alpha);
      const bool live = alpha != 0.0f && !done;
      const bool stop = live && test_T < 0.0001f;
      const bool upd = live && !stop;
      const float w = alpha * T;
      const float4 gc = s_c[j];
      const float dz = s_b[j].z;
      const float n0 = fmaf(gc.x, w, C0), n1 = fmaf(gc.y, w, C1), n2 = fmaf(gc.z, w, C2), nd = fmaf(dz, w, D);
      C0 = upd ? n0 : C0;
      C1 = upd ? n1 : C1;
      C2 = upd ? n2 : C2;
      D = upd ? nd : D;
      T = upd ? test_T : T;
      last = upd ? first + (uint32_t)j : last;
      done = done || stop;
    };
    // four list entries per step: their alphas do not depend on each other,
    // so the four exponentials are in flight together (the heaviest tiles run
    // nearly alone at the end of the launch, where this latency is the time);
    // they are applied in order.  The indices come in one 4-byte LDS read;
    // slots past nw hold stale indices of staged splats and are masked.
    for (int i = 0; !done && i < nw; i += 4) {
      const uint32_t jj = *reinterpret_cast<const uint32_t*>(&s_list[wv][i]);
      const int j0 = jj & 0xffu, j1 = (jj >> 8) & 0xffu, j2 = (jj >> 16) & 0xffu, j3 = jj >> 24;
      const int left = nw - i;
      const float a0 = alpha_at(j0);
      const float a1 = left > 1 ? alpha_at(j1) : 0.0f;
      const float a2 = left > 2 ? alpha_at(j2) : 0.0f;
      const float a3 = left > 3 ? alpha_at(j3) : 0.0f;
      apply(j0, a0);
      apply(j1, a1);
      apply(j2, a2);
      apply(j3, a3);
    }
  }
  if (inside) {
    const int pix = py * a.W + px;
    const int HW = a.H * a.W;
    a.final_T[pix] = T;
    a.n_contrib[pix] = last;
    a.out_color[pix] = fmaf(T, a.bg[0], C0);
    a.out_color[HW + pix] = fmaf(T, a.bg[1], C1);
    a.out_color[2 * HW + pix] = fmaf(T, a.bg[2], C2);
    a.out_depth[pix] = D;
    if (a.out_alpha) a.out_alpha[pix] = 1.0f - T;
    a.accum[pix] = C0;
    a.accum[HW + pix] = C1;
    a.accum[2 * HW + pix] = C2;
    a.accum[3 * HW + pix] = D;
    if (last) atomicMax(&s_maxc, last);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.tile_maxc[tile] = s_maxc;
    if (batch > 0) a.bweight[a.ckpt_base[tile] + tile + batch - 1] = s_bw;  // the last batch run
  }
  __syncthreads();
  // The backward's work items of this tile: its 256-entry chunks up to the
  // last contributor, each with a cost class from the longest per-wave list
  // the blend had in the same batch (class 0 = longest).
  if (threadIdx.x < 64) {
    const uint32_t nch = (s_maxc + 255u) / 256u;
    const uint32_t wbase = a.ckpt_base[tile] + tile;
    for (uint32_t c = lane; c < nch; c += 64) {
      const uint32_t w = a.bweight[wbase + c];
      const uint32_t cls = 63u - min(w >> 2, 63u);
      const uint32_t slot = atomicAdd(&a.item_log[0], 1u);
      a.item_log[1 + 2 * slot] = tile;
      a.item_log[2 + 2 * slot] = (c << 6) | cls;
    }
  }
}

__global__ __launch_bounds__(256) void k_blend(BlendArgs a) { blend_tile(a, a.tile_order[blockIdx.x]); }

// Frame f of a set whose per-frame block counts have the exclusive prefix
// `prefix` (n + 1 entries): the last f with prefix[f] <= b.
__device__ __forceinline__ int frame_of(const uint32_t* prefix, int n, uint32_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Blend of many frames in one launch: one workgroup per tile of every frame
// (each frame's tiles in its longest-first order).
__global__ __launch_bounds__(256) void k_blend_multi(const uint32_t* prefix, const BlendArgs* args, int n) {
  const uint32_t b = blockIdx.x;
  const int f = frame_of(prefix, n, b);
  const BlendArgs a = args[f];
  blend_tile(a, a.tile_order[b - prefix[f]]);
}

// One workgroup: checkpoint base of every tile = exclusive prefix over tiles
// of (batches in its list - 1); ckpt_base[n_tiles] = total checkpoints.
// It also clears the blend's work-item log.
// Also turns the ranges of empty tiles, left at [UINT_MAX, 0) by the
// depth-first list sort, into [0, 0), and orders the tiles for the blend,
// longest list first (a counting sort on the list length in 64-entry steps,
// capped at 255 steps; the order within a step is whatever the LDS atomics
// give, which no result depends on).  A launch lasts as long as its last
// tiles: the longest lists start first, and consecutive workgroups go to
// different XCDs, so the long lists of one image region spread over the GPU.
__device__ __forceinline__ void tile_ckpt_scan_block(uint32_t* ranges, int nt, uint32_t* ckpt_base,
                                                     uint32_t* item_log, uint32_t* tile_order) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t s_carry;
  __shared__ uint32_t s_cls[256];
  const int t = threadIdx.x;
  if (t == 0) {
    s_carry = 0;
    item_log[0] = 0;
  }
  __syncthreads();
  for (int b0 = 0; b0 < nt; b0 += 1024) {
    const int b = b0 + t;
    uint32_t v = 0;
    if (b < nt) {
      uint32_t r0 = ranges[2 * b];
      const uint32_t r1 = ranges[2 * b + 1];
      if (r0 > r1) {
        r0 = r1;  // r1 = 0: an empty tile
        ranges[2 * b] = 0u;
      }
      const uint32_t len = r1 - r0;
      const uint32_t nb = (len + 255) / 256;
      v = nb > 0 ? nb - 1 : 0;
    }
    uint32_t tot;
    const uint32_t incl = block_scan_incl<1024>(v, s, tot);
    const uint32_t carry = s_carry;
    if (b < nt) ckpt_base[b] = carry + incl - v;
    __syncthreads();
    if (t == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (t == 0) ckpt_base[nt] = s_carry;
  if (t < 256) s_cls[t] = 0;
  __syncthreads();
  auto cls_of = [&](int b) { return min((ranges[2 * b + 1] - ranges[2 * b] + 63u) / 64u, 255u); };
  for (int b = t; b < nt; b += 1024) atomicAdd(&s_cls[cls_of(b)], 1u);
  __syncthreads();
  {
    // first position of each class, longest class first
    const uint32_t v = t < 256 ? s_cls[255 - t] : 0u;
    uint32_t tot;
    const uint32_t incl = block_scan_incl<1024>(v, s, tot);
    __syncthreads();
    if (t < 256) s_cls[255 - t] = incl - v;
  }
  __syncthreads();
  for (int b = t; b < nt; b += 1024) tile_order[atomicAdd(&s_cls[cls_of(b)], 1u)] = (uint32_t)b;
}

__global__ __launch_bounds__(1024) void k_tile_ckpt_scan(uint32_t* ranges, int nt, uint32_t* ckpt_base,
                                                         uint32_t* item_log, uint32_t* tile_order) {
  tile_ckpt_scan_block(ranges, nt, ckpt_base, item_log, tile_order);
}

struct CkptScan {
  uint32_t* ranges;
  uint32_t* ckpt_base;
  uint32_t* item_log;
  uint32_t* tile_order;
  int nt;
};

// Workgroup f: the checkpoint scan of frame f.
__global__ __launch_bounds__(1024) void k_tile_ckpt_scan_multi(const CkptScan* cs) {
  const CkptScan c = cs[blockIdx.x];
  tile_ckpt_scan_block(c.ranges, c.nt, c.ckpt_base, c.item_log, c.tile_order);
}

__global__ __launch_bounds__(256) void k_mark_visible(const float* means, int P, const float* view,
                                                      uint8_t* visible) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const f3 p = {means[3 * i], means[3 * i + 1], means[3 * i + 2]};
  visible[i] = xf43(view, p).z > 0.2f ? 1 : 0;
}

}  // namespace gs

// ---------------------------------------------------------------------------
// C ABI
using namespace gs;

namespace {
inline int grid_tiles(int W, int H, int& gx, int& gy) {
  gx = (W + kBX - 1) / kBX;
  gy = (H + kBY - 1) / kBY;
  return gx * gy;
}
inline int nblocks256(long long n) { return (int)((n + 255) / 256); }
}  // namespace

extern "C" int gs_geom_layout(int P, gs_geom_layout_t* out) {
  if (P < 0 || !out) return GS_E_ARG;
  const size_t p = (size_t)P;
  size_t o = 0;
  out->splat = o;
  o = align_up(o + 48 * p, 256);
  out->cov3D = o;
  o = align_up(o + 24 * p, 256);
  out->tiles = o;
  o = align_up(o + 4 * p, 256);
  out->offsets = o;
  o = align_up(o + 4 * p, 256);
  out->scan_tmp = o;
  o = align_up(o + 8 * (size_t)nblocks256(P) + 8, 256);
  out->total = o;
  return GS_OK;
}

extern "C" int gs_binning_layout(long long capacity, int P, int width, int height, int binning,
                                 gs_binning_layout_t* out) {
  if (capacity < 0 || capacity > 0x7fffffffLL || P < 0 || width <= 0 || height <= 0 || !out) return GS_E_ARG;
  if (binning != GS_BINNING_DEPTH_FIRST && binning != GS_BINNING_SORT64) return GS_E_ARG;
  int gx, gy;
  const int nt = grid_tiles(width, height, gx, gy);
  const size_t c = (size_t)capacity;
  const bool df = binning == GS_BINNING_DEPTH_FIRST;
  const size_t kb = df ? 4 : 8;
  size_t o = 0;
  out->keys_dup = o;
  o = align_up(o + kb * c, 256);
  out->keys_alt = o;
  o = align_up(o + kb * c, 256);
  out->vals_dup = o;
  o = align_up(o + 4 * c, 256);
  out->vals_alt = o;
  o = align_up(o + 4 * c, 256);
  const size_t p = df ? (size_t)P : 0;
  out->dkeys_a = o;
  o = align_up(o + 4 * p, 256);
  out->dvals_a = o;
  o = align_up(o + 4 * p, 256);
  out->dkeys_b = o;
  o = align_up(o + 4 * p, 256);
  out->dvals_b = o;
  o = align_up(o + 4 * p, 256);
  out->dscan = o;
  o = align_up(o + 4 * p, 256);
  out->dblocks = o;
  o = align_up(o + 4 * (p / 256 + 2), 256);
  out->drect = o;
  o = align_up(o + 8 * p, 256);
  out->hist = o;
  {
    // histogram space for both sorts that share it (their tile sizes may differ)
    const size_t t_list = sort_tmp_bytes(capacity), t_depth = sort_tmp_bytes((long long)p);
    o = align_up(o + (t_list > t_depth ? t_list : t_depth), 256);
  }
  // blend checkpoints: at most capacity / 256 batch boundaries, 5 x 256 floats each
  out->ckpt = o;
  o = align_up(o + (size_t)(c / 256 + 1) * 5 * 256 * sizeof(float), 256);
  // per (tile, batch) cost estimates: tile t's batches start at ckpt_base[t] + t
  out->bweight = o;
  o = align_up(o + 4 * (c / 256 + 2 * (size_t)nt + 2), 256);
  out->items = o;
  o = align_up(o + 4 * (1 + 2 * (c / 256 + (size_t)nt + 1)), 256);
  out->item_log = o;
  o = align_up(o + 4 * (1 + 2 * (c / 256 + (size_t)nt + 1)), 256);
  out->total = o;
  const int end_bit = (df ? 0 : 32) + bits_for((uint32_t)nt);
  out->sort_passes = sort_passes(end_bit > 0 ? end_bit : 1);
  out->key_bytes = (int)kb;
  const bool in_b = (out->sort_passes & 1) != 0;
  out->keys_sorted = in_b ? out->keys_alt : out->keys_dup;
  out->vals_sorted = in_b ? out->vals_alt : out->vals_dup;
  // one word per list entry when the Gaussian id and the tile id fit together
  const int idb = bits_for((uint32_t)(P > 1 ? P : 2));
  out->id_bits = (df && idb + bits_for((uint32_t)nt) <= 32) ? idb : 0;
  if (out->id_bits) out->vals_sorted = out->keys_sorted;
  return GS_OK;
}

extern "C" int gs_image_layout(int width, int height, gs_image_layout_t* out) {
  if (width <= 0 || height <= 0 || !out) return GS_E_ARG;
  int gx, gy;
  const int nt = grid_tiles(width, height, gx, gy);
  const size_t hw = (size_t)width * height;
  size_t o = 0;
  out->ranges = o;
  o = align_up(o + 8 * (size_t)nt, 256);
  out->final_T = o;
  o = align_up(o + 4 * hw, 256);
  out->n_contrib = o;
  o = align_up(o + 4 * hw, 256);
  out->accum = o;
  o = align_up(o + 16 * hw, 256);
  out->ckpt_base = o;
  o = align_up(o + 4 * ((size_t)nt + 1), 256);
  out->tile_maxc = o;
  o = align_up(o + 4 * (size_t)nt, 256);
  out->tile_order = o;
  o = align_up(o + 4 * (size_t)nt, 256);
  out->total = o;
  return GS_OK;
}

static int check_frame(const gs_frame_t* f) {
  if (!f || f->width <= 0 || f->height <= 0 || !f->viewmatrix || !f->projmatrix || !f->bg) return GS_E_ARG;
  if (f->sh_degree < 0 || f->sh_degree > 3) return GS_E_ARG;
  return GS_OK;
}

// The checked preprocess arguments of one frame.
static int make_pre_args(const gs_frame_t* f, const gs_gaussians_t* g, int* radii, void* geom, size_t geom_bytes,
                         const uint32_t* counters, PreArgs& a) {
  if (check_frame(f) != GS_OK || !g || g->P < 0 || !counters) return GS_E_ARG;
  const int P = g->P;
  if (P > 0 && (!g->means3D || !g->opacities || !radii || !geom)) return GS_E_ARG;
  // (empty inputs carry no pointers, so the exactly-one-of rules apply for P > 0)
  if (P > 0 && (g->shs == nullptr) == (g->colors_precomp == nullptr)) return GS_E_ARG;
  const bool has_sr = g->scales && g->rotations;
  if (P > 0 && has_sr == (g->cov3D_precomp != nullptr)) return GS_E_ARG;
  if (g->shs && (f->sh_degree + 1) * (f->sh_degree + 1) > g->M) return GS_E_ARG;
  if (g->shs && !f->campos) return GS_E_ARG;
  gs_geom_layout_t L;
  gs_geom_layout(P, &L);
  if (geom_bytes < L.total) return GS_E_WORKSPACE;
  char* base = (char*)geom;
  a.P = P;
  a.M = g->M;
  a.D = f->sh_degree;
  a.W = f->width;
  a.H = f->height;
  grid_tiles(a.W, a.H, a.gx, a.gy);
  a.prefiltered = f->prefiltered;
  a.tanfovx = f->tanfovx;
  a.tanfovy = f->tanfovy;
  a.fx = (float)a.W / (2.0f * f->tanfovx);
  a.fy = (float)a.H / (2.0f * f->tanfovy);
  a.scale_mod = f->scale_modifier;
  a.view = f->viewmatrix;
  a.proj = f->projmatrix;
  a.campos = f->campos;
  a.pose = f->pose;
  a.means = g->means3D;
  a.shs = g->shs;
  a.colors = g->colors_precomp;
  a.opac = g->opacities;
  a.scales = g->scales;
  a.rots = g->rotations;
  a.cov_pre = g->cov3D_precomp;
  a.radii = radii;
  a.splat = (float4*)(base + L.splat);
  a.cov3D = (float*)(base + L.cov3D);
  a.tiles = (uint32_t*)(base + L.tiles);
  a.local_incl = (uint32_t*)(base + L.offsets);
  a.block_sum = (uint32_t*)(base + L.scan_tmp);
  return GS_OK;
}

// Dynamic LDS of a preprocess launch: the staged SH rows of 256 Gaussians
// (an odd row stride).
static size_t preprocess_lds(const PreArgs& a) {
  const int nc = (a.D + 1) * (a.D + 1);
  return a.shs ? (size_t)256 * ((3 * nc) | 1) * sizeof(float) : 0;
}

extern "C" int gs_preprocess(const gs_frame_t* f, const gs_gaussians_t* g, int* radii, void* geom,
                             size_t geom_bytes, uint32_t* counters, void* stream) {
  PreArgs a;
  const int rc = make_pre_args(f, g, radii, geom, geom_bytes, counters, a);
  if (rc != GS_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int nb = nblocks256(a.P);
  if (nb > 0) {
    const size_t shm = preprocess_lds(a);
    timing::Scope ts(timing::kPreprocess, s);
    hipLaunchKernelGGL(k_preprocess, dim3(nb), dim3(256), shm, s, a);
  }
  {
    timing::Scope ts(timing::kScan, s);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, a.block_sum, nb, counters);
  }
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// Binning of one frame: everything of the forward but the blend (ranges
// cleared, depth-first or 64-bit-key binning, ranges, checkpoint bases).
// Phases of the binning: kPrepare is everything that does not read
// num_rendered (clearing the tile ranges; in the depth-first schedule also
// the depth keys, the 32-bit depth sort and the prefix of the tile counts in
// depth order), so a caller can queue it before it waits for the readback;
// kResume is the rest, after kPrepare with the same arguments.  The other
// phases split the depth-first schedule around its two sorts instead, so
// that one segmented sort can serve many frames: kKeys is the depth keys
// alone; kAfterSort everything after the depth sort; kDup the tile-count
// prefix and the duplicate (after the depth sort, before the list sort);
// kTail what follows the list sort (the checkpoint scan).
enum BinPhase { kBinAll = 0, kPrepare = 1, kResume = 2, kKeys = 3, kAfterSort = 4, kDup = 5, kTail = 6 };

static int render_binning(const gs_frame_t* f, int P, const void* geom, uint32_t* counters, long long capacity,
                          void* binning, size_t binning_bytes, void* image, size_t image_bytes, hipStream_t s,
                          int phase = kBinAll) {
  const bool pre = phase == kBinAll || phase == kPrepare, post = phase == kBinAll || phase == kResume;
  const bool split = phase == kKeys || phase == kAfterSort || phase == kDup || phase == kTail;
  if (check_frame(f) != GS_OK || P < 0 || !counters || !image) return GS_E_ARG;
  if (P > 0 && !geom) return GS_E_ARG;
  if (capacity < 0 || capacity > 0x7fffffffLL) return GS_E_ARG;
  gs_geom_layout_t LG;
  gs_binning_layout_t LB;
  gs_image_layout_t LI;
  gs_geom_layout(P, &LG);
  if (gs_binning_layout(capacity, P, f->width, f->height, f->binning, &LB) != GS_OK) return GS_E_ARG;
  gs_image_layout(f->width, f->height, &LI);
  if (binning_bytes < LB.total || image_bytes < LI.total) return GS_E_WORKSPACE;
  if (capacity > 0 && !binning) return GS_E_ARG;
  int gx, gy;
  const int nt = grid_tiles(f->width, f->height, gx, gy);
  const char* gb = (const char*)geom;
  char* bb = (char*)binning;
  char* ib = (char*)image;
  uint32_t* ranges = (uint32_t*)(ib + LI.ranges);
  const bool df = capacity > 0 && P > 0 && f->binning == GS_BINNING_DEPTH_FIRST;
  if (split && !df) return GS_E_ARG;  // only the depth-first schedule has a depth sort to split around
  // depth-first stages: keys | depth sort | tile-count prefix | duplicate | list sort | checkpoint scan
  const bool do_keys = pre || phase == kKeys, do_dsort = pre;
  const bool do_dscan = pre || phase == kAfterSort || phase == kDup;
  const bool do_dup = post || phase == kAfterSort || phase == kDup, do_lsort = post || phase == kAfterSort;
  const bool do_ckpt = post || phase == kAfterSort || phase == kTail;
  if (pre && !df)  // the depth-first schedule clears them in k_df_keys
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(nblocks256(2LL * nt), 1024)), dim3(256), 0, s, ranges, 2LL * nt,
                       0u);
  if (df) {
    if (gx > 0xffff || gy > 0xffff) return GS_E_ARG;  // tile rects are packed in 16-bit fields
    const float4* splat = (const float4*)(gb + LG.splat);
    const uint32_t* tiles = (const uint32_t*)(gb + LG.tiles);
    uint32_t* dka = (uint32_t*)(bb + LB.dkeys_a);
    uint32_t* dva = (uint32_t*)(bb + LB.dvals_a);
    const int nbp = nblocks256(P);
    hipError_t e = hipSuccess;
    int in_b = sort_passes(32) & 1;  // where the depth sort (or a segmented sort of it) leaves its result
    if (do_keys || do_dsort) {
      {
        timing::Scope ts(timing::kSort, s);
        if (do_keys)
          hipLaunchKernelGGL(k_df_keys, dim3(std::max(nbp, nblocks256(2LL * nt))), dim3(256), 0, s, P, tiles,
                             splat, gx, gy, dka, dva, (uint2*)(bb + LB.drect), counters, ranges, 2 * nt);
        if (do_dsort)
          e = sort_pairs_u32(dka, dva, (uint32_t*)(bb + LB.dkeys_b), (uint32_t*)(bb + LB.dvals_b), counters + 2, P,
                             32, bb + LB.hist, s, &in_b);
      }
      if (e != hipSuccess) return GS_E_LAUNCH;
    }
    const uint32_t* dvals = (const uint32_t*)(bb + (in_b ? LB.dvals_b : LB.dvals_a));
    if (do_dscan) {
      timing::Scope ts(timing::kDuplicate, s);
      hipLaunchKernelGGL(k_df_scan, dim3(nbp), dim3(256), 0, s, P, dvals, tiles, (uint32_t*)(bb + LB.dscan),
                         (uint32_t*)(bb + LB.dblocks));
      hipLaunchKernelGGL(k_excl_scan_1wg, dim3(1), dim3(1024), 0, s, (uint32_t*)(bb + LB.dblocks), nbp);
    }
    if (do_dup) {
      timing::Scope ts(timing::kDuplicate, s);
      DfDupArgs d;
      d.P = P;
      d.gx = gx;
      d.gy = gy;
      d.capacity = capacity;
      d.rect = (const uint2*)(bb + LB.drect);
      d.tiles = tiles;
      d.dvals = dvals;
      d.dscan = (const uint32_t*)(bb + LB.dscan);
      d.dblocks = (const uint32_t*)(bb + LB.dblocks);
      d.counters = counters;
      d.keys = (uint32_t*)(bb + LB.keys_dup);
      d.vals = (uint32_t*)(bb + LB.vals_dup);
      d.id_bits = LB.id_bits;
      hipLaunchKernelGGL(k_df_duplicate, dim3(nbp), dim3(256), 0, s, d);
    }
    if (do_lsort) {
      timing::Scope ts(timing::kSort, s);
      const int tb = bits_for((uint32_t)nt) > 0 ? bits_for((uint32_t)nt) : 1;
      // the last pass also records the tile ranges
      if (LB.id_bits)  // one word per entry: sort on its tile bits only
        e = sort_keys_u32((uint32_t*)(bb + LB.keys_dup), (uint32_t*)(bb + LB.keys_alt), counters, capacity,
                          LB.id_bits, LB.id_bits + tb, bb + LB.hist, s, &in_b, ranges);
      else
        e = sort_pairs_u32((uint32_t*)(bb + LB.keys_dup), (uint32_t*)(bb + LB.vals_dup),
                           (uint32_t*)(bb + LB.keys_alt), (uint32_t*)(bb + LB.vals_alt), counters, capacity, tb,
                           bb + LB.hist, s, &in_b, ranges);
    }
    if (e != hipSuccess) return GS_E_LAUNCH;
  } else if (capacity > 0 && P > 0) {
    if (!post) return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
    DupArgs d;
    d.P = P;
    d.gx = gx;
    d.gy = gy;
    d.capacity = capacity;
    d.splat = (const float4*)(gb + LG.splat);
    d.tiles = (const uint32_t*)(gb + LG.tiles);
    d.local_incl = (const uint32_t*)(gb + LG.offsets);
    d.block_excl = (const uint32_t*)(gb + LG.scan_tmp);
    d.counters = counters;
    d.keys = (uint64_t*)(bb + LB.keys_dup);
    d.vals = (uint32_t*)(bb + LB.vals_dup);
    {
      timing::Scope ts(timing::kDuplicate, s);
      hipLaunchKernelGGL(k_duplicate, dim3(nblocks256(P)), dim3(256), 0, s, d);
    }
    const int end_bit = 32 + bits_for((uint32_t)nt);
    int in_b = 0;
    hipError_t e;
    {
      timing::Scope ts_sort(timing::kSort, s);
      e = sort_pairs_u64((uint64_t*)(bb + LB.keys_dup), (uint32_t*)(bb + LB.vals_dup), (uint64_t*)(bb + LB.keys_alt),
                         (uint32_t*)(bb + LB.vals_alt), counters, capacity, end_bit, bb + LB.hist, s, &in_b);
    }
    if (e != hipSuccess) return GS_E_LAUNCH;
    timing::Scope ts(timing::kRanges, s);
    hipLaunchKernelGGL(k_ranges, dim3(nblocks256(capacity)), dim3(256), 0, s,
                       (const uint64_t*)(bb + LB.keys_sorted), (const uint32_t*)counters, capacity, ranges);
  }
  if (post || (df && do_ckpt))
    hipLaunchKernelGGL(k_tile_ckpt_scan, dim3(1), dim3(1024), 0, s, ranges, nt, (uint32_t*)(ib + LI.ckpt_base),
                       (uint32_t*)(bb + LB.item_log), (uint32_t*)(ib + LI.tile_order));
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// Blend arguments of one frame (its workspaces as binned by render_binning).
static BlendArgs make_blend_args(const gs_frame_t* f, int P, const void* geom, long long capacity,
                                 void* binning, void* image, float* out_color, float* out_depth,
                                 float* out_alpha) {
  gs_geom_layout_t LG;
  gs_binning_layout_t LB;
  gs_image_layout_t LI;
  gs_geom_layout(P, &LG);
  gs_binning_layout(capacity, P, f->width, f->height, f->binning, &LB);
  gs_image_layout(f->width, f->height, &LI);
  int gx, gy;
  grid_tiles(f->width, f->height, gx, gy);
  const char* gb = (const char*)geom;
  char* bb = (char*)binning;
  char* ib = (char*)image;
  BlendArgs b;
  b.W = f->width;
  b.H = f->height;
  b.gx = gx;
  b.ranges = (const uint32_t*)(ib + LI.ranges);
  b.list = capacity > 0 ? (const uint32_t*)(bb + LB.vals_sorted) : nullptr;
  b.id_mask = LB.id_bits ? (1u << LB.id_bits) - 1u : 0xffffffffu;
  b.splat = (const float4*)(gb + LG.splat);
  b.bg = f->bg;
  b.out_color = out_color;
  b.out_depth = out_depth;
  b.out_alpha = out_alpha;
  b.final_T = (float*)(ib + LI.final_T);
  b.n_contrib = (uint32_t*)(ib + LI.n_contrib);
  b.ckpt_base = (const uint32_t*)(ib + LI.ckpt_base);
  b.ckpt = (float*)(bb + LB.ckpt);
  b.tile_maxc = (uint32_t*)(ib + LI.tile_maxc);
  b.tile_order = (const uint32_t*)(ib + LI.tile_order);
  b.bweight = (uint32_t*)(bb + LB.bweight);
  b.item_log = (uint32_t*)(bb + LB.item_log);
  b.accum = (float*)(ib + LI.accum);
  return b;
}

extern "C" int gs_render_binning(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                 long long capacity, void* binning, size_t binning_bytes, void* image,
                                 size_t image_bytes, void* stream) {
  return render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                        (hipStream_t)stream);
}

static int render_blend_from(int phase, const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                             long long capacity, void* binning, size_t binning_bytes, void* image, size_t image_bytes,
                             float* out_color, float* out_depth, float* out_alpha, hipStream_t s) {
  if (!out_color || !out_depth) return GS_E_ARG;
  const int rc = render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes, s, phase);
  if (rc != GS_OK) return rc;
  const BlendArgs b = make_blend_args(f, P, geom, capacity, binning, image, out_color, out_depth, out_alpha);
  int gx, gy;
  const int nt = grid_tiles(f->width, f->height, gx, gy);
  {
    timing::Scope ts(timing::kBlend, s);
    hipLaunchKernelGGL(k_blend, dim3(nt), dim3(256), 0, s, b);
  }
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_render(const gs_frame_t* f, int P, const void* geom, uint32_t* counters, long long capacity,
                         void* binning, size_t binning_bytes, void* image, size_t image_bytes, float* out_color,
                         float* out_depth, float* out_alpha, void* stream) {
  return render_blend_from(kBinAll, f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                           out_color, out_depth, out_alpha, (hipStream_t)stream);
}

extern "C" int gs_render_prepare(const gs_frame_t* f, int P, const void* geom, uint32_t* counters, long long capacity,
                                 void* binning, size_t binning_bytes, void* image, size_t image_bytes, void* stream) {
  return render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                        (hipStream_t)stream, kPrepare);
}

extern "C" int gs_render_resume(const gs_frame_t* f, int P, const void* geom, uint32_t* counters, long long capacity,
                                void* binning, size_t binning_bytes, void* image, size_t image_bytes,
                                float* out_color, float* out_depth, float* out_alpha, void* stream) {
  return render_blend_from(kResume, f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                           out_color, out_depth, out_alpha, (hipStream_t)stream);
}

// ---- the depth-first binning around a depth sort shared by many frames -----
extern "C" int gs_render_depth_keys(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                    long long capacity, void* binning, size_t binning_bytes, void* image,
                                    size_t image_bytes, void* stream) {
  return render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                        (hipStream_t)stream, kKeys);
}

// ---- many frames' preprocess and depth keys in three launches --------------
// Set layout: uint32 prefix[n + 1] of preprocess workgroup bounds per frame,
// uint32 prefix[n + 1] of depth-key workgroup bounds, then PreArgs[n],
// ScanTop[n] and DfKeysArgs[n], each array at a 256-byte boundary.
static size_t mpk_prefix2_offset(int n) { return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256); }
static size_t mpk_pre_offset(int n) { return 2 * mpk_prefix2_offset(n); }
static size_t mpk_top_offset(int n) { return mpk_pre_offset(n) + align_up(sizeof(PreArgs) * (size_t)n, 256); }
static size_t mpk_keys_offset(int n) { return mpk_top_offset(n) + align_up(sizeof(ScanTop) * (size_t)n, 256); }

extern "C" size_t gs_multi_preprocess_keys_bytes(int n) {
  return n <= 0 ? 0 : mpk_keys_offset(n) + sizeof(DfKeysArgs) * (size_t)n;
}

extern "C" int gs_multi_preprocess_keys_build(int n, const gs_frame_t* frames, const gs_gaussians_t* gauss,
                                              int* const* radii, void* const* geom, const size_t* geom_bytes,
                                              uint32_t* const* counters, const long long* capacity,
                                              void* const* binning, const size_t* binning_bytes,
                                              void* const* image, const size_t* image_bytes, void* host_set,
                                              size_t set_bytes, unsigned* launch) {
  if (n <= 0 || !frames || !gauss || !radii || !geom || !geom_bytes || !counters || !capacity || !binning ||
      !binning_bytes || !image || !image_bytes || !host_set || !launch)
    return GS_E_ARG;
  if (set_bytes < gs_multi_preprocess_keys_bytes(n)) return GS_E_WORKSPACE;
  uint32_t* pprefix = (uint32_t*)host_set;
  uint32_t* kprefix = (uint32_t*)((char*)host_set + mpk_prefix2_offset(n));
  PreArgs* pre = (PreArgs*)((char*)host_set + mpk_pre_offset(n));
  ScanTop* top = (ScanTop*)((char*)host_set + mpk_top_offset(n));
  DfKeysArgs* keys = (DfKeysArgs*)((char*)host_set + mpk_keys_offset(n));
  unsigned long long ptotal = 0, ktotal = 0;
  size_t shm = 0;
  for (int i = 0; i < n; ++i) {
    const gs_frame_t* f = &frames[i];
    const int rc = make_pre_args(f, &gauss[i], radii[i], geom[i], geom_bytes[i], counters[i], pre[i]);
    if (rc != GS_OK) return rc;
    const int P = gauss[i].P;
    // the depth keys of the depth-first schedule, as gs_render_depth_keys
    // takes them
    if (P <= 0 || f->binning != GS_BINNING_DEPTH_FIRST || capacity[i] <= 0 || capacity[i] > 0x7fffffffLL ||
        !binning[i] || !image[i])
      return GS_E_ARG;
    gs_geom_layout_t LG;
    gs_binning_layout_t LB;
    gs_image_layout_t LI;
    gs_geom_layout(P, &LG);
    if (gs_binning_layout(capacity[i], P, f->width, f->height, f->binning, &LB) != GS_OK) return GS_E_ARG;
    gs_image_layout(f->width, f->height, &LI);
    if (binning_bytes[i] < LB.total || image_bytes[i] < LI.total) return GS_E_WORKSPACE;
    int gx, gy;
    const int nt = grid_tiles(f->width, f->height, gx, gy);
    if (gx > 0xffff || gy > 0xffff) return GS_E_ARG;  // tile rects are packed in 16-bit fields
    const int nbp = nblocks256(P);
    pprefix[i] = (uint32_t)ptotal;
    ptotal += (unsigned long long)nbp;
    kprefix[i] = (uint32_t)ktotal;
    ktotal += (unsigned long long)std::max(nbp, nblocks256(2LL * nt));
    if (ptotal > 0x7fffffffULL || ktotal > 0x7fffffffULL) return GS_E_ARG;
    shm = std::max(shm, preprocess_lds(pre[i]));
    top[i].block_sum = pre[i].block_sum;
    top[i].counters = counters[i];
    top[i].nblocks = nbp;
    const char* gb = (const char*)geom[i];
    char* bb = (char*)binning[i];
    DfKeysArgs& k = keys[i];
    k.P = P;
    k.gx = gx;
    k.gy = gy;
    k.n_ranges = 2 * nt;
    k.tiles = (const uint32_t*)(gb + LG.tiles);
    k.splat = (const float4*)(gb + LG.splat);
    k.keys = (uint32_t*)(bb + LB.dkeys_a);
    k.vals = (uint32_t*)(bb + LB.dvals_a);
    k.rect = (uint2*)(bb + LB.drect);
    k.counters = counters[i];
    k.ranges = (uint32_t*)((char*)image[i] + LI.ranges);
  }
  pprefix[n] = (uint32_t)ptotal;
  kprefix[n] = (uint32_t)ktotal;
  launch[0] = (unsigned)ptotal;
  launch[1] = (unsigned)ktotal;
  launch[2] = (unsigned)shm;
  return GS_OK;
}

extern "C" int gs_multi_preprocess_keys(int n, const void* dev_set, const unsigned* launch, void* stream) {
  if (n <= 0 || !dev_set || !launch) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const char* d = (const char*)dev_set;
  if (launch[0] > 0) {
    timing::Scope ts(timing::kPreprocess, s);
    hipLaunchKernelGGL(k_preprocess_multi, dim3(launch[0]), dim3(256), (size_t)launch[2], s, (const uint32_t*)d,
                       (const PreArgs*)(d + mpk_pre_offset(n)), n);
  }
  {
    timing::Scope ts(timing::kScan, s);
    hipLaunchKernelGGL(k_scan_top_multi, dim3(n), dim3(1024), 0, s, (const ScanTop*)(d + mpk_top_offset(n)));
  }
  if (launch[1] > 0) {
    timing::Scope ts(timing::kSort, s);
    hipLaunchKernelGGL(k_df_keys_multi, dim3(launch[1]), dim3(256), 0, s,
                       (const uint32_t*)(d + mpk_prefix2_offset(n)), (const DfKeysArgs*)(d + mpk_keys_offset(n)), n);
  }
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// ---- many frames' binning steps around the list sort ----------------------
// Set layout: uint32 prefix[n + 1] of per-Gaussian workgroup bounds per frame,
// then DfMid[n] and CkptScan[n], each array at a 256-byte boundary.
static size_t mbm_mid_offset(int n) { return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256); }
static size_t mbm_ckpt_offset(int n) { return mbm_mid_offset(n) + align_up(sizeof(DfMid) * (size_t)n, 256); }

extern "C" size_t gs_multi_binning_steps_bytes(int n) {
  return n <= 0 ? 0 : mbm_ckpt_offset(n) + sizeof(CkptScan) * (size_t)n;
}

extern "C" int gs_multi_binning_steps_build(int n, const gs_frame_t* frames, const int* P, const void* const* geom,
                                            uint32_t* const* counters, const long long* capacity,
                                            void* const* binning, const size_t* binning_bytes, void* const* image,
                                            const size_t* image_bytes, void* host_set, size_t set_bytes,
                                            unsigned* total_blocks) {
  if (n <= 0 || !frames || !P || !geom || !counters || !capacity || !binning || !binning_bytes || !image ||
      !image_bytes || !host_set || !total_blocks)
    return GS_E_ARG;
  if (set_bytes < gs_multi_binning_steps_bytes(n)) return GS_E_WORKSPACE;
  uint32_t* prefix = (uint32_t*)host_set;
  DfMid* mid = (DfMid*)((char*)host_set + mbm_mid_offset(n));
  CkptScan* ck = (CkptScan*)((char*)host_set + mbm_ckpt_offset(n));
  unsigned long long total = 0;
  for (int i = 0; i < n; ++i) {
    const gs_frame_t* f = &frames[i];
    // the depth-first schedule's split phases, as gs_render_binning_duplicate
    // and gs_render_binning_after_list_sort take them
    if (check_frame(f) != GS_OK || P[i] <= 0 || !geom[i] || !counters[i] || !binning[i] || !image[i] ||
        capacity[i] <= 0 || capacity[i] > 0x7fffffffLL || f->binning != GS_BINNING_DEPTH_FIRST)
      return GS_E_ARG;
    gs_geom_layout_t LG;
    gs_binning_layout_t LB;
    gs_image_layout_t LI;
    gs_geom_layout(P[i], &LG);
    if (gs_binning_layout(capacity[i], P[i], f->width, f->height, f->binning, &LB) != GS_OK) return GS_E_ARG;
    gs_image_layout(f->width, f->height, &LI);
    if (binning_bytes[i] < LB.total || image_bytes[i] < LI.total) return GS_E_WORKSPACE;
    int gx, gy;
    const int nt = grid_tiles(f->width, f->height, gx, gy);
    if (gx > 0xffff || gy > 0xffff) return GS_E_ARG;  // tile rects are packed in 16-bit fields
    const int nbp = nblocks256(P[i]);
    prefix[i] = (uint32_t)total;
    total += (unsigned long long)nbp;
    if (total > 0x7fffffffULL) return GS_E_ARG;
    const char* gb = (const char*)geom[i];
    char* bb = (char*)binning[i];
    char* ib = (char*)image[i];
    // where the depth sort (per frame or segmented) leaves its result
    const int in_b = sort_passes(32) & 1;
    DfDupArgs& d = mid[i].dup;
    d.P = P[i];
    d.gx = gx;
    d.gy = gy;
    d.capacity = capacity[i];
    d.rect = (const uint2*)(bb + LB.drect);
    d.tiles = (const uint32_t*)(gb + LG.tiles);
    d.dvals = (const uint32_t*)(bb + (in_b ? LB.dvals_b : LB.dvals_a));
    d.dscan = (const uint32_t*)(bb + LB.dscan);
    d.dblocks = (const uint32_t*)(bb + LB.dblocks);
    d.counters = counters[i];
    d.keys = (uint32_t*)(bb + LB.keys_dup);
    d.vals = (uint32_t*)(bb + LB.vals_dup);
    d.id_bits = LB.id_bits;
    mid[i].dscan = (uint32_t*)(bb + LB.dscan);
    mid[i].dblocks = (uint32_t*)(bb + LB.dblocks);
    mid[i].nblocks = nbp;
    ck[i].ranges = (uint32_t*)(ib + LI.ranges);
    ck[i].ckpt_base = (uint32_t*)(ib + LI.ckpt_base);
    ck[i].item_log = (uint32_t*)(bb + LB.item_log);
    ck[i].tile_order = (uint32_t*)(ib + LI.tile_order);
    ck[i].nt = nt;
  }
  prefix[n] = (uint32_t)total;
  *total_blocks = (unsigned)total;
  return GS_OK;
}

extern "C" int gs_multi_binning_duplicate(int n, const void* dev_set, unsigned total_blocks, void* stream) {
  if (n <= 0 || !dev_set) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t* prefix = (const uint32_t*)dev_set;
  const DfMid* mid = (const DfMid*)((const char*)dev_set + mbm_mid_offset(n));
  timing::Scope ts(timing::kDuplicate, s);
  if (total_blocks > 0) hipLaunchKernelGGL(k_df_scan_multi, dim3(total_blocks), dim3(256), 0, s, prefix, mid, n);
  hipLaunchKernelGGL(k_excl_scan_multi, dim3(n), dim3(1024), 0, s, mid);
  if (total_blocks > 0)
    hipLaunchKernelGGL(k_df_duplicate_multi, dim3(total_blocks), dim3(256), 0, s, prefix, mid, n);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_multi_binning_after_list_sort(int n, const void* dev_set, void* stream) {
  if (n <= 0 || !dev_set) return GS_E_ARG;
  hipLaunchKernelGGL(k_tile_ckpt_scan_multi, dim3(n), dim3(1024), 0, (hipStream_t)stream,
                     (const CkptScan*)((const char*)dev_set + mbm_ckpt_offset(n)));
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_render_binning_after_depth_sort(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                                  long long capacity, void* binning, size_t binning_bytes,
                                                  void* image, size_t image_bytes, void* stream) {
  return render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                        (hipStream_t)stream, kAfterSort);
}

extern "C" int gs_render_after_depth_sort(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                          long long capacity, void* binning, size_t binning_bytes, void* image,
                                          size_t image_bytes, float* out_color, float* out_depth, float* out_alpha,
                                          void* stream) {
  return render_blend_from(kAfterSort, f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                           out_color, out_depth, out_alpha, (hipStream_t)stream);
}

extern "C" size_t gs_multi_depth_sort_bytes(int n) { return seg_sort_desc_bytes(n, sort_passes(32)); }

extern "C" int gs_multi_depth_sort_build(int n, const gs_frame_t* frames, const int* P, const long long* capacity,
                                         void* const* binning, const size_t* binning_bytes,
                                         uint32_t* const* counters, void* host_set, size_t set_bytes,
                                         unsigned* blocks) {
  if (n <= 0 || !frames || !P || !capacity || !binning || !binning_bytes || !counters || !host_set || !blocks)
    return GS_E_ARG;
  if (set_bytes < gs_multi_depth_sort_bytes(n)) return GS_E_WORKSPACE;
  std::vector<SortSeg> segs((size_t)n);
  for (int i = 0; i < n; ++i) {
    const gs_frame_t* f = &frames[i];
    if (check_frame(f) != GS_OK || f->binning != GS_BINNING_DEPTH_FIRST || P[i] <= 0 || capacity[i] <= 0 ||
        capacity[i] > 0x7fffffffLL || P[i] > kSegMaxKeys || !binning[i] || !counters[i])
      return GS_E_ARG;
    gs_binning_layout_t LB;
    if (gs_binning_layout(capacity[i], P[i], f->width, f->height, f->binning, &LB) != GS_OK) return GS_E_ARG;
    if (binning_bytes[i] < LB.total) return GS_E_WORKSPACE;
    char* bb = (char*)binning[i];
    segs[i] = SortSeg{(uint32_t*)(bb + LB.dkeys_a), (uint32_t*)(bb + LB.dvals_a), (uint32_t*)(bb + LB.dkeys_b),
                      (uint32_t*)(bb + LB.dvals_b), counters[i] + 2, (long long)P[i], bb + LB.hist, 0, 32,
                      nullptr};
  }
  unsigned per[kSegMaxPasses];
  int passes = 0;
  if (!seg_sort_build(n, segs.data(), host_set, set_bytes, per, &passes)) return GS_E_ARG;
  *blocks = per[0];  // every frame runs all four passes
  return GS_OK;
}

extern "C" int gs_multi_depth_sort(int n, const void* dev_set, unsigned blocks, void* stream) {
  if (n <= 0 || !dev_set || blocks == 0) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  timing::Scope ts(timing::kSort, s);
  const unsigned per[kSegMaxPasses] = {blocks, blocks, blocks, blocks};
  static_assert(kSegMaxPasses == 4, "the depth sort runs four 8-bit passes");
  return seg_sort_launch(n, dev_set, per, true, s) == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_render_binning_duplicate(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                           long long capacity, void* binning, size_t binning_bytes, void* image,
                                           size_t image_bytes, void* stream) {
  return render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                        (hipStream_t)stream, kDup);
}

extern "C" size_t gs_multi_list_sort_bytes(int n) { return seg_sort_desc_bytes(n, kSegMaxPasses); }

extern "C" int gs_multi_list_sort_build(int n, const gs_frame_t* frames, const int* P, const long long* capacity,
                                        void* const* binning, const size_t* binning_bytes,
                                        uint32_t* const* counters, void* const* image, void* host_set,
                                        size_t set_bytes, unsigned* blocks) {
  if (n <= 0 || !frames || !P || !capacity || !binning || !binning_bytes || !counters || !image || !host_set ||
      !blocks)
    return GS_E_ARG;
  if (set_bytes < gs_multi_list_sort_bytes(n)) return GS_E_WORKSPACE;
  std::vector<SortSeg> segs((size_t)n);
  for (int i = 0; i < n; ++i) {
    const gs_frame_t* f = &frames[i];
    if (check_frame(f) != GS_OK || f->binning != GS_BINNING_DEPTH_FIRST || P[i] <= 0 || capacity[i] <= 0 ||
        capacity[i] > kSegMaxKeys || !binning[i] || !counters[i] || !image[i])
      return GS_E_ARG;
    gs_binning_layout_t LB;
    gs_image_layout_t LI;
    if (gs_binning_layout(capacity[i], P[i], f->width, f->height, f->binning, &LB) != GS_OK) return GS_E_ARG;
    if (!LB.id_bits) return GS_E_ARG;  // the shared list sort takes one-word entries only
    if (binning_bytes[i] < LB.total) return GS_E_WORKSPACE;
    gs_image_layout(f->width, f->height, &LI);
    int gx, gy;
    const int nt = grid_tiles(f->width, f->height, gx, gy);
    const int tb = bits_for((uint32_t)nt) > 0 ? bits_for((uint32_t)nt) : 1;
    char* bb = (char*)binning[i];
    segs[i] = SortSeg{(uint32_t*)(bb + LB.keys_dup), nullptr, (uint32_t*)(bb + LB.keys_alt), nullptr, counters[i],
                      capacity[i], bb + LB.hist, LB.id_bits, LB.id_bits + tb,
                      (uint32_t*)((char*)image[i] + LI.ranges)};
  }
  int passes = 0;
  return seg_sort_build(n, segs.data(), host_set, set_bytes, blocks, &passes) ? GS_OK : GS_E_ARG;
}

extern "C" int gs_multi_list_sort(int n, const void* dev_set, const unsigned* blocks, void* stream) {
  if (n <= 0 || !dev_set || !blocks || blocks[0] == 0) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  timing::Scope ts(timing::kSort, s);
  return seg_sort_launch(n, dev_set, blocks, false, s) == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_render_binning_after_list_sort(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                                 long long capacity, void* binning, size_t binning_bytes,
                                                 void* image, size_t image_bytes, void* stream) {
  return render_binning(f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                        (hipStream_t)stream, kTail);
}

extern "C" int gs_render_after_list_sort(const gs_frame_t* f, int P, const void* geom, uint32_t* counters,
                                         long long capacity, void* binning, size_t binning_bytes, void* image,
                                         size_t image_bytes, float* out_color, float* out_depth, float* out_alpha,
                                         void* stream) {
  return render_blend_from(kTail, f, P, geom, counters, capacity, binning, binning_bytes, image, image_bytes,
                           out_color, out_depth, out_alpha, (hipStream_t)stream);
}

// ---- many frames in one blend launch --------------------------------------
// Set layout: uint32 prefix[n + 1] of tiles per frame, then BlendArgs[n] at a
// 256-byte boundary.
static size_t multi_blend_args_offset(int n) { return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256); }

extern "C" size_t gs_multi_blend_bytes(int n) {
  return n <= 0 ? 0 : multi_blend_args_offset(n) + sizeof(BlendArgs) * (size_t)n;
}

extern "C" int gs_multi_blend_build(int n, const gs_frame_t* frames, const int* P, const void* const* geom,
                                    const long long* capacity, void* const* binning, void* const* image,
                                    float* const* out_color, float* const* out_depth, float* const* out_alpha,
                                    void* host_set, size_t set_bytes, unsigned* total_tiles) {
  if (n <= 0 || !frames || !P || !geom || !capacity || !binning || !image || !out_color || !out_depth || !host_set ||
      !total_tiles)
    return GS_E_ARG;
  if (set_bytes < gs_multi_blend_bytes(n)) return GS_E_WORKSPACE;
  uint32_t* prefix = (uint32_t*)host_set;
  BlendArgs* args = (BlendArgs*)((char*)host_set + multi_blend_args_offset(n));
  unsigned long long total = 0;
  for (int i = 0; i < n; ++i) {
    if (check_frame(&frames[i]) != GS_OK || P[i] < 0 || !image[i] || !out_color[i] || !out_depth[i]) return GS_E_ARG;
    if (capacity[i] < 0 || capacity[i] > 0x7fffffffLL || (P[i] > 0 && !geom[i]) || (capacity[i] > 0 && !binning[i]))
      return GS_E_ARG;
    int gx, gy;
    const int nt = grid_tiles(frames[i].width, frames[i].height, gx, gy);
    prefix[i] = (uint32_t)total;
    total += (unsigned long long)nt;
    if (total > 0x7fffffffULL) return GS_E_ARG;
    args[i] = make_blend_args(&frames[i], P[i], geom[i], capacity[i], binning[i], image[i], out_color[i],
                              out_depth[i], out_alpha ? out_alpha[i] : nullptr);
  }
  prefix[n] = (uint32_t)total;
  *total_tiles = (unsigned)total;
  return GS_OK;
}

extern "C" int gs_multi_blend(int n, const void* dev_set, unsigned total_tiles, void* stream) {
  if (n <= 0 || !dev_set || total_tiles == 0) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t* prefix = (const uint32_t*)dev_set;
  const BlendArgs* args = (const BlendArgs*)((const char*)dev_set + multi_blend_args_offset(n));
  {
    timing::Scope ts(timing::kBlend, s);
    hipLaunchKernelGGL(k_blend_multi, dim3(total_tiles), dim3(256), 0, s, prefix, args, n);
  }
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_mark_visible(const float* means3D, int P, const float* viewmatrix, const float* projmatrix,
                               uint8_t* visible, void* stream) {
  (void)projmatrix;
  if (P < 0 || (P > 0 && (!means3D || !viewmatrix || !visible))) return GS_E_ARG;
  if (P == 0) return GS_OK;
  hipLaunchKernelGGL(k_mark_visible, dim3(nblocks256(P)), dim3(256), 0, (hipStream_t)stream, means3D, P,
                     viewmatrix, visible);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" const char* gs_error_string(int code) {
  switch (code) {
    case GS_OK: return "ok";
    case GS_E_ARG: return "invalid argument";
    case GS_E_WORKSPACE: return "workspace too small";
    case GS_E_LAUNCH: return "kernel launch failed";
    case GS_E_STATIC_LIMIT: return "input larger than the entry point's static scratch allows";
    case GS_E_SCRATCH: return "static scratch address lookup failed";
    default: return "unknown error";
  }
}

extern "C" int gs_abi_version(void) { return 4; }
