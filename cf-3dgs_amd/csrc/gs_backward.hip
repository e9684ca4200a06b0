// gs_backward.hip -- backward rasteriser for gfx950.
//
//   k_bwd_prepare    one launch, two jobs: workgroup 0 builds the work items
//                    (256-entry chunks of each tile's list up to the tile's
//                    largest contributor count, longest first, from the
//                    forward's item log); the others clear the per-Gaussian
//                    render-gradient accumulators (a kernel, not a memset, so
//                    the sequence can be captured in a graph)
//   k_render_bwd     one chunk per workgroup (16x16 pixels), back to front.
//                    A chunk that is not the tile's last starts from the
//                    forward's checkpoint at its end: T_b, and the suffix sums
//                    (C_final - C_b) / T_b that the back-to-front recurrence
//                    would have accumulated, so heavy tiles spread over many
//                    workgroups instead of serialising in one.  T is then
//                    recovered by division as in the forward's algebra.  The
//                    10 per-(pixel, Gaussian) terms are summed over the
//                    wave with permlane swaps + DPP (wave_sum12), the wave
//                    sums go to LDS accumulators, and each Gaussian of the
//                    chunk receives one global atomic per term.
//                    Deterministic mode: each wave keeps its own LDS
//                    partials (summed in wave order), and the chunk writes one
//                    10-float partial per (Gaussian, tile) pair into a slot
//                    of its own (the Gaussian's offset in the Gaussian-order
//                    tile prefix + the tile's rank in its rect) -- no atomics;
//   k_det_gather     deterministic mode: per Gaussian, its partials summed in
//                    rect order (the pose-only backward sums its 5-term
//                    partials the same way inside k_preprocess_bwd);
//   k_preprocess_bwd per Gaussian: conic -> 2-D cov -> 3-D cov / mean,
//                    projection, depth, SH direction; scale/rotation; the
//                    rigid pose gradient is reduced per wave (DPP) and each
//                    wave stores its 12 partial sums;
//   k_pose_finalize  sums the wave partials in a fixed tree order, then
//                    d/dR, d/dt -> d/d(q, t) through the quaternion
//                    normalisation.
#include <cmath>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {

// render-gradient record: m2x m2y ca cb cc opac r g b depth - -, where ca and
// cc are the conic gradients times -2: their factor -0.5 is applied to the
// Gaussian's sum instead of to every pixel's term (a power of two, so the
// sums and their rounding are exactly the same)
constexpr int kRG = 12;

struct RBArgs {
  int W, H, gx, gy, nt;
  const uint32_t* ranges;
  const uint32_t* list;  // sorted entries; the Gaussian id is entry & id_mask
  uint32_t id_mask;
  const float4* splat;
  const float* bg;
  const float* final_T;
  const uint32_t* n_contrib;
  const float* accum;         // [4][H*W] colour sums before the background, depth
  const float* ckpt;          // [..][5][256] forward state at batch boundaries
  const uint32_t* ckpt_base;  // [nt + 1]
  const uint32_t* tile_maxc;  // [nt]
  const uint32_t* items;      // [0] count, then (tile, chunk) pairs, longest estimated first
  const uint32_t* item_log;   // the blend's item records: [0] count, then (tile, chunk << 6 | class)
  const float* dL_dcolor;
  const float* dL_ddepth;
  const float* dL_dalpha;
  float* g_render;
  // deterministic mode
  const uint32_t* local_incl;  // [P] tiles prefix within each 256-Gaussian block
  const uint32_t* block_excl;  // [ceil(P/256)] exclusive block prefix
  float* part;                 // [capacity][10] partial per (Gaussian, tile) pair
};

// One workgroup orders the backward's work items.  A tile's items are its
// 256-entry chunks up to its last contributor; the forward's blend logged
// them with a cost class from the longest per-wave list it had in the same
// batch, which the render backward's time per chunk follows closely.  Items
// are handed out longest first (longest-processing-time order): a large frame
// has more items than fit on the GPU at once, and an item that starts late
// and runs long keeps the launch going after the rest have drained.  This is
// a counting sort on the 64 classes.  The order within a class is whatever
// the LDS atomics give: no result depends on the order of items
// (deterministic mode writes one slot per pair; the atomic mode is
// order-dependent in any case).
__device__ __forceinline__ void bwd_items(const uint32_t* item_log, uint32_t* items) {
  __shared__ uint32_t s_cls[64];
  const int t = threadIdx.x;
  const uint32_t n = item_log[0];
  if (t < 64) s_cls[t] = 0;
  __syncthreads();
  for (uint32_t i = t; i < n; i += 1024) atomicAdd(&s_cls[item_log[2 + 2 * i] & 63u], 1u);
  __syncthreads();
  if (t < 64) {
    // exclusive scan of the 64 class counts within wave 0
    const uint32_t v = s_cls[t];
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t up = __shfl_up(incl, off, 64);
      if (t >= off) incl += up;
    }
    s_cls[t] = incl - v;
  }
  __syncthreads();
  for (uint32_t i = t; i < n; i += 1024) {
    const uint32_t tile = item_log[1 + 2 * i], cc = item_log[2 + 2 * i];
    const uint32_t pos = atomicAdd(&s_cls[cc & 63u], 1u);
    items[1 + 2 * pos] = tile;
    items[2 + 2 * pos] = cc >> 6;
  }
  if (t == 0) items[0] = n;
}

// The backward's set-up in one launch: workgroup 0 builds the work-item table
// (bwd_items; none when items is null), the others clear the zn accumulator
// floats at zp.
// The set-up of one frame: zn floats at zp to clear, or, when terms > 0
// (the deterministic mode's pair partials), only the first num_rendered *
// terms of them (counters[0] = num_rendered; no pair slot at or past it is
// written or read).
struct BwdPrep {
  float* zp;
  long long zn;
  const uint32_t* item_log;
  uint32_t* items;
  const uint32_t* counters;
  int terms;
};

// Workgroup bid of a set-up over nblocks workgroups.
__device__ __forceinline__ void bwd_prepare_block(const BwdPrep& p, int bid, int nblocks) {
  const int first = p.items ? 1 : 0;
  if (first && bid == 0) {
    bwd_items(p.item_log, p.items);
    return;
  }
  long long zn = p.zn;
  if (p.terms > 0) zn = min(zn, (long long)p.counters[0] * p.terms);
  const long long nb = (long long)nblocks - first;
  for (long long k = ((long long)bid - first) * 1024 + threadIdx.x; k < zn; k += nb * 1024) p.zp[k] = 0.0f;
}

__global__ __launch_bounds__(1024) void k_bwd_prepare(BwdPrep p) { bwd_prepare_block(p, blockIdx.x, gridDim.x); }

// Many frames in one launch: frame f owns workgroups [prefix[f], prefix[f +
// 1]), as many as its own launch would have.
__global__ __launch_bounds__(1024) void k_bwd_prepare_multi(const uint32_t* prefix, const BwdPrep* prep, int n) {
  const uint32_t b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const BwdPrep p = prep[lo];
  bwd_prepare_block(p, (int)(b - prefix[lo]), (int)(prefix[lo + 1] - prefix[lo]));
}

// Sum over the 16 lanes of each DPP row; every lane of the row gets the sum.
__device__ __forceinline__ float row_sum16(float v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x140>(v);  // row_mirror
  return v;
}

// Wave-wide sums of 12 values at once (reduce-scatter).  Stage 1 swaps the
// upper half-wave of v[2i] with the lower half of v[2i+1] (v_permlane32_swap)
// and adds: value 2i now sits, half-summed, in lanes 0-31 and 2i+1 in lanes
// 32-63.  Stage 2 does the same with 16-lane rows (v_permlane16_swap) on
// pairs of those, leaving four values per register, one per row, each lane
// holding the sum of its 4 lanes at stride 16.  Stage 3 sums each 16-lane row
// (DPP).  Afterwards every lane of row q of w[m] holds the wave sum of value
// 4m + kRowValue[q], kRowValue = {0, 2, 1, 3}.  The order of additions is
// fixed, so the sums are bitwise reproducible.  30 VALU operations for 12
// values, against 4 DPP steps per value for separate row sums.
__device__ __forceinline__ void swap_add32(float a, float b, float& out) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  out = __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_add16(float a, float b, float& out) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  out = __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ void wave_sum12(const float v[12], float w[3]) {
  float u[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) swap_add32(v[2 * i], v[2 * i + 1], u[i]);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    swap_add16(u[2 * m], u[2 * m + 1], w[m]);
    w[m] = row_sum16(w[m]);
  }
}

// wave_sum12 of v[0..9] with v[10] = v[11] = 0 (the render backward's 10
// terms): the last stage-1 sum is the literal 0, not a swap of two zeros.
__device__ __forceinline__ void wave_sum10(const float v[10], float w[3]) {
  float u[6];
#pragma unroll
  for (int i = 0; i < 5; ++i) swap_add32(v[2 * i], v[2 * i + 1], u[i]);
  u[5] = 0.0f;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    swap_add16(u[2 * m], u[2 * m + 1], w[m]);
    w[m] = row_sum16(w[m]);
  }
  // keep the last row step with the reduction: the sums are consumed only
  // by the storing lanes, and an add sunk into that branch would leave its
  // DPP permutation as a separate move
  asm volatile("" : : "v"(w[0]), "v"(w[1]), "v"(w[2]));
}

// DEP: the frame has a depth or an alpha cotangent.  Without one (a colour
// loss, as in the training step and the pose solve) the depth and alpha
// terms, their running sums and the depth gradient are identically zero and
// are not computed.
// wave_sum10 restricted to v[0..4] (the pose-only backward's terms): each of
// these five sums is formed by the same operations, in the same lanes, as in
// wave_sum10, so the two give the same bits.  w[0] rows hold values
// {0, 2, 1, 3}, row 0 of w[1] value 4.
__device__ __forceinline__ void wave_sum5(const float v[5], float w[2]) {
  float u[3];
  swap_add32(v[0], v[1], u[0]);
  swap_add32(v[2], v[3], u[1]);
  swap_add32(v[4], 0.0f, u[2]);
  swap_add16(u[0], u[1], w[0]);
  w[0] = row_sum16(w[0]);
  swap_add16(u[2], 0.0f, w[1]);
  w[1] = row_sum16(w[1]);
  asm volatile("" : : "v"(w[0]), "v"(w[1]));  // as in wave_sum10
}

template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

template <bool DET, bool DEP, bool PO>
__device__ __forceinline__ void render_bwd_item(const RBArgs& a, uint32_t item) {
  // PO (pose-only, fixed-order, colour cotangent only): the caller wants only
  // the pose gradient, which at active SH degree 0 depends on a splat's
  // screen position and conic alone, so only those five terms are summed and
  // stored per (Gaussian, tile) pair.  Their arithmetic is the full path's,
  // term for term, so the pose gradient has the same bits either way.
  static_assert(!PO || (DET && !DEP), "the pose-only backward is fixed-order and colour-only");
  constexpr int kTerms = PO ? 5 : 10;
  constexpr int kParts = DET ? 4 : 1;  // LDS partial sets: one per wave when deterministic
  // The item's 256-entry chunk is staged 128 entries at a time (back half
  // first): half the LDS per workgroup, so twice as many workgroups fit on a
  // CU (deterministic mode: 5 instead of 2).  The pixel state carries over.
  constexpr int kSub = 128;
  __shared__ uint32_t s_id[kSub];
  __shared__ float4 s_a[kSub];
  __shared__ float4 s_b[kSub];
  __shared__ float4 s_c[kSub];
  __shared__ uint32_t s_mask[kSub];    // quadrants each staged splat can reach (splat_quads)
  __shared__ uint8_t s_list[4][kSub];  // per wave: the staged splats that reach its pixels
  __shared__ float s_acc[kTerms * kParts][kSub];
  if (item >= a.items[0]) return;
  const int tile = (int)a.items[1 + 2 * item];
  const uint32_t chunk = a.items[2 + 2 * item];
  const uint32_t maxc = a.tile_maxc[tile];
  const uint32_t nchunk = (maxc + 255) / 256;
  const uint32_t cstart = chunk * 256;
  const uint32_t cend = min(cstart + 256, maxc);
  const int tx = tile % a.gx, ty = tile / a.gx;
  int pdx, pdy;
  tile_pixel(threadIdx.x, pdx, pdy);
  const int px = tx * kBX + pdx, py = ty * kBY + pdy;
  const bool inside = px < a.W && py < a.H;
  const int pix = py * a.W + px;
  const int HW = a.W * a.H;
  const float pfx = (float)px, pfy = (float)py;
  const uint32_t r0 = a.ranges[2 * tile];
  const float T_fin = inside ? a.final_T[pix] : 0.0f;
  const uint32_t last = inside ? a.n_contrib[pix] : 0u;
  float dpix0 = 0.f, dpix1 = 0.f, dpix2 = 0.f, dD = 0.f, dA = 0.f;
  // per-pixel state at position cend (the end of this chunk)
  float T = T_fin;
  float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc_d = 0.f, acc_a = 0.f;
  if (inside && last > cstart) {
    dpix0 = a.dL_dcolor[pix];
    dpix1 = a.dL_dcolor[HW + pix];
    dpix2 = a.dL_dcolor[2 * HW + pix];
    if constexpr (DEP) {
      if (a.dL_ddepth) dD = a.dL_ddepth[pix];
      if (a.dL_dalpha) dA = a.dL_dalpha[pix];
    }
    if (chunk + 1 < nchunk) {
      // suffix sums from the forward's state at the boundary: sum_{k >= b} c_k a_k T_k = C_final - C_b
      const float* ck = a.ckpt + (size_t)(a.ckpt_base[tile] + chunk) * 5 * 256 + threadIdx.x;
      const float Tb = ck[0];
      const float inv = 1.0f / Tb;
      T = Tb;
      acc0 = (a.accum[pix] - ck[256]) * inv;
      acc1 = (a.accum[HW + pix] - ck[512]) * inv;
      acc2 = (a.accum[2 * HW + pix] - ck[768]) * inv;
      if constexpr (DEP) {
        acc_d = (a.accum[3 * HW + pix] - ck[1024]) * inv;
        acc_a = 1.0f - T_fin * inv;
      }
    }
  }
  const float bg_dot = (a.bg[0] * dpix0 + a.bg[1] * dpix1) + a.bg[2] * dpix2;
  const bool bg_black = a.bg[0] == 0.0f && a.bg[1] == 0.0f && a.bg[2] == 0.0f;
  const float ddelx = 0.5f * (float)a.W, ddely = 0.5f * (float)a.H;
  const float a_min = 1.0f / 255.0f;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  // entries at or past the last contributor of every pixel of the wave are
  // inactive on every lane
  const uint32_t wlast = wave_max_u32(last);
#pragma unroll 1
  for (int sb = 1; sb >= 0; --sb) {
    const uint32_t sbase = cstart + (uint32_t)(sb * kSub);
    if (sbase >= cend) continue;  // (uniform) a chunk of at most 128 entries
    const uint32_t send = min(sbase + (uint32_t)kSub, cend);
    __syncthreads();  // the previous half's staged data has been flushed
    if (threadIdx.x < kSub) {
      const uint32_t q = sbase + threadIdx.x;
      uint32_t m = 0u;
      if (q < send) {
        const uint32_t g = a.list[r0 + q] & a.id_mask;
        s_id[threadIdx.x] = g;
        const float4 sa = a.splat[3 * g];
        s_a[threadIdx.x] = sa;
        float4 b = a.splat[3 * g + 1];
        b.w = skip_threshold(b.y);
        s_b[threadIdx.x] = b;
        s_c[threadIdx.x] = a.splat[3 * g + 2];
        m = splat_quads(sa.x, sa.y, sa.z, sa.w, b.x, b.w, (float)(tx * kBX), (float)(ty * kBY));
      }
      s_mask[threadIdx.x] = m;
    }
    for (int k = threadIdx.x; k < kTerms * kParts * kSub; k += 256) (&s_acc[0][0])[k] = 0.0f;
    __syncthreads();
    // the wave's own list: splats that miss its 8 x 8 pixels are ones no lane would use
    const int lim = wlast > sbase ? (int)min(wlast - sbase, (uint32_t)kSub) : 0;
    // (a sum of ballot counts, so wave-uniform; read from the first lane so
    // that the pair loop below counts in a scalar register)
    const int nw = __builtin_amdgcn_readfirstlane(compact_wave_list<kSub / 64>(s_mask, s_list[wv], wv, lane, lim));
    // Branch-free per pair: masked lanes cost a SIMD as much as active ones,
    // so every lane evaluates the pair; the only branches are wave-uniform
    // skips when no lane is active.  An inactive lane runs with alpha = G = 0,
    // which leaves its state unchanged (T / (1 - 0) = T, 0 c + 1 acc = acc) and
    // makes each of its gradient terms 0, so no per-lane selects are needed.
    // The running sums acc = sum_{k > current} c_k alpha_k T_k / T_{current+1}
    // are advanced right after each splat (the reference advances them
    // lazily at the next one, from last_alpha / last_color; same expression,
    // same rounding).
    // The pair loop in two forms: a black background (the reference's
    // default) adds nothing to dL/dalpha -- its term is a signed zero -- so
    // that form leaves the term out (the choice is wave-uniform).
    auto pair_loop = [&](auto bg_on) {
      for (int i = nw - 1; i >= 0; --i) {
        // Gradient arithmetic may fuse multiply-adds (the library is built
        // without contraction for the forward's bit-exactness).  What decides
        // which pairs contribute -- blend_power, spec_expf and the single
        // product alpha = o G -- is outside this scope or has no add, so it
        // stays exactly the forward's.
  #pragma clang fp contract(fast)
        const int j = s_list[wv][i];
        const uint32_t q = sbase + j;  // 0-based list position; contributor index q + 1
        const float4 ga = s_a[j];
        const float4 gb = s_b[j];
        const float dx = ga.x - pfx, dy = ga.y - pfy;
        const float power = blend_power(ga.z, ga.w, gb.x, dx, dy);
        // gb.w: alpha < 1/255 for sure below it.  (q < last implies inside:
        // last is 0 outside the image.)  The wave-uniform skips test the
        // product of one ballot per compare, which is the ballot of the
        // conjunction: a ballot of a single compare is that compare's lane mask.
        const bool in_list = q < last, p_le0 = !(power > 0.0f), p_ge = !(power < gb.w);
        uint64_t live = __builtin_amdgcn_ballot_w64(in_list) & __builtin_amdgcn_ballot_w64(p_le0) &
                        __builtin_amdgcn_ballot_w64(p_ge);
        if (live == 0ull) continue;
        const float Graw = spec_expf(power);
        const float alpha_raw = fminf(0.99f, gb.y * Graw);
        const bool a_ok = !(alpha_raw < a_min);
        live &= __builtin_amdgcn_ballot_w64(a_ok);
        if (live == 0ull) continue;
        const bool active = in_list && p_le0 && p_ge && a_ok;
        const float G = active ? Graw : 0.0f;
        const float alpha = active ? alpha_raw : 0.0f;
        const float4 gc = s_c[j];
        // T before this splat; the gradient tolerates the 1-ulp reciprocal
        const float inv_1ma = __builtin_amdgcn_rcpf(1.0f - alpha);
        const float Tn = T * inv_1ma;
        const float w = alpha * Tn;
        float dL_da = (gc.x - acc0) * dpix0;
        dL_da += (gc.y - acc1) * dpix1;
        dL_da += (gc.z - acc2) * dpix2;
        if constexpr (DEP) {
          dL_da += (gb.z - acc_d) * dD;
          dL_da += (1.0f - acc_a) * dA;
        }
        dL_da *= Tn;
        if constexpr (decltype(bg_on)::value) dL_da += (-T_fin * inv_1ma) * bg_dot;
        const float dL_dG = gb.y * dL_da;
        const float gdx = G * dx, gdy = G * dy;
        const float dG_dx = -gdx * ga.z - gdy * ga.w;
        const float dG_dy = -gdy * gb.x - gdx * ga.w;
        float v[10];
        v[0] = dL_dG * dG_dx * ddelx;
        v[1] = dL_dG * dG_dy * ddely;
        v[2] = gdx * dx * dL_dG;  // times -0.5 in k_preprocess_bwd
        v[3] = -gdx * dy * dL_dG;
        v[4] = gdy * dy * dL_dG;  // times -0.5 in k_preprocess_bwd
        if constexpr (!PO) {
          v[5] = G * dL_da;
          v[6] = w * dpix0;
          v[7] = w * dpix1;
          v[8] = w * dpix2;
          v[9] = DEP ? w * dD : 0.0f;
        }
        T = Tn;
        const float oma = 1.0f - alpha;
        acc0 = alpha * gc.x + oma * acc0;
        acc1 = alpha * gc.y + oma * acc1;
        acc2 = alpha * gc.z + oma * acc2;
        if constexpr (DEP) {
          acc_d = alpha * gb.z + oma * acc_d;
          acc_a = alpha + oma * acc_a;
        }
        if constexpr (PO) {
          float w2[2];
          wave_sum5(v, w2);
          if ((lane & 15) == 0) {
            const int q = lane >> 4;
            const int o = q == 1 ? 2 : (q == 2 ? 1 : q);
            s_acc[wv * kTerms + o][j] = w2[0];
            if (q == 0) s_acc[wv * kTerms + 4][j] = w2[1];
          }
        } else {
          float w[3];
          wave_sum10(v, w);
          if ((lane & 15) == 0) {
            const int q = lane >> 4;
            const int o = q == 1 ? 2 : (q == 2 ? 1 : q);  // value of row q: 4m + {0, 2, 1, 3}[q]
  #pragma unroll
            for (int m = 0; m < 3; ++m) {
              const int k = 4 * m + o;
              if (k < (DEP ? 10 : 9)) {
                if constexpr (DET) {
                  s_acc[wv * 10 + k][j] = w[m];
                } else {
                  atomicAdd(&s_acc[k][j], w[m]);  // (a zero adds nothing; testing for it costs more)
                }
              }
            }
          }
        }
      }
    };
    if (bg_black) pair_loop(BoolTag<false>{});
    else pair_loop(BoolTag<true>{});
    __syncthreads();
    const uint32_t t = threadIdx.x;
    if (t < send - sbase) {
      if constexpr (DET) {
        // slot of this (Gaussian, tile) pair: the Gaussian's first slot in the
        // Gaussian-order prefix of tiles touched + the tile's row-major rank in
        // its rect (every pair of the list has exactly one chunk)
        const uint32_t g = s_id[t];
        int x0, y0, x1, y1;
        tile_rect(s_a[t].x, s_a[t].y, __float_as_int(s_c[t].w), a.gx, a.gy, x0, y0, x1, y1);
        const uint32_t w = (uint32_t)(x1 - x0);
        const uint32_t ntg = w * (uint32_t)(y1 - y0);
        const uint32_t slot =
            a.block_excl[g >> 8] + a.local_incl[g] - ntg + (uint32_t)(ty - y0) * w + (uint32_t)(tx - x0);
        float* dst = a.part + (size_t)slot * kTerms;
#pragma unroll
        for (int k = 0; k < kTerms; ++k)
          dst[k] = ((s_acc[k][t] + s_acc[kTerms + k][t]) + s_acc[2 * kTerms + k][t]) + s_acc[3 * kTerms + k][t];
      } else {
        float* dst = a.g_render + (size_t)s_id[t] * kRG;
#pragma unroll
        for (int k = 0; k < 10; ++k) {
          const float s = s_acc[k][t];
          if (s != 0.0f) unsafeAtomicAdd(dst + k, s);
        }
      }
    }
  }
}

// The grid is an upper bound on the item count, so items are taken in
// dispatch order: live items then go round-robin over all 8 XCDs (a
// contiguous-per-XCD remap would put them all on the first few).
template <bool DET, bool DEP, bool PO>
__global__ __launch_bounds__(256) void k_render_bwd(RBArgs a) {
  render_bwd_item<DET, DEP, PO>(a, blockIdx.x);
}

// Many frames in one launch: frame f owns blocks [prefix[f], prefix[f + 1]),
// an upper bound on its work items.
template <bool DET, bool DEP, bool PO>
__global__ __launch_bounds__(256) void k_render_bwd_multi(const uint32_t* prefix, const RBArgs* args, int n) {
  const uint32_t b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const RBArgs a = args[lo];
  render_bwd_item<DET, DEP, PO>(a, b - prefix[lo]);
}

// Deterministic mode: per Gaussian, the partials of its (Gaussian, tile)
// pairs (NT terms each) summed in rect order into its render-gradient
// record.  Nothing is read when the frame
// overflowed the binning capacity (no pairs were made).
template <int NT>
__global__ __launch_bounds__(256) void k_det_gather(int P, const uint32_t* tiles, const uint32_t* local_incl,
                                                    const uint32_t* block_excl, const uint32_t* counters,
                                                    long long capacity, const float* part, float* g_render) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= P) return;
  float acc[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) acc[k] = 0.f;
  const uint32_t nt = tiles[g];
  if ((long long)counters[0] <= capacity && nt > 0) {
    const uint32_t off = block_excl[g >> 8] + local_incl[g] - nt;
    for (uint32_t e = off; e < off + nt; ++e) {
      const float* src = part + (size_t)e * NT;
#pragma unroll
      for (int k = 0; k < NT; ++k) acc[k] += src[k];
    }
  }
  float4* dst = reinterpret_cast<float4*>(g_render + (size_t)g * kRG);
  dst[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
  dst[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
  dst[2] = make_float4(acc[8], acc[9], 0.f, 0.f);
}

struct PBArgs {
  int P, M, D, W, H;
  float tanfovx, tanfovy, fx, fy, scale_mod;
  const float* view;
  const float* proj;
  const float* campos;
  const float* pose;
  const float* means;
  const float* shs;
  const float* scales;
  const float* rots;
  const float* cov_pre;
  const int* radii;
  const float4* splat;
  const float* cov3D;
  const float* g_render;
  float* dL_dmeans2D;
  float* dL_dcolors;
  float* dL_dopacity;
  float* dL_dmeans3D;
  float* dL_dcov3D;
  float* dL_dsh;
  float* dL_dscales;
  float* dL_drots;
  float* pose_part;  // [4 * gridDim.x][12] per-wave sums: dL/dt (3), dL/dR row-major (9)
  // pose-only backward: the render gradient comes straight from the fixed-order
  // (Gaussian, tile) partials (5 terms each), summed here in rect order
  // instead of by a separate gather launch; g_render is then not read
  const float* part5;
  const uint32_t* g_tiles;
  const uint32_t* local_incl;
  const uint32_t* block_excl;
  const uint32_t* counters;
  long long capacity;
  // gs_backward_adam: one Adam step of the means and SH coefficients instead
  // of writing their gradients (adam_on == 0 otherwise)
  int adam_on;
  float* means_p;
  float* means_m;
  float* means_v;
  float* shs_p;
  float* shs_m;
  float* shs_v;
  float means_step, shs_step, shs_lead_step;
  int shs_lead;
  float b1, b2, eps, inv_sqrt_bc2;
};

// The fused Adam step of one Gaussian's mean (gradient gp); all nine loads
// before the first store.
__device__ __forceinline__ void means_adam(const PBArgs& a, int i, const float gp[3]) {
  float p[3], m[3], v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t e = 3 * (size_t)i + k;
    p[k] = a.means_p[e];
    m[k] = a.means_m[e];
    v[k] = a.means_v[e];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t e = 3 * (size_t)i + k;
    adam_one(p[k], gp[k], m[k], v[k], a.b1, a.b2, a.eps, a.inv_sqrt_bc2, a.means_step);
    a.means_p[e] = p[k];
    a.means_m[e] = m[k];
    a.means_v[e] = v[k];
  }
}

// The fused Adam step's moments for the workgroup's SH rows [base, base +
// nloc): the gradients are in LDS (s_g, rows of ldsrow >= row floats; ldsrow
// == row unless row is a whole number of 16-byte vectors) and each is
// replaced there by its element's decrement step * m / denom; the moments are
// in global memory and the parameters are not touched.  Element r of a row
// steps with shs_lead_step when r < shs_lead.  Four float4 of each moment
// array per thread are loaded before any is updated.
__device__ __forceinline__ void block_adam_moments(const PBArgs& a, float* s_g, int base, int nloc, int row,
                                                   int ldsrow) {
  const size_t off = (size_t)base * row;
  float* M = a.shs_m + off;
  float* V = a.shs_v + off;
  const int total = nloc * row;
  auto upd4 = [&](float4& g, float4& m, float4& v, int e) {
    const int r = e % row;
    const float s0 = r < a.shs_lead ? a.shs_lead_step : a.shs_step;
    const float s1 = r + 1 < a.shs_lead ? a.shs_lead_step : a.shs_step;
    const float s2 = r + 2 < a.shs_lead ? a.shs_lead_step : a.shs_step;
    const float s3 = r + 3 < a.shs_lead ? a.shs_lead_step : a.shs_step;
    g.x = adam_delta(g.x, m.x, v.x, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s0);
    g.y = adam_delta(g.y, m.y, v.y, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s1);
    g.z = adam_delta(g.z, m.z, v.z, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s2);
    g.w = adam_delta(g.w, m.w, v.w, a.b1, a.b2, a.eps, a.inv_sqrt_bc2, s3);
  };
  if ((row & 3) == 0) {
    const int n4 = total >> 2, per4 = row >> 2, lper4 = ldsrow >> 2;
    float4* g4 = reinterpret_cast<float4*>(s_g);
    auto lds4 = [&](int k) { return (k / per4) * lper4 + k % per4; };  // vector k of the rows, in LDS
    int k = threadIdx.x;
    for (; k + 768 < n4; k += 1024) {
      float4 m[4], v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = 4 * (k + 256 * u);
        m[u] = ld_nt(M + q);
        v[u] = ld_nt(V + q);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = 4 * (k + 256 * u);
        const int l = lds4(k + 256 * u);
        float4 g = g4[l];
        upd4(g, m[u], v[u], q);
        g4[l] = g;
        st_nt(M + q, m[u]);
        st_nt(V + q, v[u]);
      }
    }
    for (; k < n4; k += 256) {
      float4 m = ld_nt(M + 4 * k), v = ld_nt(V + 4 * k);
      const int l = lds4(k);
      float4 g = g4[l];
      upd4(g, m, v, 4 * k);
      g4[l] = g;
      st_nt(M + 4 * k, m);
      st_nt(V + 4 * k, v);
    }
  } else {
    for (int k = threadIdx.x; k < total; k += 256) {
      float m = M[k], v = V[k];
      s_g[k] = adam_delta(s_g[k], m, v, a.b1, a.b2, a.eps, a.inv_sqrt_bc2,
                          k % row < a.shs_lead ? a.shs_lead_step : a.shs_step);
      M[k] = m;
      V[k] = v;
    }
  }
}

// Workgroup bid of one frame's per-Gaussian backward: Gaussians [256 bid,
// 256 bid + 256).  s_sh holds the [256][3M] SH rows of this workgroup:
// coefficients are read from and their gradients written back to the
// thread's own row, then the whole block is stored with coalesced 16-byte
// writes (M <= 16; larger M goes direct)
__device__ __forceinline__ void preprocess_bwd_block(const PBArgs& a, int bid, float* s_sh) {
  const int i = bid * 256 + threadIdx.x;
  const int base = bid * 256;
  const int nloc = min(256, a.P - base);
  const bool sh_grad = a.shs && (a.dL_dsh || a.adam_on);
  const bool staged = sh_grad && a.M <= 16;
  const int row = 3 * a.M;
  // LDS row stride: a row of whole 16-byte vectors gets one more vector, so
  // a thread's own-row accesses, at a stride of an odd number of vectors,
  // spread a wave over the banks (a 48-word stride puts 16 lanes on a bank)
  const int ldsrow = (row & 3) == 0 ? row + 4 : row;
  // With the fused Adam the thread's SH gradient row (b[k] * gc[c] for the
  // nc coefficients of degree D, zero beyond) stays in registers until the
  // parameter row it is computed from has been read back out of LDS, so the
  // parameters are read from global memory once.
  const bool swap_rows = staged && a.adam_on;
  float sb[16], sgc[3] = {0.f, 0.f, 0.f};
  int row_nc = 0;  // coefficients whose gradient row was set (0: none)
#pragma unroll
  for (int k = 0; k < 16; ++k) sb[k] = 0.f;
  if (staged) {
    const float* src = a.shs + (size_t)base * row;
    const int total = nloc * row;
    if ((a.M & 3) == 0) {
      rows_to_lds(s_sh, reinterpret_cast<const float4*>(src), nloc, row >> 2, ldsrow);
    } else {
      for (int k = threadIdx.x; k < total; k += 256) s_sh[k] = src[k];
    }
  }
  __syncthreads();
  float pacc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) pacc[k] = 0.f;
  Pose ps;
  if (a.pose) make_pose(a.pose, ps);
  if (i < a.P) {
    const bool vis = a.radii[i] > 0;
    float4 rA, rB, rC;  // m2x m2y ca cb | cc opac r g | b depth - -
    if (a.part5) {
      // the sums k_det_gather<5> forms, in its order (terms 5..9 are zero)
      float acc[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      const uint32_t nt = a.g_tiles[i];
      if ((long long)a.counters[0] <= a.capacity && nt > 0) {
        const uint32_t off = a.block_excl[i >> 8] + a.local_incl[i] - nt;
        for (uint32_t e = off; e < off + nt; ++e) {
          const float* src = a.part5 + (size_t)e * 5;
#pragma unroll
          for (int k = 0; k < 5; ++k) acc[k] += src[k];
        }
      }
      rA = make_float4(acc[0], acc[1], acc[2], acc[3]);
      rB = make_float4(acc[4], 0.f, 0.f, 0.f);
      rC = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      const float* rg = a.g_render + (size_t)i * kRG;
      rA = reinterpret_cast<const float4*>(rg)[0];
      rB = reinterpret_cast<const float4*>(rg)[1];
      rC = reinterpret_cast<const float4*>(rg)[2];
    }
    // (every per-Gaussian output is optional: a pose-only solve asks for none)
    if (a.dL_dmeans2D) {
      a.dL_dmeans2D[3 * i] = vis ? rA.x : 0.f;
      a.dL_dmeans2D[3 * i + 1] = vis ? rA.y : 0.f;
      a.dL_dmeans2D[3 * i + 2] = 0.f;
    }
    if (a.dL_dopacity) a.dL_dopacity[i] = vis ? rB.y : 0.f;
    if (a.dL_dcolors) {
      a.dL_dcolors[3 * i] = vis ? rB.z : 0.f;
      a.dL_dcolors[3 * i + 1] = vis ? rB.w : 0.f;
      a.dL_dcolors[3 * i + 2] = vis ? rC.x : 0.f;
    }
    float gp[3] = {0.f, 0.f, 0.f};
    float gS[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int nc = (a.D + 1) * (a.D + 1);
    if (vis) {
      f3 p0 = {a.means[3 * i], a.means[3 * i + 1], a.means[3 * i + 2]};
      const f3 p = a.pose ? pose_apply(ps, p0) : p0;
      const float* c3 = a.cov3D + 6 * i;
      // ---- conic -> 2-D covariance -> 3-D covariance and mean (through J)
      const f3 t = xf43(a.view, p);
      const float limx = 1.3f * a.tanfovx, limy = 1.3f * a.tanfovy;
      const float txtz = t.x / t.z, tytz = t.y / t.z;
      const float tx = fminf(limx, fmaxf(-limx, txtz)) * t.z;
      const float ty = fminf(limy, fmaxf(-limy, tytz)) * t.z;
      const float xm = (txtz < -limx || txtz > limx) ? 0.f : 1.f;
      const float ym = (tytz < -limy || tytz > limy) ? 0.f : 1.f;
      const float tz = t.z;
      const float J00 = a.fx / tz, J02 = -(a.fx * tx) / (tz * tz);
      const float J11 = a.fy / tz, J12 = -(a.fy * ty) / (tz * tz);
      float T0[3], T1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T0[c] = J00 * a.view[c * 4 + 0] + J02 * a.view[c * 4 + 2];
        T1[c] = J11 * a.view[c * 4 + 1] + J12 * a.view[c * 4 + 2];
      }
      const float S[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
      float ST0[3], ST1[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ST0[k] = S[k][0] * T0[0] + S[k][1] * T0[1] + S[k][2] * T0[2];
        ST1[k] = S[k][0] * T1[0] + S[k][1] * T1[1] + S[k][2] * T1[2];
      }
      const float ca = T0[0] * ST0[0] + T0[1] * ST0[1] + T0[2] * ST0[2] + 0.3f;
      const float cb = T0[0] * ST1[0] + T0[1] * ST1[1] + T0[2] * ST1[2];
      const float cc = T1[0] * ST1[0] + T1[1] * ST1[1] + T1[2] * ST1[2] + 0.3f;
      const float den = ca * cc - cb * cb;
      const float gA = -0.5f * rA.z, gB = rA.w, gC = -0.5f * rB.x;  // (the record's ca and cc are -2x)
      const float d2i = 1.0f / (den * den + 0.0000001f);
      float dLa = 0.f, dLb = 0.f, dLc = 0.f;
      if (d2i != 0.0f) {
        dLa = d2i * (-cc * cc * gA + cb * cc * gB + (den - ca * cc) * gC);
        dLc = d2i * (-ca * ca * gC + ca * cb * gB + (den - ca * cc) * gA);
        dLb = d2i * (2.0f * cb * cc * gA - (den + 2.0f * cb * cb) * gB + 2.0f * ca * cb * gC);
        gS[0] = T0[0] * T0[0] * dLa + T0[0] * T1[0] * dLb + T1[0] * T1[0] * dLc;
        gS[3] = T0[1] * T0[1] * dLa + T0[1] * T1[1] * dLb + T1[1] * T1[1] * dLc;
        gS[5] = T0[2] * T0[2] * dLa + T0[2] * T1[2] * dLb + T1[2] * T1[2] * dLc;
        gS[1] = 2.0f * T0[0] * T0[1] * dLa + (T0[0] * T1[1] + T0[1] * T1[0]) * dLb + 2.0f * T1[0] * T1[1] * dLc;
        gS[2] = 2.0f * T0[0] * T0[2] * dLa + (T0[0] * T1[2] + T0[2] * T1[0]) * dLb + 2.0f * T1[0] * T1[2] * dLc;
        gS[4] = 2.0f * T0[1] * T0[2] * dLa + (T0[1] * T1[2] + T0[2] * T1[1]) * dLb + 2.0f * T1[1] * T1[2] * dLc;
      }
      float gJ00 = 0.f, gJ02 = 0.f, gJ11 = 0.f, gJ12 = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float gT0 = 2.0f * dLa * ST0[c] + dLb * ST1[c];
        const float gT1 = 2.0f * dLc * ST1[c] + dLb * ST0[c];
        gJ00 += gT0 * a.view[c * 4 + 0];
        gJ02 += gT0 * a.view[c * 4 + 2];
        gJ11 += gT1 * a.view[c * 4 + 1];
        gJ12 += gT1 * a.view[c * 4 + 2];
      }
      const float iz = 1.0f / tz, iz2 = iz * iz, iz3 = iz2 * iz;
      const float gtx = xm * -a.fx * iz2 * gJ02;
      const float gty = ym * -a.fy * iz2 * gJ12;
      const float gtz = -a.fx * iz2 * gJ00 - a.fy * iz2 * gJ11 + (2.0f * a.fx * tx) * iz3 * gJ02 +
                        (2.0f * a.fy * ty) * iz3 * gJ12;
#pragma unroll
      for (int k = 0; k < 3; ++k) gp[k] = a.view[k * 4 + 0] * gtx + a.view[k * 4 + 1] * gty + a.view[k * 4 + 2] * gtz;
      // ---- projection of the mean (NDC units)
      const f3 ph = xf43(a.proj, p);
      const float mw = 1.0f / (xf44w(a.proj, p) + 0.0000001f);
      const float mul1 = ph.x * mw * mw, mul2 = ph.y * mw * mw;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        gp[k] += (a.proj[4 * k + 0] * mw - a.proj[4 * k + 3] * mul1) * rA.x +
                 (a.proj[4 * k + 1] * mw - a.proj[4 * k + 3] * mul2) * rA.y;
      // ---- depth = view-space z
#pragma unroll
      for (int k = 0; k < 3; ++k) gp[k] += a.view[4 * k + 2] * rC.y;
      // ---- SH colour (at degree 0 the colour does not depend on the mean)
      if (a.shs && (sh_grad || a.D > 0)) {
        const float cbits = a.splat[3 * i + 1].w;
        const uint32_t cm = __float_as_uint(cbits);
        const float gc[3] = {(cm & 1u) ? 0.f : rB.z, (cm & 2u) ? 0.f : rB.w, (cm & 4u) ? 0.f : rC.x};
        const f3 dv = {p.x - a.campos[0], p.y - a.campos[1], p.z - a.campos[2]};
        const float len = sqrtf(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z);
        const float dx = dv.x / len, dy = dv.y / len, dz = dv.z / len;
        float b[16], bg3[16][3];
        sh_basis(a.D, dx, dy, dz, b);
        sh_basis_grad(a.D, dx, dy, dz, bg3);
        const float* sh = staged ? s_sh + threadIdx.x * ldsrow : a.shs + (size_t)i * row;
        float* gsh = (!sh_grad || swap_rows) ? nullptr
                                             : (staged ? s_sh + threadIdx.x * ldsrow : a.dL_dsh + (size_t)i * row);
        float gdir0 = 0.f, gdir1 = 0.f, gdir2 = 0.f;
        // unrolled to the 16 coefficients of degree 3, so b / bg3 are indexed
        // by constants and live in registers (a runtime bound puts them in scratch)
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (k < nc) {
            const float s0 = sh[3 * k], s1 = sh[3 * k + 1], s2 = sh[3 * k + 2];
            if (gsh) {
              gsh[3 * k] = b[k] * gc[0];
              gsh[3 * k + 1] = b[k] * gc[1];
              gsh[3 * k + 2] = b[k] * gc[2];
            }
            const float sd = s0 * gc[0] + s1 * gc[1] + s2 * gc[2];
            gdir0 += sd * bg3[k][0];
            gdir1 += sd * bg3[k][1];
            gdir2 += sd * bg3[k][2];
          }
        }
        if (gsh)
          for (int k = 3 * nc; k < 3 * a.M; ++k) gsh[k] = 0.f;
        if (swap_rows) {
#pragma unroll
          for (int k = 0; k < 16; ++k) sb[k] = k < nc ? b[k] : 0.f;
          sgc[0] = gc[0];
          sgc[1] = gc[1];
          sgc[2] = gc[2];
          row_nc = nc;
        }
        const float ndot = dx * gdir0 + dy * gdir1 + dz * gdir2;
        gp[0] += (gdir0 - dx * ndot) / len;
        gp[1] += (gdir1 - dy * ndot) / len;
        gp[2] += (gdir2 - dz * ndot) / len;
      }
      // ---- scale / rotation
      if (!a.cov_pre && (a.dL_dscales || a.dL_drots)) {
        const float4 q = reinterpret_cast<const float4*>(a.rots)[i];
        float R[9];
        quat_to_R(q.x, q.y, q.z, q.w, R);
        const float s[3] = {a.scale_mod * a.scales[3 * i], a.scale_mod * a.scales[3 * i + 1],
                            a.scale_mod * a.scales[3 * i + 2]};
        const float G[9] = {gS[0], 0.5f * gS[1], 0.5f * gS[2], 0.5f * gS[1], gS[3],
                            0.5f * gS[4], 0.5f * gS[2], 0.5f * gS[4], gS[5]};
        float L[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int k = 0; k < 3; ++k) L[3 * r + k] = R[3 * r + k] * s[k];
        float gR[9];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float gsk = 0.f;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float gL = 2.0f * (G[3 * r] * L[k] + G[3 * r + 1] * L[3 + k] + G[3 * r + 2] * L[6 + k]);
            gsk += gL * R[3 * r + k];
            gR[3 * r + k] = gL * s[k];
          }
          if (a.dL_dscales) a.dL_dscales[3 * i + k] = gsk;
        }
        float dq[4];
        quat_R_grad(q.x, q.y, q.z, q.w, gR, dq);
        if (a.dL_drots) reinterpret_cast<float4*>(a.dL_drots)[i] = make_float4(dq[0], dq[1], dq[2], dq[3]);
      }
      if (a.pose) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          pacc[r] = gp[r];
          pacc[3 + 3 * r + 0] = gp[r] * p0.x;
          pacc[3 + 3 * r + 1] = gp[r] * p0.y;
          pacc[3 + 3 * r + 2] = gp[r] * p0.z;
        }
        if (a.dL_dmeans3D)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            a.dL_dmeans3D[3 * i + k] = ps.R[k] * gp[0] + ps.R[3 + k] * gp[1] + ps.R[6 + k] * gp[2];
      } else if (a.adam_on) {
        means_adam(a, i, gp);
      } else if (a.dL_dmeans3D) {
        a.dL_dmeans3D[3 * i] = gp[0];
        a.dL_dmeans3D[3 * i + 1] = gp[1];
        a.dL_dmeans3D[3 * i + 2] = gp[2];
      }
    } else {
      if (a.adam_on) {
        const float g0[3] = {0.f, 0.f, 0.f};
        means_adam(a, i, g0);
      } else if (a.dL_dmeans3D) {
        a.dL_dmeans3D[3 * i] = a.dL_dmeans3D[3 * i + 1] = a.dL_dmeans3D[3 * i + 2] = 0.f;
      }
      if (sh_grad && !swap_rows) {
        float* gsh = staged ? s_sh + threadIdx.x * ldsrow : a.dL_dsh + (size_t)i * row;
        for (int k = 0; k < row; ++k) gsh[k] = 0.f;
      }
      if (a.dL_dscales) a.dL_dscales[3 * i] = a.dL_dscales[3 * i + 1] = a.dL_dscales[3 * i + 2] = 0.f;
      if (a.dL_drots) reinterpret_cast<float4*>(a.dL_drots)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (a.dL_dcov3D)
#pragma unroll
      for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * i + k] = gS[k];
  }
  if (swap_rows) {
    // parameter row out of LDS, gradient row in; the moments' pass replaces
    // the gradients by the decrements; p - decrement goes back to LDS and the
    // block is stored with coalesced 16-byte writes.  The own-row passes use
    // 16-byte accesses when the row is whole vectors.
    float prow[48];
    float* own = s_sh + threadIdx.x * ldsrow;
    const bool mine = (int)threadIdx.x < nloc;
    auto gval = [&](int e) { return e / 3 < row_nc ? sb[e / 3] * sgc[e % 3] : 0.f; };
    if (mine) {
      if ((row & 3) == 0) {
        float4* own4 = reinterpret_cast<float4*>(own);
#pragma unroll
        for (int q = 0; q < 12; ++q)
          if (4 * q < row) {
            const float4 v = own4[q];
            prow[4 * q] = v.x;
            prow[4 * q + 1] = v.y;
            prow[4 * q + 2] = v.z;
            prow[4 * q + 3] = v.w;
            own4[q] = make_float4(gval(4 * q), gval(4 * q + 1), gval(4 * q + 2), gval(4 * q + 3));
          }
      } else {
#pragma unroll
        for (int k = 0; k < 48; ++k)
          if (k < row) {
            prow[k] = own[k];
            own[k] = gval(k);
          }
      }
    }
    __syncthreads();
    block_adam_moments(a, s_sh, base, nloc, row, ldsrow);
    __syncthreads();
    if (mine) {
      if ((row & 3) == 0) {
        float4* own4 = reinterpret_cast<float4*>(own);
#pragma unroll
        for (int q = 0; q < 12; ++q)
          if (4 * q < row) {
            const float4 d = own4[q];
            own4[q] = make_float4(prow[4 * q] - d.x, prow[4 * q + 1] - d.y, prow[4 * q + 2] - d.z,
                                  prow[4 * q + 3] - d.w);
          }
      } else {
#pragma unroll
        for (int k = 0; k < 48; ++k)
          if (k < row) own[k] = prow[k] - own[k];
      }
    }
    __syncthreads();
    float* P = a.shs_p + (size_t)base * row;
    const int total = nloc * row;
    if ((row & 3) == 0) {
      const float4* s4 = reinterpret_cast<const float4*>(s_sh);
      const int per4 = row >> 2, lper4 = ldsrow >> 2;
      for (int k = threadIdx.x; k < (total >> 2); k += 256) st_nt(P + 4 * k, s4[(k / per4) * lper4 + k % per4]);
    } else {
      for (int k = threadIdx.x; k < total; k += 256) P[k] = s_sh[k];
    }
  } else if (staged) {
    __syncthreads();
    float* dst = a.dL_dsh + (size_t)base * row;
    const int total = nloc * row;
    if ((a.M & 3) == 0) {
      float4* d4 = reinterpret_cast<float4*>(dst);
      const float4* s4 = reinterpret_cast<const float4*>(s_sh);
      const int per4 = row >> 2, lper4 = ldsrow >> 2;
      for (int k = threadIdx.x; k < (total >> 2); k += 256) d4[k] = s4[(k / per4) * lper4 + k % per4];
    } else {
      for (int k = threadIdx.x; k < total; k += 256) dst[k] = s_sh[k];
    }
  }
  if (a.pose) {
    // all lanes reach here (no early return above); every wave of the grid
    // stores its 12 sums, so the finalize reads a fully written array
    float* dst = a.pose_part + ((size_t)bid * 4 + (threadIdx.x >> 6)) * 12;
    float w[3];
    wave_sum12(pacc, w);
    const int lane = threadIdx.x & 63;
    if ((lane & 15) == 0) {
      const int q = lane >> 4;
      const int o = q == 1 ? 2 : (q == 2 ? 1 : q);  // value of row q: 4m + {0, 2, 1, 3}[q]
#pragma unroll
      for (int m = 0; m < 3; ++m) dst[4 * m + o] = w[m];
    }
  }
}

__global__ __launch_bounds__(256) void k_preprocess_bwd(PBArgs a) {
  extern __shared__ float s_sh[];
  preprocess_bwd_block(a, blockIdx.x, s_sh);
}

// Many frames in one launch: frame f owns workgroups [prefix[f], prefix[f +
// 1]) = ceil(P_f / 256) of them.  The dynamic LDS covers the largest frame's
// SH rows.
__global__ __launch_bounds__(256) void k_preprocess_bwd_multi(const uint32_t* prefix, const PBArgs* args, int n) {
  extern __shared__ float s_sh[];
  const uint32_t b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const PBArgs a = args[lo];
  preprocess_bwd_block(a, (int)(b - prefix[lo]), s_sh);
}

// One frame's pose gradient from its nparts wave partials (12 terms each), by
// one workgroup of 256 threads.
__device__ __forceinline__ void pose_finalize_block(const float* pose7, const float* part, int nparts,
                                                    float* dL_dpose) {
  __shared__ float s[12][256];
  const int t = threadIdx.x;
  float sum[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) sum[k] = 0.f;
  for (int w = t; w < nparts; w += 256)
#pragma unroll
    for (int k = 0; k < 12; ++k) sum[k] += part[(size_t)w * 12 + k];
#pragma unroll
  for (int k = 0; k < 12; ++k) s[k][t] = sum[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
#pragma unroll
      for (int k = 0; k < 12; ++k) s[k][t] += s[k][t + off];
    __syncthreads();
  }
  if (t != 0) return;
  float acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = s[k][0];
  Pose ps;
  make_pose(pose7, ps);
  float GR[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) GR[k] = acc[3 + k];
  float dqh[4];
  quat_R_grad(ps.q[0], ps.q[1], ps.q[2], ps.q[3], GR, dqh);
  const float dot = ps.q[0] * dqh[0] + ps.q[1] * dqh[1] + ps.q[2] * dqh[2] + ps.q[3] * dqh[3];
  float dq[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) dq[k] = (dqh[k] - ps.q[k] * dot) * ps.inv_norm;
  dL_dpose[0] = acc[0];
  dL_dpose[1] = acc[1];
  dL_dpose[2] = acc[2];
  dL_dpose[3] = dq[1];
  dL_dpose[4] = dq[2];
  dL_dpose[5] = dq[3];
  dL_dpose[6] = dq[0];
}

__global__ __launch_bounds__(256) void k_pose_finalize(const float* pose7, const float* part, int nparts,
                                                        float* dL_dpose) {
  pose_finalize_block(pose7, part, nparts, dL_dpose);
}

// The pose reduction of one frame of a multi-frame finish.
struct PoseFin {
  const float* pose7;
  const float* part;
  float* dL_dpose;
  int nparts;
};

// Workgroup f reduces frame f.
__global__ __launch_bounds__(256) void k_pose_finalize_multi(const PoseFin* fin) {
  const PoseFin p = fin[blockIdx.x];
  pose_finalize_block(p.pose7, p.part, p.nparts, p.dL_dpose);
}

}  // namespace gs

using namespace gs;

// scratch: render-gradient records [P][kRG] | pose wave partials [4 * ceil(P/256)][12]
//          | deterministic mode: pair partials [capacity][10]
static size_t pose_parts(int P) { return (size_t)4 * (size_t)((P + 255) / 256); }

extern "C" size_t gs_backward_scratch_bytes(int P, long long capacity, int deterministic) {
  if (P < 0) P = 0;
  if (capacity < 0) capacity = 0;
  size_t n = align_up((size_t)P * kRG * sizeof(float), 256) + align_up(pose_parts(P) * 12 * sizeof(float), 256) + 256;
  if (deterministic) n += align_up((size_t)capacity * 10 * sizeof(float), 256);
  return n;
}

// Render-backward arguments of one frame (workspaces of its forward).
static RBArgs make_rb_args(const gs_frame_t* f, int P, const void* geom, long long capacity, void* binning,
                           const void* image, const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                           void* scratch) {
  gs_geom_layout_t LG;
  gs_binning_layout_t LB;
  gs_image_layout_t LI;
  gs_geom_layout(P, &LG);
  gs_binning_layout(capacity, P, f->width, f->height, f->binning, &LB);
  gs_image_layout(f->width, f->height, &LI);
  const int gx = (f->width + kBX - 1) / kBX, gy = (f->height + kBY - 1) / kBY;
  const char* gb = (const char*)geom;
  const char* bb = (const char*)binning;
  const char* ib = (const char*)image;
  float* g_render = (float*)scratch;
  float* pose_part = (float*)((char*)scratch + align_up((size_t)P * kRG * sizeof(float), 256));
  float* part = (float*)((char*)pose_part + align_up(pose_parts(P) * 12 * sizeof(float), 256));
  RBArgs r;
  r.W = f->width;
  r.H = f->height;
  r.gx = gx;
  r.gy = gy;
  r.nt = gx * gy;
  r.ranges = (const uint32_t*)(ib + LI.ranges);
  r.list = capacity > 0 ? (const uint32_t*)(bb + LB.vals_sorted) : nullptr;
  r.id_mask = LB.id_bits ? (1u << LB.id_bits) - 1u : 0xffffffffu;
  r.splat = (const float4*)(gb + LG.splat);
  r.bg = f->bg;
  r.final_T = (const float*)(ib + LI.final_T);
  r.n_contrib = (const uint32_t*)(ib + LI.n_contrib);
  r.accum = (const float*)(ib + LI.accum);
  r.ckpt = capacity > 0 ? (const float*)(bb + LB.ckpt) : nullptr;
  r.ckpt_base = (const uint32_t*)(ib + LI.ckpt_base);
  r.tile_maxc = (const uint32_t*)(ib + LI.tile_maxc);
  r.items = (const uint32_t*)(bb + LB.items);
  r.item_log = (const uint32_t*)(bb + LB.item_log);
  r.dL_dcolor = dL_dcolor;
  r.dL_ddepth = dL_ddepth;
  r.dL_dalpha = dL_dalpha;
  r.g_render = g_render;
  r.local_incl = (const uint32_t*)(gb + LG.offsets);
  r.block_excl = (const uint32_t*)(gb + LG.scan_tmp);
  r.part = part;
  return r;
}

// Work-item upper bound of one frame's render backward: sum over tiles of
// ceil(list length / 256) <= capacity / 256 + tiles.
static long long rb_max_items(const gs_frame_t* f, long long capacity) {
  const int gx = (f->width + kBX - 1) / kBX, gy = (f->height + kBY - 1) / kBY;
  return capacity / 256 + (long long)gx * gy + 1;
}

// The pose-only backward applies when the fixed-order reductions are on, only
// the pose gradient is wanted, the loss has no depth or alpha term, and the
// colours do not depend on the view direction (active SH degree 0, or
// precomputed colours): the pose gradient then reads only the screen-position
// and conic terms of each pair.
static bool pose_only(const gs_frame_t* f, const gs_gaussians_t* g, const gs_grads_t* out, bool dep) {
  return f->deterministic && f->pose && out->dL_dpose && !dep && !out->dL_dmeans2D && !out->dL_dcolors &&
         !out->dL_dopacity && !out->dL_dmeans3D && !out->dL_dcov3D && !out->dL_dsh && !out->dL_dscales &&
         !out->dL_drotations && (f->sh_degree == 0 || !g->shs);
}

static int bwd_check(const gs_frame_t* f, const gs_gaussians_t* g, const int* radii, const void* geom,
                     void* binning, long long capacity, const void* image, const uint32_t* counters,
                     const float* dL_dcolor, const gs_grads_t* out, void* scratch, size_t scratch_bytes) {
  if (!f || !g || !out || !counters || !image || !dL_dcolor) return GS_E_ARG;
  if (f->width <= 0 || f->height <= 0 || !f->viewmatrix || !f->projmatrix || !f->bg) return GS_E_ARG;
  const int P = g->P;
  if (P < 0) return GS_E_ARG;
  if (P > 0 && (g->shs == nullptr) == (g->colors_precomp == nullptr)) return GS_E_ARG;
  const bool has_sr = g->scales && g->rotations;
  if (P > 0 && has_sr == (g->cov3D_precomp != nullptr)) return GS_E_ARG;
  // per-Gaussian outputs are optional (NULL = not wanted)
  if (P > 0 && (!radii || !geom)) return GS_E_ARG;
  if (!g->shs && out->dL_dsh) return GS_E_ARG;
  if (!has_sr && (out->dL_dscales || out->dL_drotations)) return GS_E_ARG;
  if (out->dL_dpose && !f->pose) return GS_E_ARG;
  if (capacity < 0 || capacity > 0x7fffffffLL || (capacity > 0 && !binning)) return GS_E_ARG;
  if (scratch_bytes < gs_backward_scratch_bytes(P, capacity, f->deterministic != 0) || !scratch)
    return GS_E_WORKSPACE;
  gs_binning_layout_t LB;
  if (gs_binning_layout(capacity, P, f->width, f->height, f->binning, &LB) != GS_OK) return GS_E_ARG;
  return GS_OK;
}

// Phase 1: clear the accumulators and build the work-item table.
// One frame's set-up and the workgroups it takes (0: nothing to do).
static BwdPrep bwd_prep_of(const gs_frame_t* f, int P, long long capacity, const uint32_t* counters,
                           const RBArgs& r, bool po, unsigned* grid) {
  BwdPrep p;
  // deterministic mode writes every record in k_det_gather; pairs past a
  // tile's last contributor are never visited, so their partials start at 0
  p.zp = f->deterministic ? r.part : r.g_render;
  p.zn = f->deterministic ? capacity * (po ? 5 : 10) : (long long)P * kRG;
  p.counters = counters;
  p.terms = f->deterministic ? (po ? 5 : 10) : 0;
  // the work-item table, from the forward's item log (both live in the
  // binning workspace)
  p.item_log = r.item_log;
  p.items = capacity > 0 && P > 0 ? (uint32_t*)r.items : nullptr;
  const long long nzb = (p.zn + 1023) / 1024;
  *grid = (unsigned)((nzb < 2048 ? nzb : 2048) + (p.items ? 1 : 0));
  return p;
}

static void bwd_prepare(const gs_frame_t* f, int P, long long capacity, const uint32_t* counters, const RBArgs& r,
                        hipStream_t s, bool po) {
  unsigned grid = 0;
  const BwdPrep p = bwd_prep_of(f, P, capacity, counters, r, po, &grid);
  if (grid > 0) hipLaunchKernelGGL(k_bwd_prepare, dim3(grid), dim3(1024), 0, s, p);
}

// Phase 2: the render backward of one frame.
static void bwd_render(const gs_frame_t* f, int P, long long capacity, const RBArgs& r, hipStream_t s, bool po) {
  if (capacity <= 0 || P <= 0) return;
  const long long max_items = rb_max_items(f, capacity);
  timing::Scope ts(timing::kRenderBwd, s);
  const dim3 grid((unsigned)max_items);
  const bool dep = r.dL_ddepth || r.dL_dalpha;
  if (po) {
    hipLaunchKernelGGL((k_render_bwd<true, false, true>), grid, dim3(256), 0, s, r);
  } else if (f->deterministic) {
    if (dep) hipLaunchKernelGGL((k_render_bwd<true, true, false>), grid, dim3(256), 0, s, r);
    else hipLaunchKernelGGL((k_render_bwd<true, false, false>), grid, dim3(256), 0, s, r);
  } else {
    if (dep) hipLaunchKernelGGL((k_render_bwd<false, true, false>), grid, dim3(256), 0, s, r);
    else hipLaunchKernelGGL((k_render_bwd<false, false, false>), grid, dim3(256), 0, s, r);
  }
}

// Phase 3: per-Gaussian gradients and the pose gradient.
// Host-side values of a fused Adam step (gs_backward_adam), checked.
struct AdamFusedHost {
  float* means_p;
  float* shs_p;
  const gs_adam_fused_t* in;
  float means_step, shs_step, shs_lead_step, inv_sqrt_bc2;
};

// The per-Gaussian backward's arguments for one frame.
static PBArgs make_pb_args(const gs_frame_t* f, const gs_gaussians_t* g, const int* radii, const void* geom,
                           long long capacity, const uint32_t* counters, const gs_grads_t* out, const RBArgs& r,
                           bool po, const AdamFusedHost* ad) {
  const int P = g->P;
  gs_geom_layout_t LG;
  gs_geom_layout(P, &LG);
  const char* gb = (const char*)geom;
  float* pose_part = (float*)((char*)r.g_render + align_up((size_t)P * kRG * sizeof(float), 256));
  PBArgs b;
  b.P = P;
  b.M = g->M;
  b.D = f->sh_degree;
  b.W = f->width;
  b.H = f->height;
  b.tanfovx = f->tanfovx;
  b.tanfovy = f->tanfovy;
  b.fx = (float)f->width / (2.0f * f->tanfovx);
  b.fy = (float)f->height / (2.0f * f->tanfovy);
  b.scale_mod = f->scale_modifier;
  b.view = f->viewmatrix;
  b.proj = f->projmatrix;
  b.campos = f->campos;
  b.pose = f->pose;
  b.means = g->means3D;
  b.shs = g->shs;
  b.scales = g->scales;
  b.rots = g->rotations;
  b.cov_pre = g->cov3D_precomp;
  b.radii = radii;
  b.splat = (const float4*)(gb + LG.splat);
  b.cov3D = (const float*)(gb + LG.cov3D);
  b.g_render = r.g_render;
  b.dL_dmeans2D = out->dL_dmeans2D;
  b.dL_dcolors = out->dL_dcolors;
  b.dL_dopacity = out->dL_dopacity;
  b.dL_dmeans3D = out->dL_dmeans3D;
  b.dL_dcov3D = out->dL_dcov3D;
  b.dL_dsh = out->dL_dsh;
  b.dL_dscales = out->dL_dscales;
  b.dL_drots = out->dL_drotations;
  b.pose_part = pose_part;
  b.part5 = po ? (const float*)r.part : nullptr;
  b.g_tiles = (const uint32_t*)(gb + LG.tiles);
  b.local_incl = (const uint32_t*)(gb + LG.offsets);
  b.block_excl = (const uint32_t*)(gb + LG.scan_tmp);
  b.counters = counters;
  b.capacity = capacity;
  b.adam_on = ad ? 1 : 0;
  b.means_p = ad ? ad->means_p : nullptr;
  b.means_m = ad ? ad->in->means_m : nullptr;
  b.means_v = ad ? ad->in->means_v : nullptr;
  b.shs_p = ad ? ad->shs_p : nullptr;
  b.shs_m = ad ? ad->in->shs_m : nullptr;
  b.shs_v = ad ? ad->in->shs_v : nullptr;
  b.means_step = ad ? ad->means_step : 0.f;
  b.shs_step = ad ? ad->shs_step : 0.f;
  b.shs_lead_step = ad ? ad->shs_lead_step : 0.f;
  b.shs_lead = ad ? ad->in->shs_lead : 0;
  b.b1 = ad ? ad->in->beta1 : 0.f;
  b.b2 = ad ? ad->in->beta2 : 0.f;
  b.eps = ad ? ad->in->eps : 0.f;
  b.inv_sqrt_bc2 = ad ? ad->inv_sqrt_bc2 : 0.f;
  return b;
}

static void bwd_finish(const gs_frame_t* f, const gs_gaussians_t* g, const int* radii, const void* geom,
                       long long capacity, const uint32_t* counters, const gs_grads_t* out, const RBArgs& r,
                       hipStream_t s, bool po, const AdamFusedHost* ad = nullptr) {
  const int P = g->P;
  gs_geom_layout_t LG;
  gs_geom_layout(P, &LG);
  const char* gb = (const char*)geom;
  float* pose_part = (float*)((char*)r.g_render + align_up((size_t)P * kRG * sizeof(float), 256));
  // (the pose-only backward gathers its 5-term partials inside k_preprocess_bwd)
  if (f->deterministic && P > 0 && !po)
    hipLaunchKernelGGL(k_det_gather<10>, dim3((P + 255) / 256), dim3(256), 0, s, P,
                       (const uint32_t*)(gb + LG.tiles), (const uint32_t*)(gb + LG.offsets),
                       (const uint32_t*)(gb + LG.scan_tmp), counters, capacity, (const float*)r.part, r.g_render);
  if (P > 0) {
    const PBArgs b = make_pb_args(f, g, radii, geom, capacity, counters, out, r, po, ad);
    timing::Scope ts(timing::kPreprocessBwd, s);
    const bool sh_grad = g->shs && (out->dL_dsh || ad);
    // LDS rows as the kernel lays them out: 3M floats, plus 4 when 3M is a
    // multiple of 4
    const size_t shm = (sh_grad && g->M <= 16) ? (size_t)256 * (3 * g->M + ((g->M & 3) == 0 ? 4 : 0)) * sizeof(float)
                                               : 0;
    hipLaunchKernelGGL(k_preprocess_bwd, dim3((P + 255) / 256), dim3(256), shm, s, b);
  }
  // (P = 0: no wave partials, the pose gradient is zero)
  if (out->dL_dpose)
    hipLaunchKernelGGL(k_pose_finalize, dim3(1), dim3(256), 0, s, f->pose, pose_part, (int)pose_parts(P),
                       out->dL_dpose);
}

#define GS_BWD_ARGS                                                                                              \
  const gs_frame_t *f, const gs_gaussians_t *g, const int *radii, const void *geom, void *binning,             \
      long long capacity, const void *image, const uint32_t *counters, const float *dL_dcolor,                  \
      const float *dL_ddepth,                                                                                   \
      const float *dL_dalpha, const gs_grads_t *out, void *scratch, size_t scratch_bytes, void *stream

extern "C" int gs_backward(GS_BWD_ARGS) {
  const int rc = bwd_check(f, g, radii, geom, binning, capacity, image, counters, dL_dcolor, out, scratch,
                           scratch_bytes);
  if (rc != GS_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const RBArgs r = make_rb_args(f, g->P, geom, capacity, binning, image, dL_dcolor, dL_ddepth, dL_dalpha, scratch);
  const bool po = pose_only(f, g, out, dL_ddepth || dL_dalpha);
  bwd_prepare(f, g->P, capacity, counters, r, s, po);
  bwd_render(f, g->P, capacity, r, s, po);
  bwd_finish(f, g, radii, geom, capacity, counters, out, r, s, po);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_backward_adam(const gs_frame_t* f, const gs_gaussians_t* g, const int* radii, const void* geom,
                                void* binning, long long capacity, const void* image, const uint32_t* counters,
                                const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                                const gs_grads_t* out, const gs_adam_fused_t* adam, void* scratch,
                                size_t scratch_bytes, void* stream) {
  const int rc = bwd_check(f, g, radii, geom, binning, capacity, image, counters, dL_dcolor, out, scratch,
                           scratch_bytes);
  if (rc != GS_OK) return rc;
  if (!adam || f->pose || out->dL_dmeans3D || out->dL_dsh || adam->step < 1) return GS_E_ARG;
  const int P = g->P;
  if (P > 0 && (!g->means3D || !adam->means_m || !adam->means_v)) return GS_E_ARG;
  if (P > 0 && g->shs && (g->M > 16 || !adam->shs_m || !adam->shs_v || adam->shs_lead < 0 ||
                          adam->shs_lead > 3 * g->M))
    return GS_E_ARG;
  const uintptr_t al = (uintptr_t)g->means3D | (uintptr_t)adam->means_m | (uintptr_t)adam->means_v |
                       (uintptr_t)(g->shs ? g->shs : nullptr) | (uintptr_t)(g->shs ? adam->shs_m : nullptr) |
                       (uintptr_t)(g->shs ? adam->shs_v : nullptr);
  if (al & 15) return GS_E_ARG;
  // the bias corrections as gs_adam_step_rows computes them
  const double bc1 = 1.0 - std::pow((double)adam->beta1, (double)adam->step);
  const double bc2 = 1.0 - std::pow((double)adam->beta2, (double)adam->step);
  AdamFusedHost ad;
  ad.in = adam;
  ad.means_p = const_cast<float*>(g->means3D);  // updated in place
  ad.shs_p = const_cast<float*>(g->shs);
  ad.means_step = (float)(adam->means_lr / bc1);
  ad.shs_step = (float)(adam->shs_lr / bc1);
  ad.shs_lead_step = (float)(adam->shs_lead_lr / bc1);
  ad.inv_sqrt_bc2 = (float)(1.0 / std::sqrt(bc2));
  hipStream_t s = (hipStream_t)stream;
  const RBArgs r = make_rb_args(f, P, geom, capacity, binning, image, dL_dcolor, dL_ddepth, dL_dalpha, scratch);
  bwd_prepare(f, P, capacity, counters, r, s, false);
  bwd_render(f, P, capacity, r, s, false);
  bwd_finish(f, g, radii, geom, capacity, counters, out, r, s, false, &ad);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_backward_prepare(GS_BWD_ARGS) {
  const int rc = bwd_check(f, g, radii, geom, binning, capacity, image, counters, dL_dcolor, out, scratch,
                           scratch_bytes);
  if (rc != GS_OK) return rc;
  const RBArgs r = make_rb_args(f, g->P, geom, capacity, binning, image, dL_dcolor, dL_ddepth, dL_dalpha, scratch);
  bwd_prepare(f, g->P, capacity, counters, r, (hipStream_t)stream, pose_only(f, g, out, dL_ddepth || dL_dalpha));
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_backward_finish(GS_BWD_ARGS) {
  const int rc = bwd_check(f, g, radii, geom, binning, capacity, image, counters, dL_dcolor, out, scratch,
                           scratch_bytes);
  if (rc != GS_OK) return rc;
  const RBArgs r = make_rb_args(f, g->P, geom, capacity, binning, image, dL_dcolor, dL_ddepth, dL_dalpha, scratch);
  bwd_finish(f, g, radii, geom, capacity, counters, out, r, (hipStream_t)stream,
             pose_only(f, g, out, dL_ddepth || dL_dalpha));
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// ---- many frames in one render-backward launch -----------------------------
// Set layout: uint32 prefix[n + 1] of work-item bounds per frame, then
// RBArgs[n] at a 256-byte boundary.
static size_t multi_rb_args_offset(int n) { return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256); }

extern "C" size_t gs_multi_render_bwd_bytes(int n) {
  return n <= 0 ? 0 : multi_rb_args_offset(n) + sizeof(RBArgs) * (size_t)n;
}

extern "C" int gs_multi_render_bwd_build(int n, const gs_frame_t* frames, const int* P, const void* const* geom,
                                         const long long* capacity, void* const* binning, const void* const* image,
                                         const float* const* dL_dcolor, const float* const* dL_ddepth,
                                         const float* const* dL_dalpha, void* const* scratch, void* host_set,
                                         size_t set_bytes, unsigned* total_blocks) {
  if (n <= 0 || !frames || !P || !geom || !capacity || !binning || !image || !dL_dcolor || !scratch || !host_set ||
      !total_blocks)
    return GS_E_ARG;
  if (set_bytes < gs_multi_render_bwd_bytes(n)) return GS_E_WORKSPACE;
  uint32_t* prefix = (uint32_t*)host_set;
  RBArgs* args = (RBArgs*)((char*)host_set + multi_rb_args_offset(n));
  unsigned long long total = 0;
  const int det = frames[0].deterministic != 0;
  for (int i = 0; i < n; ++i) {
    const gs_frame_t* f = &frames[i];
    if (f->width <= 0 || f->height <= 0 || !f->bg || P[i] < 0 || !image[i] || !dL_dcolor[i] || !scratch[i])
      return GS_E_ARG;
    if ((f->deterministic != 0) != det) return GS_E_ARG;  // one kernel variant per launch
    if (capacity[i] < 0 || capacity[i] > 0x7fffffffLL || (P[i] > 0 && !geom[i]) || (capacity[i] > 0 && !binning[i]))
      return GS_E_ARG;
    prefix[i] = (uint32_t)total;
    if (capacity[i] > 0 && P[i] > 0) total += (unsigned long long)rb_max_items(f, capacity[i]);
    if (total > 0x7fffffffULL) return GS_E_ARG;
    args[i] = make_rb_args(f, P[i], geom[i], capacity[i], binning[i], image[i], dL_dcolor[i],
                           dL_ddepth ? dL_ddepth[i] : nullptr, dL_dalpha ? dL_dalpha[i] : nullptr, scratch[i]);
  }
  prefix[n] = (uint32_t)total;
  *total_blocks = (unsigned)total;
  return GS_OK;
}

extern "C" int gs_multi_render_bwd(int n, const void* dev_set, unsigned total_blocks, int flags, void* stream) {
  if (n <= 0 || !dev_set || (flags & ~7)) return GS_E_ARG;
  const bool det = flags & 1, dep = !(flags & GS_RB_COLOR_ONLY), po = flags & GS_RB_POSE_ONLY;
  if (po && (!det || dep)) return GS_E_ARG;
  if (total_blocks == 0) return GS_OK;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t* prefix = (const uint32_t*)dev_set;
  const RBArgs* args = (const RBArgs*)((const char*)dev_set + multi_rb_args_offset(n));
  timing::Scope ts(timing::kRenderBwd, s);
  const dim3 grid(total_blocks);
  if (po) {
    hipLaunchKernelGGL((k_render_bwd_multi<true, false, true>), grid, dim3(256), 0, s, prefix, args, n);
  } else if (det) {
    if (dep) hipLaunchKernelGGL((k_render_bwd_multi<true, true, false>), grid, dim3(256), 0, s, prefix, args, n);
    else hipLaunchKernelGGL((k_render_bwd_multi<true, false, false>), grid, dim3(256), 0, s, prefix, args, n);
  } else {
    if (dep) hipLaunchKernelGGL((k_render_bwd_multi<false, true, false>), grid, dim3(256), 0, s, prefix, args, n);
    else hipLaunchKernelGGL((k_render_bwd_multi<false, false, false>), grid, dim3(256), 0, s, prefix, args, n);
  }
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

// ---- the pose-only backward set-up and finish of many frames ---------------
// Set layout: uint32 prefix[n + 1] of per-Gaussian workgroup bounds per frame,
// uint32 prefix[n + 1] of set-up workgroup bounds, then BwdPrep[n], PBArgs[n]
// and PoseFin[n], each array at a 256-byte boundary.
static size_t mpb_prep_prefix_offset(int n) { return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256); }
static size_t mpb_prep_offset(int n) { return 2 * mpb_prep_prefix_offset(n); }
static size_t mpb_args_offset(int n) { return mpb_prep_offset(n) + align_up(sizeof(BwdPrep) * (size_t)n, 256); }
static size_t mpb_fin_offset(int n) { return mpb_args_offset(n) + align_up(sizeof(PBArgs) * (size_t)n, 256); }

extern "C" size_t gs_multi_pose_bwd_bytes(int n) {
  return n <= 0 ? 0 : mpb_fin_offset(n) + sizeof(PoseFin) * (size_t)n;
}

extern "C" int gs_multi_pose_bwd_build(int n, const gs_frame_t* frames, const gs_gaussians_t* gauss,
                                       const int* const* radii, const void* const* geom, void* const* binning,
                                       const long long* capacity, const void* const* image,
                                       const uint32_t* const* counters, const float* const* dL_dcolor,
                                       const gs_grads_t* out, void* const* scratch, const size_t* scratch_bytes,
                                       void* host_set, size_t set_bytes, unsigned* blocks) {
  if (n <= 0 || !frames || !gauss || !radii || !geom || !binning || !capacity || !image || !counters ||
      !dL_dcolor || !out || !scratch || !scratch_bytes || !host_set || !blocks)
    return GS_E_ARG;
  if (set_bytes < gs_multi_pose_bwd_bytes(n)) return GS_E_WORKSPACE;
  uint32_t* prefix = (uint32_t*)host_set;
  uint32_t* pprefix = (uint32_t*)((char*)host_set + mpb_prep_prefix_offset(n));
  BwdPrep* prep = (BwdPrep*)((char*)host_set + mpb_prep_offset(n));
  PBArgs* args = (PBArgs*)((char*)host_set + mpb_args_offset(n));
  PoseFin* fin = (PoseFin*)((char*)host_set + mpb_fin_offset(n));
  unsigned long long total = 0, ptotal = 0;
  for (int i = 0; i < n; ++i) {
    const gs_frame_t* f = &frames[i];
    const gs_gaussians_t* g = &gauss[i];
    const int rc = bwd_check(f, g, radii[i], geom[i], binning[i], capacity[i], image[i], counters[i],
                             dL_dcolor[i], &out[i], scratch[i], scratch_bytes[i]);
    if (rc != GS_OK) return rc;
    // only the pose-only backward (no per-Gaussian outputs, no SH rows in LDS)
    if (!pose_only(f, g, &out[i], false)) return GS_E_ARG;
    const int P = g->P;
    const RBArgs r = make_rb_args(f, P, geom[i], capacity[i], binning[i], image[i], dL_dcolor[i], nullptr, nullptr,
                                  scratch[i]);
    unsigned pgrid = 0;
    prep[i] = bwd_prep_of(f, P, capacity[i], counters[i], r, true, &pgrid);
    pprefix[i] = (uint32_t)ptotal;
    ptotal += pgrid;
    prefix[i] = (uint32_t)total;
    total += (unsigned long long)((P + 255) / 256);
    if (total > 0x7fffffffULL || ptotal > 0x7fffffffULL) return GS_E_ARG;
    args[i] = make_pb_args(f, g, radii[i], geom[i], capacity[i], counters[i], &out[i], r, true, nullptr);
    fin[i].pose7 = f->pose;
    fin[i].part = args[i].pose_part;
    fin[i].dL_dpose = out[i].dL_dpose;
    fin[i].nparts = (int)pose_parts(P);
  }
  prefix[n] = (uint32_t)total;
  pprefix[n] = (uint32_t)ptotal;
  blocks[0] = (unsigned)ptotal;
  blocks[1] = (unsigned)total;
  return GS_OK;
}

extern "C" int gs_multi_pose_bwd_prepare(int n, const void* dev_set, const unsigned* blocks, void* stream) {
  if (n <= 0 || !dev_set || !blocks) return GS_E_ARG;
  if (blocks[0] > 0)
    hipLaunchKernelGGL(k_bwd_prepare_multi, dim3(blocks[0]), dim3(1024), 0, (hipStream_t)stream,
                       (const uint32_t*)((const char*)dev_set + mpb_prep_prefix_offset(n)),
                       (const BwdPrep*)((const char*)dev_set + mpb_prep_offset(n)), n);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_multi_pose_bwd_finish(int n, const void* dev_set, const unsigned* blocks, void* stream) {
  if (n <= 0 || !dev_set || !blocks) return GS_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  {
    timing::Scope ts(timing::kPreprocessBwd, s);
    if (blocks[1] > 0)
      hipLaunchKernelGGL(k_preprocess_bwd_multi, dim3(blocks[1]), dim3(256), 0, s, (const uint32_t*)dev_set,
                         (const PBArgs*)((const char*)dev_set + mpb_args_offset(n)), n);
  }
  hipLaunchKernelGGL(k_pose_finalize_multi, dim3(n), dim3(256), 0, s,
                     (const PoseFin*)((const char*)dev_set + mpb_fin_offset(n)));
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_backward_is_pose_only(const gs_frame_t* f, const gs_gaussians_t* g, const gs_grads_t* out,
                                        int has_depth_or_alpha) {
  if (!f || !g || !out) return 0;
  return pose_only(f, g, out, has_depth_or_alpha != 0) ? 1 : 0;
}
