// gs_sort.hip -- stable LSD radix sort of (key, uint32 value) pairs, with
// 64-bit keys (tile << 32 | depth bits, the reference's key) or 32-bit keys
// (depth bits, or tile ids in the depth-first binning).
//
// The element count is read on the device (so the binning pipeline never
// needs the host to know num_rendered); grids are sized by capacity and
// workgroups past the count exit at once.
//
// Keys-only sorts (32-bit keys that carry their payload in the low bits)
// sort bits [start_bit, end_bit) and skip every value load and store.
//
// Per 8-bit digit pass:
//   k_sort_upsweep  1024 or 2048 keys per workgroup (4 waves x 4 or 8 rounds
//                   x 64 lanes), all loaded up front, counted with one LDS
//                   atomic each -> per-workgroup digit counts (digit-major,
//                   so one row scan per digit orders blocks)
//   k_sort_scan     one workgroup per digit: exclusive scan of its row, total
//   k_sort_scatter  recomputes the matches; rank = prior count of the digit in
//                   this wave + lanes below with the same digit; waves are
//                   ordered by an exclusive prefix over the 4 wave counts, so
//                   the order of equal digits is the input order (stable).
//                   The workgroup's keys are first placed in LDS in digit
//                   order, then written out by consecutive threads: each
//                   digit's run leaves as contiguous stores instead of one
//                   scattered 4-byte store per lane.
// Upsweep and scatter map workgroups to key blocks XCD-aware (xcd_remap):
// dispatch deals workgroups round-robin over the 8 XCDs, and each XCD gets a
// contiguous run of blocks.  A cache line of a digit's counts covers 32
// consecutive blocks, so it is then written and read within one XCD's L2
// instead of once by each of the eight.
#include "gs_common.hpp"
#include "gs_internal.hpp"

#include <algorithm>

namespace gs {

// Lanes whose digit equals this lane's digit (only lanes with valid set).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid, int bits) {
  uint64_t m = __ballot(valid);
  for (int b = 0; b < bits; ++b) {
    const bool on = (d >> b) & 1u;
    const uint64_t v = __ballot(on);
    m &= on ? v : ~v;
  }
  return m;
}

// Inclusive prefix sum over a 256-thread workgroup (thread t holds v); the
// workgroup total in `total`.  Each wave scans by shuffles, then one barrier
// combines the four wave totals (instead of a barrier pair per Hillis-Steele
// step).  s_w (4 words of LDS) may be rewritten only after a later barrier.
__device__ __forceinline__ uint32_t block_incl_scan256(uint32_t v, uint32_t* s_w, uint32_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  const uint32_t w0 = s_w[0], w1 = s_w[1], w2 = s_w[2], w3 = s_w[3];
  total = (w0 + w1) + (w2 + w3);
  return x + (w > 0 ? w0 : 0u) + (w > 1 ? w1 : 0u) + (w > 2 ? w2 : 0u);
}

__device__ __forceinline__ uint32_t n_eff(const uint32_t* n_dev, long long cap) {
  const uint32_t n = *n_dev;
  return ((long long)n > cap) ? 0u : n;
}

// Workgroup `bid` of an upsweep over one key array (the kernels below pass
// their block index, or, in a segmented sort, the block index within the
// segment).
template <typename K, int ITEMS, int R>
__device__ __forceinline__ void upsweep_block(const K* __restrict__ keys, const uint32_t* n_dev, long long cap,
                                              int shift, int bits, uint32_t* hist, int nb_max, uint32_t bid) {
  __shared__ uint32_t s_h[4][R];
  const uint32_t n = n_eff(n_dev, cap);
  constexpr uint32_t TILE = kSortThreads * ITEMS;
  const uint32_t nblk = (n + TILE - 1) / TILE;
  if (bid >= nblk) return;
  const uint32_t blk = xcd_remap(bid, nblk);  // consecutive blocks on one XCD (see the file comment)
  const uint32_t start = blk * TILE;
  for (int k = threadIdx.x; k < 4 * R; k += kSortThreads) (&s_h[0][0])[k] = 0;
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t mask = (1u << bits) - 1u;
  const uint32_t wbase = start + (uint32_t)w * (ITEMS * 64);
  // a histogram needs no ranks: one LDS atomic per key into the wave's own
  // row, with every key of the thread loaded before the first atomic.  Order
  // does not matter here, so 32-bit keys come four to a lane in 16-byte loads
  // (the key arrays are 16-byte aligned).
  if constexpr (sizeof(K) == 4 && ITEMS % 4 == 0) {
    constexpr int V = ITEMS / 4;
    uint4 kv[V];
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const uint32_t idx = wbase + 4u * (uint32_t)(r * 64 + lane);
      kv[r] = idx + 3u < n ? *reinterpret_cast<const uint4*>(keys + idx)
                           : make_uint4(idx < n ? (uint32_t)keys[idx] : 0u, idx + 1u < n ? (uint32_t)keys[idx + 1] : 0u,
                                        idx + 2u < n ? (uint32_t)keys[idx + 2] : 0u, 0u);
    }
    if (bits <= 3) {
      // at most 8 bins (a short top digit): LDS atomics would pile onto a few
      // words, so each lane counts in registers, four 8-bit fields per word
      // (a lane sees at most 4 V <= 16 keys), and the wave adds the counts
      // after widening them to 16-bit fields (a wave sees at most 1024 keys)
      uint32_t lo = 0, hi = 0;
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const uint32_t idx = wbase + 4u * (uint32_t)(r * 64 + lane);
        const uint32_t e[4] = {kv[r].x, kv[r].y, kv[r].z, kv[r].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t d = (e[c] >> shift) & mask;
          const uint32_t inc = idx + (uint32_t)c < n ? 1u << ((d & 3u) * 8u) : 0u;
          lo += d < 4u ? inc : 0u;
          hi += d < 4u ? 0u : inc;
        }
      }
      // 16-bit fields: w0 = bins (0, 2), w1 = (1, 3), w2 = (4, 6), w3 = (5, 7)
      uint32_t cw[4] = {lo & 0x00ff00ffu, (lo >> 8) & 0x00ff00ffu, hi & 0x00ff00ffu, (hi >> 8) & 0x00ff00ffu};
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) cw[k] += __shfl_xor(cw[k], off, 64);
      }
      if (lane < 8) {
        const int k = (lane >> 2) * 2 + (lane & 1);
        const uint32_t word = k == 0 ? cw[0] : (k == 1 ? cw[1] : (k == 2 ? cw[2] : cw[3]));
        const uint32_t cnt = ((lane >> 1) & 1) ? word >> 16 : word & 0xffffu;
        if ((uint32_t)lane <= mask) s_h[w][lane] = cnt;
      }
    } else if (bits <= 4) {
      // few bins (a top digit): the wave's lanes would pile onto a handful of
      // LDS words, so count each bin with a ballot instead; lane b keeps bin b
      uint32_t cnt = 0;
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const uint32_t idx = wbase + 4u * (uint32_t)(r * 64 + lane);
        const uint32_t e[4] = {kv[r].x, kv[r].y, kv[r].z, kv[r].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bool valid = idx + (uint32_t)c < n;
          const uint32_t d = (e[c] >> shift) & mask;
          for (uint32_t b = 0; b <= mask; ++b) {
            const uint32_t m = (uint32_t)__popcll(__ballot(valid && d == b));
            if ((uint32_t)lane == b) cnt += m;
          }
        }
      }
      if ((uint32_t)lane <= mask) s_h[w][lane] = cnt;
    } else {
#pragma unroll
      for (int r = 0; r < V; ++r) {
        const uint32_t idx = wbase + 4u * (uint32_t)(r * 64 + lane);
        const uint32_t e[4] = {kv[r].x, kv[r].y, kv[r].z, kv[r].w};
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (idx + (uint32_t)c < n) atomicAdd(&s_h[w][(e[c] >> shift) & mask], 1u);
      }
    }
  } else {
    K kk[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t idx = wbase + r * 64 + lane;
      kk[r] = idx < n ? keys[idx] : (K)0;
    }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t idx = wbase + r * 64 + lane;
      if (idx < n) atomicAdd(&s_h[w][(uint32_t)(kk[r] >> shift) & mask], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kSortThreads)
    hist[(size_t)d * nb_max + blk] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
}

template <typename K, int ITEMS, int R>
__global__ __launch_bounds__(kSortThreads) void k_sort_upsweep(const K* __restrict__ keys, const uint32_t* n_dev,
                                                               long long cap, int shift, int bits, uint32_t* hist,
                                                               int nb_max) {
  upsweep_block<K, ITEMS, R>(keys, n_dev, cap, shift, bits, hist, nb_max, blockIdx.x);
}

// One workgroup per digit: exclusive scan over the active blocks' counts.
// Each thread takes kScanRun consecutive counts per round (all loaded before
// any is used) and scans them in registers around one workgroup scan of the
// run sums, so a row of up to 256 * kScanRun blocks costs one workgroup scan.
constexpr int kScanRun = 8;
__device__ __forceinline__ void scan_row(uint32_t* hist, uint32_t* totals, const uint32_t* n_dev, long long cap,
                                         int nb_max, int tile, int d) {
  __shared__ uint32_t s_w[4];
  const uint32_t n = n_eff(n_dev, cap);
  const int nb = (int)((n + tile - 1) / tile);
  uint32_t* row = hist + (size_t)d * nb_max;
  const int t = threadIdx.x;
  uint32_t carry = 0;
  for (int b0 = 0; b0 < nb; b0 += 256 * kScanRun) {
    const int first = b0 + t * kScanRun;
    uint32_t v[kScanRun], sum = 0;
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) v[j] = first + j < nb ? row[first + j] : 0u;
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) sum += v[j];
    uint32_t tot;
    const uint32_t incl = block_incl_scan256(sum, s_w, tot);
    uint32_t run = carry + incl - sum;
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) {
      if (first + j < nb) row[first + j] = run;
      run += v[j];
    }
    carry += tot;
    __syncthreads();  // s_w is rewritten by the next round
  }
  if (t == 0) totals[d] = carry;
}

__global__ __launch_bounds__(256) void k_sort_scan(uint32_t* hist, uint32_t* totals, const uint32_t* n_dev,
                                                   long long cap, int nb_max, int tile) {
  scan_row(hist, totals, n_dev, cap, nb_max, tile, blockIdx.x);
}

template <typename K, int ITEMS, int R, bool V>
__device__ __forceinline__ void scatter_block(const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                              K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                              const uint32_t* n_dev, long long cap, int shift, int bits,
                                              const uint32_t* hist, const uint32_t* totals, int nb_max,
                                              uint32_t* ranges, int tshift, uint32_t bid) {
  static_assert(R == kSortThreads, "one digit per thread");
  constexpr int TILE = kSortThreads * ITEMS;
  __shared__ uint32_t s_cnt[4][R];
  __shared__ uint32_t s_scan[R];
  __shared__ uint32_t s_w[4];
  __shared__ K s_key[TILE];
  __shared__ uint32_t s_val[V ? TILE : 1];  // keys-only sorts carry no values
  const uint32_t n = n_eff(n_dev, cap);
  const uint32_t nblk = (n + (uint32_t)TILE - 1) / (uint32_t)TILE;
  if (bid >= nblk) return;
  const uint32_t blk = xcd_remap(bid, nblk);  // the upsweep's mapping
  const uint32_t start = blk * (uint32_t)TILE;
  const int t = threadIdx.x;
  // global base of each digit for this block = exclusive scan of totals + row offset
  {
    const uint32_t v = totals[t];
    uint32_t tot;
    s_scan[t] = block_incl_scan256(v, s_w, tot) - v + hist[(size_t)t * nb_max + blk];
  }
  for (int k = t; k < 4 * R; k += kSortThreads) (&s_cnt[0][0])[k] = 0;
  __syncthreads();
  const int w = t >> 6, lane = t & 63;
  const uint32_t mask = (1u << bits) - 1u;
  const uint32_t wbase = start + (uint32_t)w * (ITEMS * 64);
  K key[ITEMS];
  uint32_t val[ITEMS];
  uint32_t pos[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t idx = wbase + r * 64 + lane;
    const bool valid = idx < n;
    key[r] = valid ? keys_in[idx] : (K)0;
    if constexpr (V) val[r] = valid ? vals_in[idx] : 0u;
    const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
    const uint64_t peers = match_digit(d, valid, bits);
    const uint64_t below = peers & lanemask_lt();
    const uint32_t prior = s_cnt[w][d];
    pos[r] = prior + (uint32_t)__popcll(below);
    if (valid && below == 0) s_cnt[w][d] = prior + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // digit t: its waves' local bases (exclusive scan over digits of the block
  // counts, then over waves) and the offset from local to global position
  const uint32_t c0 = s_cnt[0][t], c1 = s_cnt[1][t], c2 = s_cnt[2][t], c3 = s_cnt[3][t];
  const uint32_t c = c0 + c1 + c2 + c3;
  uint32_t tot;
  const uint32_t lbase = block_incl_scan256(c, s_w, tot) - c;
  s_cnt[0][t] = lbase;
  s_cnt[1][t] = lbase + c0;
  s_cnt[2][t] = lbase + c0 + c1;
  s_cnt[3][t] = lbase + c0 + c1 + c2;
  s_scan[t] = s_scan[t] - lbase;  // global = s_scan[d] + local position
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t idx = wbase + r * 64 + lane;
    if (idx < n) {
      const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
      const uint32_t lp = s_cnt[w][d] + pos[r];
      s_key[lp] = key[r];
      if constexpr (V) s_val[lp] = val[r];
    }
  }
  __syncthreads();
  const uint32_t nloc = min((uint32_t)TILE, n - start);
  for (uint32_t i = t; i < nloc; i += kSortThreads) {
    const K k = s_key[i];
    const uint32_t d = (uint32_t)(k >> shift) & mask;
    const uint32_t dst = s_scan[d] + i;
    keys_out[dst] = k;
    if constexpr (V) vals_out[dst] = s_val[i];
    if (ranges) {
      // last pass of a tile-list sort: within a digit run the block's keys are
      // in final order, so a run of one tile id is contiguous here and in the
      // output; its ends give the tile's [start, end) across blocks by
      // atomic min / max (ranges start as [UINT_MAX, 0))
      const uint32_t tile = (uint32_t)(k >> tshift);
      if (i == 0 || (uint32_t)(s_key[i - 1] >> tshift) != tile) atomicMin(&ranges[2 * tile], dst);
      if (i + 1 == nloc || (uint32_t)(s_key[i + 1] >> tshift) != tile) atomicMax(&ranges[2 * tile + 1], dst + 1);
    }
  }
}

template <typename K, int ITEMS, int R, bool V>
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const K* __restrict__ keys_in,
                                                               const uint32_t* __restrict__ vals_in,
                                                               K* __restrict__ keys_out,
                                                               uint32_t* __restrict__ vals_out,
                                                               const uint32_t* n_dev, long long cap, int shift,
                                                               int bits, const uint32_t* hist,
                                                               const uint32_t* totals, int nb_max, uint32_t* ranges,
                                                               int tshift) {
  scatter_block<K, ITEMS, R, V>(keys_in, vals_in, keys_out, vals_out, n_dev, cap, shift, bits, hist, totals, nb_max,
                                ranges, tshift, blockIdx.x);
}

// Keys per workgroup: short sorts use 1024 (4 per thread) so that they still
// fill the 256 CUs (at 500k keys, 2048 per workgroup sorts slower); long
// ones 2048 (8 per thread).  Measured at 5.7M keys, 8 per thread sorts faster
// than 16 or 12: the scans are cheap enough that more, smaller workgroups pay.
int sort_tile(long long capacity) { return capacity <= (1LL << 21) ? kSortThreads * 4 : kSortThreads * 8; }

size_t sort_tmp_bytes(long long capacity) {
  const long long tile = sort_tile(capacity);
  const long long nb = (capacity + tile - 1) / tile;
  return align_up((size_t)kRadix * 4 * (size_t)(nb > 0 ? nb : 1), 256) + kRadix * 4;
}

template <typename K, int ITEMS, int R>
void sort_pass(const K* kin, const uint32_t* vin, K* kout, uint32_t* vout, const uint32_t* n_dev, long long capacity,
               int shift, int bits, uint32_t* hist, uint32_t* totals, int nb_max, int tile, hipStream_t s,
               uint32_t* ranges, int tshift) {
  hipLaunchKernelGGL((k_sort_upsweep<K, ITEMS, R>), dim3(nb_max), dim3(kSortThreads), 0, s, kin, n_dev, capacity,
                     shift, bits, hist, nb_max);
  hipLaunchKernelGGL(k_sort_scan, dim3(R), dim3(256), 0, s, hist, totals, n_dev, capacity, nb_max, tile);
  if (vin)
    hipLaunchKernelGGL((k_sort_scatter<K, ITEMS, R, true>), dim3(nb_max), dim3(kSortThreads), 0, s, kin, vin, kout,
                       vout, n_dev, capacity, shift, bits, hist, totals, nb_max, ranges, tshift);
  else
    hipLaunchKernelGGL((k_sort_scatter<K, ITEMS, R, false>), dim3(nb_max), dim3(kSortThreads), 0, s, kin, vin, kout,
                       vout, n_dev, capacity, shift, bits, hist, totals, nb_max, ranges, tshift);
}

// Stable LSD sort on key bits [start_bit, end_bit); vals_a == nullptr sorts
// the keys alone.  With `ranges` (tile lists: key >> start_bit is a tile id),
// the last pass also records every tile's [start, end) into ranges, which the
// caller initialises to [UINT_MAX, 0) per tile (empty tiles keep that).
template <typename K>
hipError_t sort_pairs(K* keys_a, uint32_t* vals_a, K* keys_b, uint32_t* vals_b, const uint32_t* n_dev,
                      long long capacity, int start_bit, int end_bit, void* tmp, hipStream_t s, int* result_in_b,
                      uint32_t* ranges = nullptr) {
  const int passes = sort_passes(end_bit - start_bit);
  *result_in_b = passes & 1;
  if (capacity <= 0) return hipSuccess;
  const int tile = sort_tile(capacity);
  const int nb_max = (int)((capacity + tile - 1) / tile);
  uint32_t* hist = (uint32_t*)tmp;
  uint32_t* totals = hist + align_up((size_t)kRadix * (size_t)nb_max, 64);
  K* kin = keys_a;
  uint32_t* vin = vals_a;
  K* kout = keys_b;
  uint32_t* vout = vals_b;
  for (int p = 0; p < passes; ++p) {
    const int shift = start_bit + kRadixBits * p;
    const int bits = end_bit - shift < kRadixBits ? end_bit - shift : kRadixBits;
    auto pass = tile == kSortThreads * 4 ? sort_pass<K, 4, kRadix> : sort_pass<K, 8, kRadix>;
    pass(kin, vin, kout, vout, n_dev, capacity, shift, bits, hist, totals, nb_max, tile, s,
         p + 1 == passes ? ranges : nullptr, start_bit);
    K* tk = kin;
    kin = kout;
    kout = tk;
    uint32_t* tv = vin;
    vin = vout;
    vout = tv;
  }
  return hipGetLastError();
}

// ---- segmented sort ---------------------------------------------------------
// Descriptor, pass by pass: uint32 prefix[n + 1] (each segment's upsweep /
// scatter workgroups in that pass, rounded up to a multiple of 8 so that a
// segment's first block starts a round of the XCD round-robin, as a launch
// of its own does; 0 once the segment has no pass left), then one SegPass
// per segment.  Every segment sorts with 1024-key workgroups (sort_tile of a
// capacity up to kSegMaxKeys).
struct SegPass {
  const uint32_t* kin;
  const uint32_t* vin;
  uint32_t* kout;
  uint32_t* vout;
  const uint32_t* n_dev;
  uint32_t* hist;
  uint32_t* totals;
  uint32_t* ranges;
  long long cap;
  int nb_max;  // 0: the segment has no such pass
  int shift;
  int bits;
  int tshift;
};
constexpr int kSegItems = 4;

static size_t seg_pass_bytes(int n) {
  return align_up(sizeof(uint32_t) * (size_t)(n + 1), 256) + align_up(sizeof(SegPass) * (size_t)n, 256);
}

__device__ __forceinline__ int seg_of(const uint32_t* prefix, int n, uint32_t b) {
  // the last segment whose range starts at or before b (empty ranges are skipped)
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kSortThreads) void k_seg_upsweep(const uint32_t* prefix, const SegPass* ps, int n) {
  const uint32_t b = blockIdx.x;
  const int f = seg_of(prefix, n, b);
  const SegPass p = ps[f];
  upsweep_block<uint32_t, kSegItems, kRadix>(p.kin, p.n_dev, p.cap, p.shift, p.bits, p.hist, p.nb_max,
                                             b - prefix[f]);
}

// 256 workgroups per segment, one per digit.
__global__ __launch_bounds__(256) void k_seg_scan(const SegPass* ps) {
  const SegPass p = ps[blockIdx.x >> kRadixBits];
  if (p.nb_max == 0) return;
  scan_row(p.hist, p.totals, p.n_dev, p.cap, p.nb_max, kSortThreads * kSegItems, (int)(blockIdx.x & (kRadix - 1)));
}

template <bool V>
__global__ __launch_bounds__(kSortThreads) void k_seg_scatter(const uint32_t* prefix, const SegPass* ps, int n) {
  const uint32_t b = blockIdx.x;
  const int f = seg_of(prefix, n, b);
  const SegPass p = ps[f];
  scatter_block<uint32_t, kSegItems, kRadix, V>(p.kin, p.vin, p.kout, p.vout, p.n_dev, p.cap, p.shift, p.bits,
                                                p.hist, p.totals, p.nb_max, p.ranges, p.tshift, b - prefix[f]);
}

size_t seg_sort_desc_bytes(int n, int passes) {
  if (n <= 0 || passes < 1 || passes > kSegMaxPasses) return 0;
  return seg_pass_bytes(n) * (size_t)passes;
}

bool seg_sort_build(int n, const SortSeg* segs, void* host, size_t bytes, unsigned* blocks, int* passes) {
  if (n <= 0 || !segs || !host || !blocks || !passes) return false;
  int np = 0;
  const bool values = segs[0].vals_a != nullptr;
  for (int f = 0; f < n; ++f) {
    const SortSeg& g = segs[f];
    if (g.capacity <= 0 || g.capacity > kSegMaxKeys || !g.keys_a || !g.keys_b || !g.n_dev || !g.tmp) return false;
    if ((g.vals_a != nullptr) != values || (values && !g.vals_b)) return false;
    if (g.start_bit < 0 || g.end_bit <= g.start_bit || g.end_bit > 32) return false;
    if (((uintptr_t)g.keys_a | (uintptr_t)g.keys_b) & 15) return false;  // 16-byte key loads
    if (sort_tile(g.capacity) != kSortThreads * kSegItems) return false;
    np = std::max(np, sort_passes(g.end_bit - g.start_bit));
  }
  if (np > kSegMaxPasses || bytes < seg_sort_desc_bytes(n, np)) return false;
  for (int p = 0; p < kSegMaxPasses; ++p) blocks[p] = 0;
  for (int p = 0; p < np; ++p) {
    char* base = (char*)host + seg_pass_bytes(n) * (size_t)p;
    uint32_t* prefix = (uint32_t*)base;
    SegPass* ps = (SegPass*)(base + align_up(sizeof(uint32_t) * (size_t)(n + 1), 256));
    unsigned long long total = 0;
    for (int f = 0; f < n; ++f) {
      const SortSeg& g = segs[f];
      const int tile = kSortThreads * kSegItems;
      const int nb_max = (int)((g.capacity + tile - 1) / tile);
      const int gp = sort_passes(g.end_bit - g.start_bit);
      uint32_t* hist = (uint32_t*)g.tmp;
      SegPass& q = ps[f];
      const bool odd = p & 1;
      q.kin = odd ? g.keys_b : g.keys_a;
      q.vin = odd ? g.vals_b : g.vals_a;
      q.kout = odd ? g.keys_a : g.keys_b;
      q.vout = odd ? g.vals_a : g.vals_b;
      q.n_dev = g.n_dev;
      q.hist = hist;
      q.totals = hist + align_up((size_t)kRadix * (size_t)nb_max, 64);  // sort_pairs' layout of tmp
      q.ranges = p + 1 == gp ? g.ranges : nullptr;
      q.cap = g.capacity;
      q.nb_max = p < gp ? nb_max : 0;
      q.shift = g.start_bit + kRadixBits * p;
      q.bits = g.end_bit - q.shift < kRadixBits ? g.end_bit - q.shift : kRadixBits;
      q.tshift = g.start_bit;
      prefix[f] = (uint32_t)total;
      total += p < gp ? align_up((size_t)nb_max, 8) : 0;
      if (total > 0x7fffffffULL) return false;
    }
    prefix[n] = (uint32_t)total;
    blocks[p] = (unsigned)total;
  }
  *passes = np;
  return true;
}

hipError_t seg_sort_launch(int n, const void* dev_desc, const unsigned* blocks, bool values, hipStream_t s) {
  for (int p = 0; p < kSegMaxPasses && blocks[p] > 0; ++p) {
    const char* base = (const char*)dev_desc + seg_pass_bytes(n) * (size_t)p;
    const uint32_t* prefix = (const uint32_t*)base;
    const SegPass* q = (const SegPass*)(base + align_up(sizeof(uint32_t) * (size_t)(n + 1), 256));
    hipLaunchKernelGGL(k_seg_upsweep, dim3(blocks[p]), dim3(kSortThreads), 0, s, prefix, q, n);
    hipLaunchKernelGGL(k_seg_scan, dim3((unsigned)n * kRadix), dim3(256), 0, s, q);
    if (values)
      hipLaunchKernelGGL(k_seg_scatter<true>, dim3(blocks[p]), dim3(kSortThreads), 0, s, prefix, q, n);
    else
      hipLaunchKernelGGL(k_seg_scatter<false>, dim3(blocks[p]), dim3(kSortThreads), 0, s, prefix, q, n);
  }
  return hipGetLastError();
}

hipError_t sort_pairs_u64(uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b,
                          const uint32_t* n_dev, long long capacity, int end_bit, void* tmp, hipStream_t s,
                          int* result_in_b) {
  return sort_pairs<uint64_t>(keys_a, vals_a, keys_b, vals_b, n_dev, capacity, 0, end_bit, tmp, s, result_in_b);
}

hipError_t sort_pairs_u32(uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b, uint32_t* vals_b,
                          const uint32_t* n_dev, long long capacity, int end_bit, void* tmp, hipStream_t s,
                          int* result_in_b, uint32_t* ranges) {
  return sort_pairs<uint32_t>(keys_a, vals_a, keys_b, vals_b, n_dev, capacity, 0, end_bit, tmp, s, result_in_b,
                              ranges);
}

hipError_t sort_keys_u32(uint32_t* keys_a, uint32_t* keys_b, const uint32_t* n_dev, long long capacity, int start_bit,
                         int end_bit, void* tmp, hipStream_t s, int* result_in_b, uint32_t* ranges) {
  return sort_pairs<uint32_t>(keys_a, nullptr, keys_b, nullptr, n_dev, capacity, start_bit, end_bit, tmp, s,
                              result_in_b, ranges);
}

}  // namespace gs

extern "C" size_t gs_sort_tmp_bytes(long long capacity) { return gs::sort_tmp_bytes(capacity); }

extern "C" int gs_sort_pairs_u64(uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b,
                                 const uint32_t* n_dev, long long capacity, int end_bit, void* tmp,
                                 size_t tmp_bytes, int* result_in_b, void* stream) {
  if (capacity < 0 || capacity > 0x7fffffffLL || end_bit < 1 || end_bit > 64 || !n_dev || !result_in_b)
    return GS_E_ARG;
  if (capacity > 0 && (!keys_a || !vals_a || !keys_b || !vals_b || !tmp)) return GS_E_ARG;
  if (tmp_bytes < gs::sort_tmp_bytes(capacity)) return GS_E_WORKSPACE;
  const hipError_t e = gs::sort_pairs_u64(keys_a, vals_a, keys_b, vals_b, n_dev, capacity, end_bit, tmp,
                                          (hipStream_t)stream, result_in_b);
  return e == hipSuccess ? GS_OK : GS_E_LAUNCH;
}

extern "C" int gs_sort_pairs_u32(uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b, uint32_t* vals_b,
                                 const uint32_t* n_dev, long long capacity, int end_bit, void* tmp,
                                 size_t tmp_bytes, int* result_in_b, void* stream) {
  if (capacity < 0 || capacity > 0x7fffffffLL || end_bit < 1 || end_bit > 32 || !n_dev || !result_in_b)
    return GS_E_ARG;
  if (capacity > 0 && (!keys_a || !vals_a || !keys_b || !vals_b || !tmp)) return GS_E_ARG;
  if (((uintptr_t)keys_a | (uintptr_t)keys_b) & 15) return GS_E_ARG;  // 16-byte key loads
  if (tmp_bytes < gs::sort_tmp_bytes(capacity)) return GS_E_WORKSPACE;
  const hipError_t e = gs::sort_pairs_u32(keys_a, vals_a, keys_b, vals_b, n_dev, capacity, end_bit, tmp,
                                          (hipStream_t)stream, result_in_b);
  return e == hipSuccess ? GS_OK : GS_E_LAUNCH;
}
