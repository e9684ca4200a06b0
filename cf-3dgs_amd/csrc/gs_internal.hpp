// gs_internal.hpp -- host-side declarations shared between the translation
// units of libgs_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gs {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per sort workgroup
constexpr int kRadixBits = 8;  // digit width of the LSD passes: 256 buckets
constexpr int kRadix = 1 << kRadixBits;

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline int bits_for(uint32_t n) {  // number of bits needed to hold values < n
  int b = 0;
  while (b < 32 && (1ull << b) < (uint64_t)n) ++b;
  return b;
}
// LSD passes over bits [0, end_bit), kRadixBits at a time.
inline int sort_passes(int end_bit) { return (end_bit + kRadixBits - 1) / kRadixBits; }

namespace timing {
enum Stage { kPreprocess = 0, kScan, kDuplicate, kSort, kRanges, kBlend, kRenderBwd, kPreprocessBwd, kKnn, kZero };
// RAII pair of events around the launches issued inside the scope (no-op
// unless gs_timing_enable(1) was called).
class Scope {
 public:
  Scope(int stage, hipStream_t s);
  ~Scope();
  Scope(const Scope&) = delete;
  Scope& operator=(const Scope&) = delete;

 private:
  int stage_;
  hipStream_t s_;
  void* a_;
};
bool enabled();
}  // namespace timing

size_t sort_tmp_bytes(long long capacity);
// Returns hipSuccess or the launch error.  *result_in_b tells where the
// sorted data ends up.
hipError_t sort_pairs_u64(uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b,
                          const uint32_t* n_dev, long long capacity, int end_bit, void* tmp, hipStream_t s,
                          int* result_in_b);
// keys only, on key bits [start_bit, end_bit).  The 32-bit sorts of tile
// lists may also record each tile's [start, end) of the sorted list in
// `ranges` (key >> start_bit is the tile id; the caller sets every range to
// [UINT_MAX, 0) beforehand, and empty tiles keep that).
hipError_t sort_keys_u32(uint32_t* keys_a, uint32_t* keys_b, const uint32_t* n_dev, long long capacity, int start_bit,
                         int end_bit, void* tmp, hipStream_t s, int* result_in_b, uint32_t* ranges = nullptr);
hipError_t sort_pairs_u32(uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b, uint32_t* vals_b,
                          const uint32_t* n_dev, long long capacity, int end_bit, void* tmp, hipStream_t s,
                          int* result_in_b, uint32_t* ranges = nullptr);

// A segmented sort: many independent 32-bit sorts, each on its own buffers
// exactly as sort_pairs_u32 / sort_keys_u32 would run it (same passes, same
// workgroup tiles, the result in the same buffer, the same tile ranges), but
// each kernel of a pass is one launch over every segment.  All segments carry
// values, or none does.  Segments of at most kSegMaxKeys keys.
struct SortSeg {
  uint32_t* keys_a;
  uint32_t* vals_a;  // nullptr: keys only
  uint32_t* keys_b;
  uint32_t* vals_b;
  const uint32_t* n_dev;
  long long capacity;
  void* tmp;  // sort_tmp_bytes(capacity)
  int start_bit, end_bit;  // key bits sorted on
  uint32_t* ranges;        // the last pass records the tile ranges (key >> start_bit), or nullptr
};
constexpr long long kSegMaxKeys = 1LL << 21;
constexpr int kSegMaxPasses = 4;
size_t seg_sort_desc_bytes(int n, int passes);
// Host-side descriptor of the segments' passes (its pointers are the
// segments' device buffers); the caller copies it to the device.  Returns
// false on a segment it cannot take.  blocks[p] is pass p's upsweep /
// scatter grid (0 past the longest segment's passes); *passes the pass count.
bool seg_sort_build(int n, const SortSeg* segs, void* host, size_t bytes, unsigned* blocks, int* passes);
hipError_t seg_sort_launch(int n, const void* dev_desc, const unsigned* blocks, bool values, hipStream_t s);

}  // namespace gs
