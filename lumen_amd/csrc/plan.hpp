// plan.hpp -- the device-side description of a frame's sort.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace lumen {

constexpr int kMaxDigitBits = 9;                 // widest digit a pass sorts
constexpr int kDigitBins = 1 << kMaxDigitBits;   // histogram rows per block

// Keys per thread of a sort block (256 threads).  Up to kSmallSortKeys keys
// a frame sorts 12 per thread: more, shorter blocks fill the GPU better
// (1M Gaussians at 1080p, 6.3M keys: 8 us less per frame than 16).  Larger
// frames sort 16, which writes a quarter fewer histogram rows per key (518M
// keys: 1.4 ms less per frame than 12).  The host and the planning kernel
// both decide with this function from the same key count.
constexpr uint64_t kSmallSortKeys = 1ull << 25;
__host__ __device__ inline int sort_items_for(uint64_t keys) {
  return keys <= kSmallSortKeys ? 12 : 16;
}

// Written on the device by the frame's planning kernel (or by the host path
// of lumen_ex_radix_sort), read by every later kernel of the frame, so that a
// frame needs no host round trip once its buffers are sized.
//
// Frame keys hold the tile id within the key's view (the sort keeps views
// apart, see SortSeg).  mode 0 sorts the 64-bit keys tile << 32 | depth_key
// in 8-bit digits.  mode 1 sorts 32-bit keys (tile - tmin) << dbits |
// (depth_key - dmin),
// used when the tile ids and depth keys of the frame vary in at most 32 bits
// together: both transforms are monotone and the fields do not overlap, so
// equal 64-bit keys map to equal frame keys and the stable order is the same.
struct SortPlan {
  int32_t passes;      // digit passes to run
  int32_t shift[8];    // first key bit of each pass's digit
  int32_t width[8];    // digit width in bits (<= kMaxDigitBits)
  uint32_t m;          // keys to sort (tile instances of the frame)
  uint32_t v;          // visible splats of the frame
  uint32_t overflow;   // set when the frame exceeded a reserved capacity
  int32_t mode;        // 0: 64-bit keys, 1: 32-bit frame keys
  uint32_t tmin, dmin; // frame-key offsets (mode 1)
  int32_t dbits;       // depth bits of a frame key (mode 1)
  uint32_t blocks;     // sort blocks over all segments (the histogram table's rows / 512)
  int32_t items;       // keys per thread of a sort block (sort_items_for)
  int32_t first;       // passes run before the sort: 1 when k_emit wrote its keys
                       // already ordered by pass 0's digit (see plan_frame), else 0
};

// One segment of a frame's keys: a view's contiguous run (emission is
// view-major) sorted on its own.  The sort's blocks are numbered through the
// segments (block_base), each block ranking keys of one segment only, so
// the view grouping survives every pass and the keys need only the view's
// own tile ids.  The histogram table is segment-major, then digit, then the
// segment's blocks, so one flat exclusive scan gives every (segment, digit,
// block) its output offset.
struct SortSeg {
  uint32_t key_base;    // first key of the segment
  uint32_t count;       // keys in the segment
  uint32_t block_base;  // first sort block of the segment
  uint32_t nblocks;     // sort blocks of the segment
};

// Digits that can differ between keys lying in [lo, hi]: those at or below the
// highest bit in which lo and hi differ (all keys in the interval share the
// bits above it).  Appends digit indices (offset by digit_base) to out.
__host__ __device__ inline int varying_digits(uint64_t lo, uint64_t hi, int digit_base,
                                              int max_digits, int32_t* out) {
  uint64_t x = lo ^ hi;
  if (x == 0) return 0;
  int top = 63;
  while (!((x >> top) & 1ull)) --top;
  int n = 0;
  for (int d = 0; d <= top / 8 && d < max_digits; ++d) out[n++] = digit_base + d;
  return n;
}

// How many digits varying_digits appends for [lo, hi].
__host__ __device__ inline int varying_count(uint64_t lo, uint64_t hi, int max_digits) {
  const uint64_t x = lo ^ hi;
  if (x == 0) return 0;
  int top = 63;
  while (!((x >> top) & 1ull)) --top;
  const int n = top / 8 + 1;
  return n < max_digits ? n : max_digits;
}

__host__ __device__ inline int bit_width(uint32_t x) {
  int n = 0;
  while (x) {
    ++n;
    x >>= 1;
  }
  return n;
}

// The sort of a frame whose depth keys lie in [dmin, dmax] and tile ids in
// [tmin, tmax] (nothing to sort when there is no key): mode 1 with
// ceil(bits / 9) passes of equal width when the two ranges need at most 32
// bits, else mode 0 over the varying 8-bit digits of the 64-bit key.  Fills
// passes, shift, width, mode, tmin, dmin, dbits.
__host__ __device__ inline void plan_digits(bool any, uint32_t dmin, uint32_t dmax, uint32_t tmin,
                                            uint32_t tmax, SortPlan& p) {
  // (every array index below is a constant once the loops unroll, so a
  // device copy keeps the plan in registers)
#pragma unroll
  for (int k = 0; k < 8; ++k) p.shift[k] = p.width[k] = 0;
  p.passes = 0;
  p.first = 0;
  p.mode = 1;
  p.tmin = tmin;
  p.dmin = dmin;
  p.dbits = 0;
  if (!any) return;
  const int db = bit_width(dmax - dmin), tb = bit_width(tmax - tmin);
  if (db + tb <= 32) {
    const int total = db + tb;
    const int np = (total + kMaxDigitBits - 1) / kMaxDigitBits;
    const int w = np ? (total + np - 1) / np : 0;
    p.dbits = db;
    p.passes = np;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k < np) {
        p.shift[k] = k * w;
        p.width[k] = total - k * w < w ? total - k * w : w;
      }
    }
    return;
  }
  // the varying 8-bit digits of the depth half (digits 0-3), then of the
  // tile half (digits 4-7), as varying_digits lists them
  p.mode = 0;
  const int nd = varying_count(dmin, dmax, 4), nt = varying_count(tmin, tmax, 4);
  p.passes = nd + nt;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k < nd + nt) {
      p.shift[k] = 8 * (k < nd ? k : 4 + (k - nd));
      p.width[k] = 8;
    }
  }
}

// The digit k_emit sorts by when it runs a frame's first pass itself: the
// depth key's low 9 bits.  All keys of one splat share their depth key, so
// k_count can histogram this digit per block of Gaussians from the splats'
// tile counts, and k_emit can write each splat's run of keys straight to its
// place in the pass's output (pass 0 of an LSD sort over these bits).
constexpr int kEmitDigitBits = kMaxDigitBits;
constexpr uint32_t kEmitDigitMask = (1u << kEmitDigitBits) - 1u;

// The frame's sort: plan_digits, or, when `fuse` (the frame's k_count wrote
// the per-block digit histograms) and it saves a pass, the plan whose pass 0
// k_emit runs (first = 1).  That plan keeps 32-bit frame keys with the depth
// offset rounded down to a multiple of 512 and at least 9 depth bits, so key
// bits 0-8 are exactly (depth key & 511); the later passes cover bits 9 up in
// equal digits of at most 9 bits.  Every plan sorts the same keys stably, so
// all give the same order; the host and k_plan call this with the same inputs.
__host__ __device__ inline void plan_frame(bool any, uint32_t dmin, uint32_t dmax, uint32_t tmin,
                                           uint32_t tmax, bool fuse, SortPlan& p) {
  plan_digits(any, dmin, dmax, tmin, tmax, p);
  if (!fuse || !any || p.mode != 1) return;
  const uint32_t da = dmin & ~kEmitDigitMask;
  int db = bit_width(dmax - da);
  if (db < kEmitDigitBits) db = kEmitDigitBits;
  const int tb = bit_width(tmax - tmin);
  if (db + tb > 32) return;
  const int rest = db + tb - kEmitDigitBits;
  const int np = (rest + kMaxDigitBits - 1) / kMaxDigitBits;
  if (np >= p.passes) return;  // no pass saved
  const int w = np ? (rest + np - 1) / np : 0;
  p.dmin = da;
  p.dbits = db;
  p.passes = 1 + np;
  p.first = 1;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k == 0) {
      p.shift[k] = 0;
      p.width[k] = kEmitDigitBits;
    } else if (k <= np) {
      const int s = (k - 1) * w;
      p.shift[k] = kEmitDigitBits + s;
      p.width[k] = rest - s < w ? rest - s : w;
    } else {
      p.shift[k] = p.width[k] = 0;
    }
  }
}

// The view-local tile of a frame key: mode 1 (k >> dbits) + tmin, mode 0
// the high half.
__host__ __device__ inline uint32_t frame_key_tile(uint64_t k, const SortPlan& p) {
  if (p.mode != 1) return (uint32_t)(k >> 32);
  return (p.dbits < 32 ? ((uint32_t)k >> p.dbits) : 0u) + p.tmin;
}

// A frame key of a view whose tiles start at tile_base back to the 64-bit
// key tile << 32 | depth key of the reference.
__host__ __device__ inline uint64_t frame_key_to_key(uint64_t k, const SortPlan& p,
                                                     uint32_t tile_base) {
  const uint32_t tile = frame_key_tile(k, p) + tile_base;
  if (p.mode != 1) return ((uint64_t)tile << 32) | (k & 0xFFFFFFFFull);
  const uint32_t dmask = p.dbits ? (0xFFFFFFFFu >> (32 - p.dbits)) : 0u;
  return ((uint64_t)tile << 32) | (uint64_t)(((uint32_t)k & dmask) + p.dmin);
}

}  // namespace lumen
