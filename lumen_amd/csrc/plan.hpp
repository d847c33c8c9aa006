// plan.hpp -- the device-side description of a frame's sort.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace lumen {

// Written on the device by the frame's planning kernel (or by the host path
// of lumen_ex_radix_sort), read by every later kernel of the frame, so that a
// frame needs no host round trip once its buffers are sized.
struct SortPlan {
  int32_t passes;      // digit passes to run
  int32_t digit[8];    // digit index (0 = bits 0..7) of each pass
  uint32_t m;          // keys to sort (tile instances of the frame)
  uint32_t v;          // visible splats of the frame
  uint32_t overflow;   // set when the frame exceeded a reserved capacity
  uint32_t pad[4];
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

}  // namespace lumen
