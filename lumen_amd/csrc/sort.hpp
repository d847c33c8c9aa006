// sort.hpp -- stable LSD radix sort of (uint64 key, uint32 value) pairs with
// 8-bit digits, the device form of radix_sort_64bit
// (src/raster/radix_sort.mojo:39-87).
//
// One pass per digit that is not constant across all keys (a stable pass over
// a constant digit is the identity, so skipping it leaves the result
// unchanged).  Each pass:
//   k_sort_hist     per-block 256-bin histogram of the digit (LDS atomics on
//                   per-wave sub-histograms), stored digit-major so that one
//                   exclusive scan yields every block's start offset per digit
//   exclusive_scan  over the 256 x blocks table
//   k_sort_scatter  stable rank inside the block: elements go round by round
//                   (256 consecutive elements per round), a wave ranks its 64
//                   lanes with 8 ballots (lanes with equal digits = peers,
//                   rank = peers below this lane), waves are ordered through
//                   LDS counters, so equal digits keep their input order.
#pragma once

#include "common.hpp"
#include "scan.hpp"

namespace lumen {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint64_t* __restrict__ keys,
                                                           uint32_t m, int shift,
                                                           uint32_t* __restrict__ hist,
                                                           uint32_t nb) {
  __shared__ uint32_t h[4][256];
  const int t = threadIdx.x;
  const int w = t >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) h[i][t] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t idx = base + j * kSortThreads + t;
    if (idx < m) atomicAdd(&h[w][(uint32_t)(keys[idx] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[(uint32_t)t * nb + blockIdx.x] = h[0][t] + h[1][t] + h[2][t] + h[3][t];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t m, int shift, const uint32_t* __restrict__ offsets,
    uint32_t nb) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const int t = threadIdx.x;
  const int w = t >> 6;
  const uint64_t lt = lanemask_lt();
  run[t] = offsets[(uint32_t)t * nb + blockIdx.x];
  const uint32_t base = blockIdx.x * kSortTile;
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t round = base + j * kSortThreads;
    if (round >= m) break;  // uniform across the block
#pragma unroll
    for (int i = 0; i < 4; ++i) wc[i][t] = 0;
    __syncthreads();
    const uint32_t idx = round + t;
    const bool valid = idx < m;
    const uint64_t k = valid ? kin[idx] : 0ull;
    const uint32_t v = valid ? vin[idx] : 0u;
    const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      const uint32_t c0 = wc[0][t], c1 = wc[1][t], c2 = wc[2][t], c3 = wc[3][t];
      const uint32_t r0 = run[t];
      wc[0][t] = r0;
      wc[1][t] = r0 + c0;
      wc[2][t] = r0 + c0 + c1;
      wc[3][t] = r0 + c0 + c1 + c2;
      run[t] = r0 + c0 + c1 + c2 + c3;
    }
    __syncthreads();
    if (valid) {
      const uint32_t dst = wc[w][d] + rank;
      kout[dst] = k;
      vout[dst] = v;
    }
    __syncthreads();
  }
}

struct SortScratch {
  DevBuf* hist;
  DevBuf* hist_part;
};

// Sorts (k[src], v[src]) over the given digit list, ping-ponging between the
// two buffers; *final_buf receives the index of the buffer holding the result.
inline LumenStatus radix_sort_digits(uint64_t* k[2], uint32_t* v[2], uint32_t m, const int* digits,
                                     int ndigits, SortScratch sc, int* final_buf, hipStream_t st) {
  int src = 0;
  if (m > 1 && ndigits > 0) {
    const uint32_t nb = div_up(m, kSortTile);
    const uint64_t hlen = (uint64_t)nb * 256u;
    LUMEN_TRY(ensure(*sc.hist, hlen * sizeof(uint32_t)));
    LUMEN_TRY(ensure(*sc.hist_part, (uint64_t)div_up(hlen, kScanChunk) * sizeof(uint32_t)));
    uint32_t* hist = sc.hist->as<uint32_t>();
    uint32_t* part = sc.hist_part->as<uint32_t>();
    for (int p = 0; p < ndigits; ++p) {
      const int shift = 8 * digits[p];
      hipLaunchKernelGGL(k_sort_hist, dim3(nb), dim3(kSortThreads), 0, st, k[src], m, shift, hist,
                         nb);
      LUMEN_LAUNCHED();
      LUMEN_TRY(exclusive_scan<uint32_t>(hist, hist, hlen, part, nullptr, st));
      hipLaunchKernelGGL(k_sort_scatter, dim3(nb), dim3(kSortThreads), 0, st, k[src], v[src],
                         k[src ^ 1], v[src ^ 1], m, shift, hist, nb);
      LUMEN_LAUNCHED();
      src ^= 1;
    }
  }
  *final_buf = src;
  return LUMEN_SUCCESS;
}

// Digits that can differ between keys lying in [lo, hi]: those at or below the
// highest bit in which lo and hi differ (all keys in the interval share the
// bits above it).  Appends digit indices (offset by digit_base) to out.
inline int varying_digits(uint64_t lo, uint64_t hi, int digit_base, int max_digits, int* out) {
  uint64_t x = lo ^ hi;
  int n = 0;
  if (x == 0) return 0;
  int top = 63 - __builtin_clzll(x);
  for (int d = 0; d <= top / 8 && d < max_digits; ++d) out[n++] = digit_base + d;
  return n;
}

}  // namespace lumen
