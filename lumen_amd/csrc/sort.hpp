// sort.hpp -- stable LSD radix sort of (key, uint32 value) pairs, the device
// form of radix_sort_64bit (src/raster/radix_sort.mojo:39-87): 64-bit keys
// in 8-bit digits, or a frame's 32-bit keys in digits of up to 9 bits (see
// SortPlan).
//
// Only digits that can differ between keys are sorted: a stable pass over a
// digit that is equal in all keys is the identity.  The digit plan, the
// number of keys and the buffer holding the result live in device memory
// (SortPlan), so a frame's whole sort can be captured in a hipGraph with no
// host round trip; grids are sized for a capacity and surplus blocks exit.
// Each pass:
//   k_sort_hist     per-block histogram of the digit (LDS atomics on
//                   per-wave sub-histograms), stored segment-major then
//                   digit-major (SortSeg) so that one exclusive scan gives
//                   every block's start offset per digit
//   k_sort_colscan  one block per (segment, digit): the exclusive prefix of
//                   the digit's column over the segment's blocks, in place,
//                   and the column's total
//   k_sort_scatter  stable rank inside the block (256 threads of 12 or 16
//                   keys, sort_items_for): each wave owns a quarter of the
//                   block's consecutive keys and ranks them 64 at a time with one
//                   ballot per digit bit (lanes with equal digits are peers; rank = peers
//                   below this lane) and wave-private digit counters in LDS;
//                   one block scan then orders the waves, the block's keys are
//                   staged in LDS in digit order and written out so that each
//                   digit's run lands on consecutive addresses (coalesced).
//                   Equal digits keep their input order.  A digit's output
//                   start is the segment's first key, plus the totals of the
//                   segment's smaller digits (one block scan over the column
//                   totals), plus the column prefix: the offset an exclusive
//                   scan of the whole segment-major table would give.
#pragma once

#include "common.hpp"
#include "plan.hpp"
#include "scan.hpp"

namespace lumen {

constexpr int kSortThreads = 256;

// The segment and segment-local index of the sort block that grid block i
// runs; false for a surplus block of a capacity-sized grid.  Segments
// without keys have no blocks (the search takes the last segment starting at
// or before the block).
__device__ __forceinline__ bool sort_block(const SortSeg* __restrict__ segs, int nseg, uint32_t i,
                                           SortSeg& sg, uint32_t& lb, int* seg_out = nullptr) {
  const SortSeg last = segs[nseg - 1];
  const uint32_t n = last.block_base + last.nblocks;
  if (i >= n) return false;
  const uint32_t b = xcd_block(i, n);
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].block_base <= b) lo = mid;
    else hi = mid - 1;
  }
  sg = segs[lo];
  lb = b - sg.block_base;
  if (seg_out) *seg_out = lo;
  return true;
}

// Position of (digit d, block lb) of segment sg in the histogram table.
__device__ __forceinline__ size_t hist_index(const SortSeg& sg, uint32_t d, uint32_t lb) {
  return (size_t)sg.block_base * kDigitBins + (size_t)d * sg.nblocks + lb;
}

// K = uint64_t sorts mode-0 plans (64-bit keys), K = uint32_t mode-1 plans
// (32-bit frame keys held in the same buffers); a kernel of the other mode
// returns at once, so a graph can hold both.
template <typename K>
__device__ __forceinline__ bool sort_mode_is(const SortPlan* plan) {
  return plan->mode == (sizeof(K) == 4 ? 1 : 0);
}

template <typename K, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const K* k0, const K* k1,
                                                           const SortPlan* __restrict__ plan,
                                                           const SortSeg* __restrict__ segs,
                                                           int nseg, int pass,
                                                           uint32_t* __restrict__ hist) {
  constexpr int kTile = kSortThreads * ITEMS;
  __shared__ uint32_t h[4][kDigitBins];
  // the segment lookup reads only segs: it is issued with the plan's loads
  SortSeg sg;
  uint32_t lb;
  const bool mine = sort_block(segs, nseg, blockIdx.x, sg, lb);
  if (pass >= plan->passes || pass < plan->first || !sort_mode_is<K>(plan) ||
      plan->items != ITEMS)
    return;
  if (!mine) return;  // surplus block
  const int t = threadIdx.x;
  const K* __restrict__ keys = (pass & 1) ? k1 : k0;
  const int shift = plan->shift[pass];
  const uint32_t mask = (1u << plan->width[pass]) - 1u;
  const int w = t >> 6;
  for (int d = t; d < 4 * kDigitBins; d += kSortThreads) (&h[0][0])[d] = 0u;
  __syncthreads();
  const uint32_t base = sg.key_base + lb * kTile;
  const uint32_t end = sg.key_base + sg.count;
  // all loads first (a load feeding an atomic in the same step would wait
  // for each key in turn)
  K kk[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t idx = base + j * kSortThreads + t;
    kk[j] = idx < end ? keys[idx] : K(0);
  }
  // A splat's keys are consecutive and share their depth digits, so
  // neighbouring lanes often hold the same digit: each run of equal digits
  // in neighbouring lanes adds its length once, from its first lane (fewer
  // LDS atomics, and fewer lanes of one atomic on the same bin).
  const int lane = t & 63;
  const uint64_t above = ~((2ull << lane) - 1ull);  // lanes after this one
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const bool valid = base + j * kSortThreads + t < end;
    const uint32_t d = (uint32_t)(kk[j] >> shift) & mask;
    // the previous lane's digit by a DPP wave shift (lane 0's is not used)
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)d, 0x138, 0xF, 0xF, false);
    const bool head = valid && (lane == 0 || prev != d);
    const uint64_t heads = __ballot(head);
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));  // valid lanes: 0 .. nvalid - 1
    if (head) {
      const uint64_t later = heads & above;
      const uint32_t next = later ? (uint32_t)__ffsll((unsigned long long)later) - 1u : 64u;
      atomicAdd(&h[w][d], min(next, nvalid) - (uint32_t)lane);
    }
  }
  __syncthreads();
  for (int d = t; d < kDigitBins; d += kSortThreads)
    hist[hist_index(sg, (uint32_t)d, lb)] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// Exclusive prefix sums of col[0, n) in place, by one block of kSortThreads
// (s: kColChunk + kColChunk / 32 words of LDS, s_w: one per wave); returns
// the column's total.
constexpr int kColItems = 8;
constexpr int kColChunk = kSortThreads * kColItems;
__device__ inline uint32_t column_scan(uint32_t* __restrict__ col, uint32_t n_all, uint32_t* s,
                                       uint32_t* s_w) {
  const int t = threadIdx.x;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < n_all; b0 += kColChunk) {
    const uint32_t n = min((uint32_t)kColChunk, n_all - b0);
#pragma unroll
    for (int j = 0; j < kColItems; ++j) {
      const int li = j * kSortThreads + t;
      s[scan_pad(li)] = (uint32_t)li < n ? col[b0 + li] : 0u;
    }
    __syncthreads();
    uint32_t v[kColItems], sum = 0;
#pragma unroll
    for (int j = 0; j < kColItems; ++j) {
      v[j] = s[scan_pad(t * kColItems + j)];
      sum += v[j];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan<uint32_t, kSortThreads>(sum, s_w, tot) + carry;
#pragma unroll
    for (int j = 0; j < kColItems; ++j) {
      s[scan_pad(t * kColItems + j)] = ex;
      ex += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kColItems; ++j) {
      const int li = j * kSortThreads + t;
      if ((uint32_t)li < n) col[b0 + li] = s[scan_pad(li)];
    }
    carry += tot;
    __syncthreads();  // s is refilled next round
  }
  return carry;
}

// Exclusive prefix sums of each (segment, digit) column of the histogram
// table over the segment's blocks, in place, and the column totals
// (totals[segment * kDigitBins + digit]).  Grid: segments x kDigitBins.
__global__ __launch_bounds__(kSortThreads) void k_sort_colscan(const SortPlan* __restrict__ plan,
                                                              const SortSeg* __restrict__ segs,
                                                              int pass, uint32_t* __restrict__ hist,
                                                              uint32_t* __restrict__ totals) {
  __shared__ uint32_t s[kColChunk + kColChunk / 32];
  __shared__ uint32_t s_w[kSortThreads / 64];
  const uint32_t seg = blockIdx.x / kDigitBins, d = blockIdx.x % kDigitBins;
  const SortSeg sg = segs[seg];
  if (pass >= plan->passes || pass < plan->first) return;
  uint32_t* __restrict__ col = hist + (size_t)sg.block_base * kDigitBins + (size_t)d * sg.nblocks;
  const uint32_t tot = column_scan(col, sg.nblocks, s, s_w);
  if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

template <typename K, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    K* k0, K* k1, uint32_t* v0, uint32_t* v1, const SortPlan* __restrict__ plan,
    const SortSeg* __restrict__ segs, int nseg, int pass, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ col_totals) {
  constexpr int kTile = kSortThreads * ITEMS;
  __shared__ K s_key[kTile];
  __shared__ uint32_t s_val[kTile];
  // per-wave digit counts, then per-wave starts (both <= kTile; 16 bits
  // keep the block's LDS under 40 KB with 32-bit keys: four blocks per CU)
  __shared__ uint16_t s_cnt[4][kDigitBins];
  // output position of the block's key at LDS slot i of digit d: s_delta[d] + i
  // (the digit's global start minus its block-local start, modulo 2^32)
  __shared__ uint32_t s_delta[kDigitBins];
  __shared__ uint64_t s_w[4];
  SortSeg sg;
  uint32_t lb;
  int seg_index = 0;
  const bool mine = sort_block(segs, nseg, blockIdx.x, sg, lb, &seg_index);
  if (pass >= plan->passes || pass < plan->first || !sort_mode_is<K>(plan) ||
      plan->items != ITEMS)
    return;
  if (!mine) return;  // surplus block
  const uint32_t m = sg.key_base + sg.count;  // end of this block's segment
  const K* __restrict__ kin = (pass & 1) ? k1 : k0;
  const uint32_t* __restrict__ vin = (pass & 1) ? v1 : v0;
  K* __restrict__ kout = (pass & 1) ? k0 : k1;
  uint32_t* __restrict__ vout = (pass & 1) ? v0 : v1;
  const int shift = plan->shift[pass];
  const int width = plan->width[pass];
  const uint32_t mask = (1u << width) - 1u;
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const uint64_t lt = lanemask_lt();
  for (int d = t; d < kDigitBins; d += kSortThreads) {
#pragma unroll
    for (int i = 0; i < 4; ++i) s_cnt[i][d] = 0;
  }
  __syncthreads();
  // wave w owns keys [base + w kTile/4, base + (w + 1) kTile/4), 64 per round
  const uint32_t base = sg.key_base + lb * kTile;
  const uint32_t wbase = base + (uint32_t)w * (kTile / 4);
  K kk[ITEMS];
  uint32_t vv[ITEMS];
  uint32_t off[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t idx = wbase + (uint32_t)j * 64u + (uint32_t)lane;
    kk[j] = idx < m ? kin[idx] : K(0);
    vv[j] = idx < m ? vin[idx] : 0u;
  }
  // the block's column prefixes and the segment's column totals (digits 2t
  // and 2t + 1), needed only at the end: loaded after the keys so that their
  // latency hides behind the ranking
  constexpr int kPer = kDigitBins / kSortThreads;
  uint32_t glob[kPer], ctot[kPer];
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    glob[e] = offsets[hist_index(sg, (uint32_t)(kPer * t + e), lb)];
    ctot[e] = col_totals[(size_t)seg_index * kDigitBins + kPer * t + e];
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const bool valid = wbase + (uint32_t)j * 64u + (uint32_t)lane < m;
    const uint32_t d = (uint32_t)(kk[j] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < width; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const uint32_t before = valid ? s_cnt[w][d] : 0u;  // earlier rounds of this wave
    off[j] = before + rank;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid && rank == 0) s_cnt[w][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  {  // thread t: digits 2t and 2t + 1, their block-local starts and each wave's
     // start, and their starts in the segment (one scan of both: the local
     // sums stay below 2^32, so the high halves scan the column totals)
    uint32_t c[kPer][4], sum = 0, tsum = 0;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      tsum += ctot[e];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        c[e][i] = s_cnt[i][kPer * t + e];
        sum += c[e][i];
      }
    }
    uint64_t all;
    const uint64_t both =
        block_exclusive_scan<uint64_t, kSortThreads>(((uint64_t)tsum << 32) | sum, s_w, all);
    uint32_t st = (uint32_t)both;
    uint32_t dstart = sg.key_base + (uint32_t)(both >> 32);
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int d = kPer * t + e;
      s_delta[d] = dstart + glob[e] - st;
      dstart += ctot[e];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        s_cnt[i][d] = (uint16_t)st;
        st += c[e][i];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    if (wbase + (uint32_t)j * 64u + (uint32_t)lane < m) {
      const uint32_t d = (uint32_t)(kk[j] >> shift) & mask;
      const uint32_t pos = s_cnt[w][d] + off[j];
      s_key[pos] = kk[j];
      s_val[pos] = vv[j];
    }
  }
  __syncthreads();
  // the block's keys in digit order: each digit's run goes to consecutive
  // addresses of the output
  const uint32_t nblk = min((uint32_t)kTile, m - base);
  for (uint32_t i = t; i < nblk; i += kSortThreads) {
    const K key = s_key[i];
    const uint32_t d = (uint32_t)(key >> shift) & mask;
    const uint32_t dst = s_delta[d] + i;
    kout[dst] = key;
    vout[dst] = s_val[i];
  }
}

struct SortScratch {
  DevBuf* hist;
  DevBuf* hist_part;
};

// Sort blocks (the grid) for up to m_cap keys in nseg segments, the
// histogram table (one row per digit per block) and the column totals.
inline uint32_t sort_grid(uint64_t m_cap, int nseg, int items) {
  return div_up(m_cap, (uint64_t)kSortThreads * items) + (uint32_t)nseg;
}
inline uint64_t sort_hist_len(uint64_t m_cap, int nseg, int items) {
  return (uint64_t)sort_grid(m_cap, nseg, items) * kDigitBins;
}
// The column totals: one per segment and digit.
inline size_t sort_part_bytes(uint64_t m_cap, int nseg, int items) {
  (void)m_cap;
  (void)items;
  return (size_t)(nseg > 0 ? nseg : 1) * kDigitBins * sizeof(uint32_t);
}

// The passes of one tile size: pass slots [first, slots) over (k[0], v[0]).
template <int ITEMS>
inline LumenStatus sort_passes(uint64_t* k[2], uint32_t* v[2], const SortPlan* plan,
                               const SortSeg* segs, int nseg, uint64_t m_cap, int first,
                               int slots, int modes, uint32_t* hist, uint32_t* part,
                               hipStream_t st) {
  const uint32_t nb = sort_grid(m_cap, nseg, ITEMS);
  uint32_t* k32[2] = {reinterpret_cast<uint32_t*>(k[0]), reinterpret_cast<uint32_t*>(k[1])};
  for (int p = first; p < slots; ++p) {
    if (modes & 1) {
      hipLaunchKernelGGL((k_sort_hist<uint64_t, ITEMS>), dim3(nb), dim3(kSortThreads), 0, st, k[0],
                         k[1], plan, segs, nseg, p, hist);
      LUMEN_LAUNCHED();
    }
    if (modes & 2) {
      hipLaunchKernelGGL((k_sort_hist<uint32_t, ITEMS>), dim3(nb), dim3(kSortThreads), 0, st,
                         k32[0], k32[1], plan, segs, nseg, p, hist);
      LUMEN_LAUNCHED();
    }
    // the pass's column scans run only when the device plan has the pass,
    // over the columns of the frame's segments (a graph's grid covers its capacity)
    hipLaunchKernelGGL(k_sort_colscan, dim3((uint32_t)nseg * kDigitBins), dim3(kSortThreads), 0,
                       st, plan, segs, p, hist, part);
    LUMEN_LAUNCHED();
    if (modes & 1) {
      hipLaunchKernelGGL((k_sort_scatter<uint64_t, ITEMS>), dim3(nb), dim3(kSortThreads), 0, st,
                         k[0], k[1], v[0], v[1], plan, segs, nseg, p, hist, part);
      LUMEN_LAUNCHED();
    }
    if (modes & 2) {
      hipLaunchKernelGGL((k_sort_scatter<uint32_t, ITEMS>), dim3(nb), dim3(kSortThreads), 0, st,
                         k32[0], k32[1], v[0], v[1], plan, segs, nseg, p, hist, part);
      LUMEN_LAUNCHED();
    }
  }
  return LUMEN_SUCCESS;
}

// Enqueues pass slots [first, slots) over (k[0], v[0]) with the plan in
// device memory (a slot runs when plan->first <= slot < plan->passes; pass p
// reads buffer p & 1, so a frame whose pass 0 ran in k_emit starts in
// buffer 1); the result ends in buffer (plan->passes & 1), as 64-bit keys
// (mode 0) or 32-bit frame keys (mode 1) in the same buffers.  `modes` says
// which key widths to enqueue (bit 0: 64-bit, bit 1: 32-bit; a graph, which
// cannot know the frame's mode, takes both).  m_cap bounds the number of
// keys the plan may name; `items` (12 or 16) is the plan's keys per thread,
// which the segments' block counts were made with.
inline LumenStatus radix_sort_planned(uint64_t* k[2], uint32_t* v[2], const SortPlan* plan,
                                      const SortSeg* segs, int nseg, uint64_t m_cap, int first,
                                      int slots, int modes, int items, SortScratch sc,
                                      hipStream_t st) {
  if (m_cap == 0 || slots <= first) return LUMEN_SUCCESS;
  if (items != 12 && items != 16) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "sort: %d keys per thread", items);
  LUMEN_TRY(ensure(*sc.hist, sort_hist_len(m_cap, nseg, items) * sizeof(uint32_t)));
  LUMEN_TRY(ensure(*sc.hist_part, sort_part_bytes(m_cap, nseg, items)));
  uint32_t* hist = sc.hist->as<uint32_t>();
  uint32_t* part = sc.hist_part->as<uint32_t>();
  if (items == 12)
    return sort_passes<12>(k, v, plan, segs, nseg, m_cap, first, slots, modes, hist, part, st);
  return sort_passes<16>(k, v, plan, segs, nseg, m_cap, first, slots, modes, hist, part, st);
}

}  // namespace lumen
