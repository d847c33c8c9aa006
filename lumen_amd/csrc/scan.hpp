// scan.hpp -- device-wide exclusive prefix sum (reduce / spine / downsweep;
// for up to kScanFuse chunks the downsweep adds up its predecessors itself).
//
// Used for the visible-splat compaction (packed visible flag + tile count per
// Gaussian, 64-bit) and for the radix sort's digit-major block histograms
// (32-bit).  Loads and stores are coalesced; the per-thread serial part runs
// out of LDS.
#pragma once

#include "common.hpp"

namespace lumen {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ int scan_pad(int i) { return i + (i >> 5); }

// A scan that runs only when *value > above (value null: always), so that a
// graph can hold scans of sort passes the frame's device plan may not use,
// over at most *len * len_mul elements (len null: the launch's n), so that a
// graph sized for a capacity scans only what the frame filled.
struct ScanGate {
  const int32_t* value = nullptr;
  int32_t above = 0;
  const uint32_t* len = nullptr;
  uint32_t len_mul = 1;
};

__device__ __forceinline__ bool gate_closed(ScanGate g) { return g.value && !(*g.value > g.above); }

__device__ __forceinline__ uint64_t gate_len(ScanGate g, uint64_t n) {
  if (!g.len) return n;
  const uint64_t m = (uint64_t)*g.len * g.len_mul;
  return m < n ? m : n;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in, uint64_t n,
                                                              T* __restrict__ partials, ScanGate gate) {
  if (gate_closed(gate)) return;
  __shared__ T s_w[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
  n = gate_len(gate, n);
  if (base >= n) {  // past the filled length: an empty chunk
    if (threadIdx.x == 0) partials[blockIdx.x] = T(0);
    return;
  }
  T s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const uint64_t idx = base + (uint64_t)j * kScanThreads + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) t += s_w[w];
    partials[blockIdx.x] = t;
  }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_scan_spine(T* __restrict__ partials, uint32_t nb,
                                                     T* __restrict__ total_out, ScanGate gate) {
  if (gate_closed(gate)) return;
  __shared__ T s_w[1024 / 64];
  T carry = 0;
  for (uint32_t base = 0; base < nb; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const T x = i < nb ? partials[i] : T(0);
    T tot;
    const T ex = block_exclusive_scan<T, 1024>(x, s_w, tot);
    if (i < nb) partials[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const T* __restrict__ in, uint64_t n,
                                                            const T* __restrict__ partials,
                                                            T* __restrict__ out, ScanGate gate) {
  if (gate_closed(gate)) return;
  __shared__ T s[kScanChunk + kScanChunk / 32];
  __shared__ T s_w[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
  n = gate_len(gate, n);
  if (base >= n) return;  // past the filled length: nothing to write
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int li = j * kScanThreads + t;
    const uint64_t idx = base + li;
    s[scan_pad(li)] = idx < n ? in[idx] : T(0);
  }
  __syncthreads();
  T v[kScanItems];
  T sum = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = s[scan_pad(t * kScanItems + j)];
    sum += v[j];
  }
  T tot;
  T ex = block_exclusive_scan<T, kScanThreads>(sum, s_w, tot) + partials[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    s[scan_pad(t * kScanItems + j)] = ex;
    ex += v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int li = j * kScanThreads + t;
    const uint64_t idx = base + li;
    if (idx < n) out[idx] = s[scan_pad(li)];
  }
}

// k_scan_down for at most kScanFuse chunks: each block adds up the partial
// sums of the chunks before it itself (no spine launch); the last block
// writes the total.  per_chunk partials cover one chunk (1 from
// k_scan_reduce; more when the producer of the data summed smaller pieces
// of it, in order, as it wrote them).
constexpr uint32_t kScanFuse = 1024;

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down_fused(const T* __restrict__ in,
                                                                  uint64_t n,
                                                                  const T* __restrict__ partials,
                                                                  T* __restrict__ out,
                                                                  T* __restrict__ total_out, ScanGate gate,
                                                                  uint32_t per_chunk) {
  if (gate_closed(gate)) return;
  __shared__ T s[kScanChunk + kScanChunk / 32];
  __shared__ T s_w[kScanThreads / 64];
  __shared__ T s_prefix;
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
  n = gate_len(gate, n);
  // past the filled length nothing is written, except the total by the last block
  if (base >= n && !(total_out && blockIdx.x == gridDim.x - 1)) return;
  const int t = threadIdx.x;
  // the chunk's elements, loaded together with the partials before it
  T x[kScanItems];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const uint64_t idx = base + (uint64_t)(j * kScanThreads + t);
    x[j] = idx < n ? in[idx] : T(0);
  }
  {
    const uint32_t lim = blockIdx.x * per_chunk;
    T p = 0;
    for (uint32_t b0 = 0; b0 < lim; b0 += 4u * kScanThreads) {  // four loads in flight
      T q[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t b = b0 + (uint32_t)(j * kScanThreads + t);
        q[j] = b < lim ? partials[b] : T(0);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) p += q[j];
    }
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s[scan_pad(j * kScanThreads + t)] = x[j];
    p = wave_sum(p);
    if ((t & 63) == 0) s_w[t >> 6] = p;
    __syncthreads();  // (also orders the chunk's LDS writes before the reads below)
    if (t == 0) {
      T q = 0;
#pragma unroll
      for (int w = 0; w < kScanThreads / 64; ++w) q += s_w[w];
      s_prefix = q;
    }
    __syncthreads();
  }
  T v[kScanItems];
  T sum = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = s[scan_pad(t * kScanItems + j)];
    sum += v[j];
  }
  T tot;
  const T prefix = s_prefix;
  T ex = block_exclusive_scan<T, kScanThreads>(sum, s_w, tot) + prefix;
  if (t == 0 && total_out && blockIdx.x == gridDim.x - 1) *total_out = prefix + tot;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    s[scan_pad(t * kScanItems + j)] = ex;
    ex += v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int li = j * kScanThreads + t;
    const uint64_t idx = base + li;
    if (idx < n) out[idx] = s[scan_pad(li)];
  }
}

// Exclusive scan of n elements (in may equal out).  partials must hold
// div_up(n, kScanChunk) elements; total_out (device, may be null) receives
// the sum of all elements.
template <typename T>
inline LumenStatus exclusive_scan(const T* in, T* out, uint64_t n, T* partials, T* total_out,
                                  hipStream_t st, ScanGate gate = ScanGate{}) {
  const uint32_t nb = div_up(n ? n : 1, kScanChunk);
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(kScanThreads), 0, st, in, n, partials,
                     gate);
  LUMEN_LAUNCHED();
  if (nb <= kScanFuse) {
    hipLaunchKernelGGL(k_scan_down_fused<T>, dim3(nb), dim3(kScanThreads), 0, st, in, n, partials,
                       out, total_out, gate, 1u);
    LUMEN_LAUNCHED();
    return LUMEN_SUCCESS;
  }
  hipLaunchKernelGGL(k_scan_spine<T>, dim3(1), dim3(1024), 0, st, partials, nb, total_out, gate);
  LUMEN_LAUNCHED();
  hipLaunchKernelGGL(k_scan_down<T>, dim3(nb), dim3(kScanThreads), 0, st, in, n, partials, out,
                     gate);
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

// Exclusive scan of n elements (in may equal out) whose producer already
// wrote the sums of its consecutive pieces of `piece` elements to partials
// (piece divides kScanChunk; at most kScanFuse chunks): the downsweep alone.
template <typename T>
inline LumenStatus exclusive_scan_pieces(const T* in, T* out, uint64_t n, const T* partials,
                                         uint32_t piece, T* total_out, hipStream_t st) {
  const uint32_t nb = div_up(n ? n : 1, kScanChunk);
  if (nb > kScanFuse || piece == 0 || kScanChunk % piece != 0)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "scan: %u chunks, pieces of %u", nb, piece);
  hipLaunchKernelGGL(k_scan_down_fused<T>, dim3(nb), dim3(kScanThreads), 0, st, in, n, partials,
                     out, total_out, ScanGate{}, (uint32_t)kScanChunk / piece);
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

}  // namespace lumen
