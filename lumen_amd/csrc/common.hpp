// common.hpp -- error reporting, device buffers and wave64 primitives.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "lumen_ex.h"

namespace lumen {

// Last error of the calling thread (lumen_ex_last_error).
void set_error(const char* fmt, ...);
const char* last_error();
void clear_error();

inline LumenStatus fail(LumenStatus s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_error("%s", buf);
  return s;
}

#define LUMEN_HIP(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::lumen::fail(LUMEN_EX_ERROR_DEVICE, "%s failed: %s (%s:%d)", #expr,          \
                           hipGetErrorString(e_), __FILE__, __LINE__);                     \
  } while (0)

#define LUMEN_TRY(expr)                      \
  do {                                       \
    LumenStatus s_ = (expr);                 \
    if (s_ != LUMEN_SUCCESS) return s_;      \
  } while (0)

#define LUMEN_LAUNCHED() LUMEN_HIP(hipGetLastError())

// A device allocation that only grows.  hipFree on growth synchronizes the
// device, which is acceptable because growth stops after the first frames.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// What ensure() allocates for a request of `bytes` (a quarter of headroom).
inline size_t alloc_size(size_t bytes) { return bytes + bytes / 4 + 256; }

inline LumenStatus ensure(DevBuf& b, size_t bytes) {
  if (bytes <= b.cap && b.ptr) return LUMEN_SUCCESS;
  if (b.ptr) {
    LUMEN_HIP(hipFree(b.ptr));
    b.ptr = nullptr;
    b.cap = 0;
  }
  size_t want = alloc_size(bytes);
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return fail(LUMEN_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
  }
  b.cap = want;
  return LUMEN_SUCCESS;
}

inline void release(DevBuf& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.cap = 0;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- wave64 / block primitives ------------------------------------------

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = threadIdx.x & 63u;
  return (1ull << lane) - 1ull;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Wave64 reductions on DPP (cross-lane moves inside the VALU, no LDS round
// trip): row_shr 1, 2, 4, 8 leave each 16-lane row's total in its lane 15,
// row_bcast:15 and row_bcast:31 fold the four rows into lane 63.  Lanes whose
// DPP source is out of range (or whose row is masked off) take `identity`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x, uint32_t identity) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)x, CTRL, ROW_MASK, 0xF, false);
}

// row_shr by CTRL - 0x110 lanes within each 16-lane row, out-of-range lanes
// reading 0 (bound_ctrl: no old value to keep, so none need be set up).
template <int CTRL>
__device__ __forceinline__ float dpp_shr_zero(float x) {
  static_assert(CTRL > 0x110 && CTRL < 0x120, "row_shr");
  return __uint_as_float(
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(x), CTRL, 0xF, 0xF, true));
}

// Sum over the 64 lanes; the total is valid in lane 63.
__device__ __forceinline__ float wave_sum_lane63(float x) {
  x += dpp_shr_zero<0x111>(x);
  x += dpp_shr_zero<0x112>(x);
  x += dpp_shr_zero<0x114>(x);
  x += dpp_shr_zero<0x118>(x);
  x += __uint_as_float(dpp_u32<0x142, 0xA>(__float_as_uint(x), 0u));
  x += __uint_as_float(dpp_u32<0x143, 0xC>(__float_as_uint(x), 0u));
  return x;
}

// Sum over each 16-lane row; the row's total is valid in its lane 15.
__device__ __forceinline__ float row_sum_lane15(float x) {
  x += dpp_shr_zero<0x111>(x);
  x += dpp_shr_zero<0x112>(x);
  x += dpp_shr_zero<0x114>(x);
  x += dpp_shr_zero<0x118>(x);
  return x;
}

// gfx950 v_permlane32_swap: lanes 0-31 get lo[l] + lo[l + 32], lanes 32-63
// get hi[l - 32] + hi[l] (the upper half of lo trades places with the lower
// half of hi, then the two registers are added).
__device__ __forceinline__ float swap32_add(float lo, float hi) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(lo), __float_as_uint(hi), false,
                                                  false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// gfx950 v_permlane16_swap: with rows of 16 lanes, row 0 gets x.row0 +
// x.row1, row 1 y.row0 + y.row1, row 2 x.row2 + x.row3, row 3 y.row2 + y.row3.
__device__ __forceinline__ float swap16_add(float x, float y) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false,
                                                  false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Sums of ten per-lane values over the whole wave by transposition: two
// lane-swap rounds fold the values pairwise (each add now serves two values),
// then DPP finishes within 16-lane rows.  Every lane must be active.  On
// return, lane 16 r + 15 holds the sum of value sum10_slot(r) in a, of value
// 4 + sum10_slot(r) in b, and (r = 0, 2) of value 8 + r / 2 in c.
__device__ __forceinline__ int sum10_slot(int r) { return (r & 1) * 2 + (r >> 1); }  // 0 2 1 3
__device__ __forceinline__ void wave_sum10(const float v[10], float& a, float& b, float& c) {
  a = row_sum_lane15(swap16_add(swap32_add(v[0], v[1]), swap32_add(v[2], v[3])));
  b = row_sum_lane15(swap16_add(swap32_add(v[4], v[5]), swap32_add(v[6], v[7])));
  const float r = swap32_add(v[8], v[9]);
  c = row_sum_lane15(swap16_add(r, r));
}

// The first eight values of wave_sum10, in the same lanes of a and b.
__device__ __forceinline__ void wave_sum8(const float v[8], float& a, float& b) {
  a = row_sum_lane15(swap16_add(swap32_add(v[0], v[1]), swap32_add(v[2], v[3])));
  b = row_sum_lane15(swap16_add(swap32_add(v[4], v[5]), swap32_add(v[6], v[7])));
}

__device__ __forceinline__ uint32_t wave_min_u32_all(uint32_t x) {
  x = min(x, dpp_u32<0x111, 0xF>(x, 0xFFFFFFFFu));
  x = min(x, dpp_u32<0x112, 0xF>(x, 0xFFFFFFFFu));
  x = min(x, dpp_u32<0x114, 0xF>(x, 0xFFFFFFFFu));
  x = min(x, dpp_u32<0x118, 0xF>(x, 0xFFFFFFFFu));
  x = min(x, dpp_u32<0x142, 0xA>(x, 0xFFFFFFFFu));
  x = min(x, dpp_u32<0x143, 0xC>(x, 0xFFFFFFFFu));
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

__device__ __forceinline__ uint32_t wave_max_u32_all(uint32_t x) {
  x = max(x, dpp_u32<0x111, 0xF>(x, 0u));
  x = max(x, dpp_u32<0x112, 0xF>(x, 0u));
  x = max(x, dpp_u32<0x114, 0xF>(x, 0u));
  x = max(x, dpp_u32<0x118, 0xF>(x, 0u));
  x = max(x, dpp_u32<0x142, 0xA>(x, 0u));
  x = max(x, dpp_u32<0x143, 0xC>(x, 0u));
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Exclusive scan across the block; every thread must call it.  s_wave must
// hold THREADS/64 elements.  Ends with a barrier so s_wave may be reused.
template <typename T, int THREADS>
__device__ __forceinline__ T block_exclusive_scan(T x, T* s_wave, T& total) {
  constexpr int NW = THREADS / 64;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const T inc = wave_inclusive_scan(x);
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  if (w == 0) {
    T s = lane < NW ? s_wave[lane] : T(0);
    s = wave_inclusive_scan(s);
    if (lane < NW) s_wave[lane] = s;
  }
  __syncthreads();
  const T prefix = w > 0 ? s_wave[w - 1] : T(0);
  total = s_wave[NW - 1];
  __syncthreads();
  return prefix + inc - x;
}

// The blocks of a launch are dealt to the 8 XCDs in turn (block i runs on
// XCD i % 8), each XCD with its own L2.  xcd_block renumbers blocks 0..n-1 so
// that each XCD takes a contiguous run of them: neighbouring blocks' writes
// to shared cache lines (histogram rows of neighbouring sort blocks, the
// ends of their digit runs, the gradient slots of a splat that covers
// neighbouring tiles) then go through one L2, where the pieces merge before
// they reach HBM.  A bijection of [0, n).
__device__ __forceinline__ uint32_t xcd_block(uint32_t i, uint32_t n) {
  constexpr uint32_t kXcds = 8;
  const uint32_t q = n / kXcds, r = n % kXcds, x = i % kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i / kXcds;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
inline LumenStatus fill(T* p, T v, uint64_t n, hipStream_t st) {
  if (n == 0) return LUMEN_SUCCESS;
  hipLaunchKernelGGL(k_fill<T>, dim3(div_up(n, 256)), dim3(256), 0, st, p, v, n);
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

}  // namespace lumen
