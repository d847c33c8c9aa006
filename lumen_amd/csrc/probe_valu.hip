// probe_valu.hip -- the VALU issue rate of this GPU with the library's build
// flags (no packed FP32, no contraction), the ceiling the compositors' VALU
// instruction counts are read against.
//
// Every thread runs kIters rounds of 8 independent chains (no dependency
// stalls with 8 waves per SIMD), either v_mul_f32 + v_add_f32 or v_fma_f32,
// over a grid of 8 four-wave blocks per CU.  Prints one JSON line with the
// wave-level instructions per second and per SIMD-cycle at the reported peak
// clock.
#include <hip/hip_runtime.h>
#include <stdio.h>

constexpr int kIters = 4096;

__global__ __launch_bounds__(256) void k_mul_add(float* out, float s) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = threadIdx.x * 0.001f + k;
  for (int i = 0; i < kIters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = a[k] * s + 0.5f;  // v_mul_f32, v_add_f32
  }
  float r = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) r += a[k];
  out[blockIdx.x * 256 + threadIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_fma(float* out, float s) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = threadIdx.x * 0.001f + k;
  for (int i = 0; i < kIters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = fmaf(a[k], s, 0.5f);  // v_fma_f32
  }
  float r = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) r += a[k];
  out[blockIdx.x * 256 + threadIdx.x] = r;
}

#define CHECK(x)                                                   \
  do {                                                             \
    hipError_t e_ = (x);                                           \
    if (e_ != hipSuccess) {                                        \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));      \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  int cus = 0, clock_khz = 0;
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  CHECK(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, 0));
  const int blocks = cus * 8;
  float* out = nullptr;
  CHECK(hipMalloc(&out, sizeof(float) * blocks * 256));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  double rate[2];
  for (int v = 0; v < 2; ++v) {
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {  // the first launch warms up
      CHECK(hipEventRecord(e0, 0));
      if (v == 0) hipLaunchKernelGGL(k_mul_add, dim3(blocks), dim3(256), 0, 0, out, 0.999f);
      else hipLaunchKernelGGL(k_fma, dim3(blocks), dim3(256), 0, 0, out, 0.999f);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1, 0));
      CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms < best) best = ms;
    }
    const double waves = blocks * 4.0;
    const double instr = waves * kIters * 8.0 * (v == 0 ? 2.0 : 1.0);
    rate[v] = instr / (best * 1e-3);
  }
  const double simd_cycles = cus * 4.0 * clock_khz * 1e3;
  printf("{\"cus\": %d, \"clock_khz\": %d, \"mul_add_wave_instr_per_s\": %.4g, "
         "\"fma_wave_instr_per_s\": %.4g, \"mul_add_per_simd_cycle\": %.4f, "
         "\"fma_per_simd_cycle\": %.4f}\n",
         cus, clock_khz, rate[0], rate[1], rate[0] / simd_cycles, rate[1] / simd_cycles);
  CHECK(hipFree(out));
  return 0;
}
