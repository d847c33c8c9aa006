// loss.hip -- the photometric loss of the training step and its gradient in
// one pass over the images (lumen_ex_photometric_loss).
//
// Restates src/optim/loss.mojo: compute_l1_loss (:11-35, mean |pred -
// target|, subgradient +1/N where pred >= target and -1/N elsewhere),
// compute_ssim_window (:65-133, mean SSIM over non-overlapping 8x8 windows
// with the three channels pooled, partial windows at the right and bottom
// edges not counted, 1 when there is no full window) and
// compute_combined_loss (:136-144, (1 - lambda) L1 + lambda (1 - SSIM)).
//
//   k_loss_windows  one wave per 8x8 window, one lane per pixel: DPP sums of
//                   the window's values, centred squares and cross products,
//                   the window's SSIM and its analytic gradient, the L1
//                   subgradient, and per-block partial sums (|d|, SSIM);
//   k_loss_rest     pixels outside every full window: L1 only;
//   k_loss_final    one block adds the partial sums in a fixed order.
// With n = 192 values per window, A = 2 mx my + C1, B = 2 cxy + C2,
// C = mx^2 + my^2 + C1, D = vx + vy + C2 and SSIM = A B / (C D):
//   dSSIM/dx_i = (2 / n) (my B / (C D) + A dy_i / (C D) - A B mx / (C^2 D)
//                         - A B dx_i / (C D^2)),
// dx_i, dy_i the centred values (the sums of dx and dy over a window are 0).
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "context.hpp"
#include "lumen_ex.h"

namespace lumen {

constexpr float kSsimC1 = 0.0001f;  // loss.mojo:8
constexpr float kSsimC2 = 0.0009f;  // loss.mojo:9
constexpr int kLossWin = 8;

__device__ __forceinline__ float wave_total(float x) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_sum_lane63(x)), 63));
}

struct LossShape {
  uint32_t width, height, wx, hy;  // image size; full windows across and down
  float inv_n;                     // 1 / (W H 3)
  float w_l1, w_ssim;              // (1 - lambda), lambda / windows (0 without windows)
};

// partial[2 b] = sum |d| over block b's pixels, partial[2 b + 1] = sum of its windows' SSIM
__global__ __launch_bounds__(256) void k_loss_windows(const float* __restrict__ pred,
                                                      const float* __restrict__ target,
                                                      LossShape sh, float* __restrict__ grad,
                                                      float* __restrict__ partial) {
  __shared__ float s_part[4][2];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint32_t win = blockIdx.x * 4u + (uint32_t)w;
  const uint32_t nwin = sh.wx * sh.hy;
  float l1 = 0.0f, ssim = 0.0f;
  if (win < nwin) {  // uniform across the wave
    const uint32_t px = (win % sh.wx) * kLossWin + (uint32_t)(lane & 7);
    const uint32_t py = (win / sh.wx) * kLossWin + (uint32_t)(lane >> 3);
    const size_t base = ((size_t)py * sh.width + px) * 3u;
    float x[3], y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c] = pred[base + c];
      y[c] = target[base + c];
    }
    const float inv = 1.0f / 192.0f;
    const float mx = wave_total(x[0] + x[1] + x[2]) * inv;
    const float my = wave_total(y[0] + y[1] + y[2]) * inv;
    float dx[3], dy[3], sxx = 0.0f, syy = 0.0f, sxy = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dx[c] = x[c] - mx;
      dy[c] = y[c] - my;
      sxx += dx[c] * dx[c];
      syy += dy[c] * dy[c];
      sxy += dx[c] * dy[c];
    }
    const float vx = wave_total(sxx) * inv;
    const float vy = wave_total(syy) * inv;
    const float cxy = wave_total(sxy) * inv;
    const float A = 2.0f * mx * my + kSsimC1;
    const float B = 2.0f * cxy + kSsimC2;
    const float C = mx * mx + my * my + kSsimC1;
    const float D = vx + vy + kSsimC2;
    ssim = (A * B) / (C * D);
    const float k = 2.0f * inv;
    const float rCD = 1.0f / (C * D);
    const float t0 = k * (my * B * rCD - A * B * mx * rCD / C);  // the terms common to the window
    const float ty = k * A * rCD;
    const float tx = k * A * B * rCD / D;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float d = x[c] - y[c];
      l1 += d >= 0.0f ? d : -d;
      if (grad) {
        const float dssim = t0 + ty * dy[c] - tx * dx[c];
        grad[base + c] = sh.w_l1 * (d >= 0.0f ? sh.inv_n : -sh.inv_n) - sh.w_ssim * dssim;
      }
    }
    l1 = wave_total(l1);
  }
  if (lane == 0) {
    s_part[w][0] = l1;
    s_part[w][1] = ssim;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.0f, b = 0.0f;
    for (int i = 0; i < 4; ++i) {
      a += s_part[i][0];
      b += s_part[i][1];
    }
    partial[2u * blockIdx.x] = a;
    partial[2u * blockIdx.x + 1u] = b;
  }
}

// Pixels right of the last full window column (all rows) and below the last
// full window row (left part): the L1 term only.  partial[2 b] = sum |d|.
__global__ __launch_bounds__(256) void k_loss_rest(const float* __restrict__ pred,
                                                   const float* __restrict__ target, LossShape sh,
                                                   float* __restrict__ grad,
                                                   float* __restrict__ partial) {
  const uint32_t xr = sh.wx * kLossWin, yb = sh.hy * kLossWin;
  const uint64_t right = (uint64_t)(sh.width - xr) * sh.height;  // right strip pixels
  const uint64_t total = right + (uint64_t)xr * (sh.height - yb);
  float l1 = 0.0f;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * 256u) {
    uint32_t px, py;
    if (i < right) {
      const uint32_t rw = sh.width - xr;
      py = (uint32_t)(i / rw);
      px = xr + (uint32_t)(i % rw);
    } else {
      const uint64_t j = i - right;
      py = yb + (uint32_t)(j / xr);
      px = (uint32_t)(j % xr);
    }
    const size_t base = ((size_t)py * sh.width + px) * 3u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float d = pred[base + c] - target[base + c];
      l1 += d >= 0.0f ? d : -d;
      if (grad) grad[base + c] = sh.w_l1 * (d >= 0.0f ? sh.inv_n : -sh.inv_n);
    }
  }
  __shared__ float s_w[4];
  l1 = wave_total(l1);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = l1;
  __syncthreads();
  if (threadIdx.x == 0) partial[2u * blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// out[0] = L1, out[1] = SSIM, out[2] = (1 - lambda) L1 + lambda (1 - SSIM)
__global__ __launch_bounds__(1024) void k_loss_final(const float* __restrict__ pw, uint32_t nw_blocks,
                                                     const float* __restrict__ pr,
                                                     uint32_t nr_blocks, LossShape sh,
                                                     float lambda, float* __restrict__ out) {
  __shared__ float s_a[16], s_b[16];
  float a = 0.0f, b = 0.0f;
  for (uint32_t i = threadIdx.x; i < nw_blocks; i += 1024u) {
    a += pw[2u * i];
    b += pw[2u * i + 1u];
  }
  for (uint32_t i = threadIdx.x; i < nr_blocks; i += 1024u) a += pr[2u * i];
  a = wave_total(a);
  b = wave_total(b);
  if ((threadIdx.x & 63) == 0) {
    s_a[threadIdx.x >> 6] = a;
    s_b[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float sa = 0.0f, sb = 0.0f;
    for (int i = 0; i < 16; ++i) {
      sa += s_a[i];
      sb += s_b[i];
    }
    const uint32_t nwin = sh.wx * sh.hy;
    const float l1 = sa * sh.inv_n;
    const float ssim = nwin ? sb / (float)nwin : 1.0f;
    out[0] = l1;
    out[1] = ssim;
    out[2] = (1.0f - lambda) * l1 + lambda * (1.0f - ssim);
  }
}

// The partial sums of lumen_ex_photometric_loss_noalloc: one array per
// device for the life of the process, 2 floats per block of four 8x8
// windows plus at most 2048 blocks of the remaining pixels, enough for
// images up to 8192 x 8192 (2.02 MiB).  Host code never passes this name to
// a kernel: it is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress (scratch_ptr below).
constexpr size_t kLossScratchFloats = 2u * ((8192u / 8u) * (8192u / 8u) / 4u + 2048u);
static __device__ __align__(256) float scratch[kLossScratchFloats];

// Enqueues the loss of one image pair with its partial sums at `part`
// (loss_part_bytes(width, height) bytes).
static LumenStatus enqueue_loss(const float* pred, const float* target, uint32_t width,
                                uint32_t height, float lambda_dssim, float* out, float* grad,
                                float* part, hipStream_t st) {
  LossShape sh;
  sh.width = width;
  sh.height = height;
  sh.wx = width / kLossWin;
  sh.hy = height / kLossWin;
  sh.inv_n = 1.0f / (float)((uint64_t)width * height * 3u);
  const uint32_t nwin = sh.wx * sh.hy;
  sh.w_l1 = 1.0f - lambda_dssim;
  sh.w_ssim = nwin ? lambda_dssim / (float)nwin : 0.0f;
  const uint32_t bw = loss_window_blocks(width, height);
  const uint32_t br = loss_rest_blocks(width, height);
  float* pw = part;
  float* pr = pw + 2u * bw;
  hipLaunchKernelGGL(k_loss_windows, dim3(bw), dim3(256), 0, st, pred, target, sh, grad, pw);
  LUMEN_LAUNCHED();
  hipLaunchKernelGGL(k_loss_rest, dim3(br), dim3(256), 0, st, pred, target, sh, grad, pr);
  LUMEN_LAUNCHED();
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(1024), 0, st, pw, bw, pr, br, sh, lambda_dssim,
                     out);
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
constexpr int kMaxLossDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxLossDevices] = {};

static LumenStatus scratch_ptr(int device, float** p) {
  if (device < 0 || device >= kMaxLossDevices)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(LUMEN_EX_ERROR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                  hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<float*>(ptr);
  }
  *p = g_scratch_ptr[device];
  return LUMEN_SUCCESS;
}

}  // namespace lumen

using namespace lumen;

extern "C" LumenStatus lumen_ex_photometric_loss_noalloc(const float* pred, const float* target,
                                                         uint32_t width, uint32_t height,
                                                         float lambda_dssim, float* out,
                                                         float* grad, void* stream) {
  clear_error();
  if (!pred || !target || !out)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_photometric_loss_noalloc: null argument");
  if (width == 0 || height == 0)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_photometric_loss_noalloc: empty image");
  if (loss_part_bytes(width, height) > sizeof(scratch))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                "lumen_ex_photometric_loss_noalloc: %ux%u needs %zu bytes of partial sums, more "
                "than its fixed %zu (images up to 8192 x 8192)",
                width, height, loss_part_bytes(width, height), sizeof(scratch));
  int device = 0;
  LUMEN_HIP(hipGetDevice(&device));
  float* part = nullptr;
  LUMEN_TRY(scratch_ptr(device, &part));
  return enqueue_loss(pred, target, width, height, lambda_dssim, out, grad, part,
                      (hipStream_t)stream);
}

extern "C" LumenStatus lumen_ex_photometric_loss(LumenExContext* c, const float* pred,
                                                 const float* target, uint32_t width,
                                                 uint32_t height, float lambda_dssim, float* out,
                                                 float* grad, void* stream) {
  clear_error();
  if (!c || !pred || !target || !out)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_photometric_loss: null argument");
  if (width == 0 || height == 0)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_photometric_loss: empty image");
  hipStream_t st = (hipStream_t)stream;
  LUMEN_HIP(hipSetDevice(c->device));
  LUMEN_TRY(ensure(c->loss_part, loss_part_bytes(width, height)));
  return enqueue_loss(pred, target, width, height, lambda_dssim, out, grad,
                      c->loss_part.as<float>(), st);
}
