// legacy.hip -- the reference C library's own renderer (c_api/src/lumen_c.c:
// 380-575) on MI355X, bit for bit: the isotropic, DC-only splat renderer
// behind lumen_rasterize, selected with lumen_ex_rasterizer_set_mode.
//
// The reference projects every Gaussian, sorts the visible ones by depth
// (qsort, which glibc implements as a stable merge sort), then for every
// tile walks all visible splats in depth order and, for those whose box
// meets the tile, updates each pixel of the tile in its 8-bit framebuffer.
// Here:
//   k_legacy_project  one thread per Gaussian: view transform, culling,
//                     radius, variance, DC colour (the C expression trees)
//   scan + compact    visible splats in index order, depth keys
//   radix sort        stable, by depth (lumen_ex_radix_sort): equal depths
//                     keep index order, as the merge sort does
//   k_legacy_tiles    per splat in depth order, the tiles its box meets: the
//                     reference's four float comparisons are monotone in the
//                     tile index, so binary searches find the exact run
//   emit + sort       (tile, depth rank) pairs sorted by tile: each tile's
//                     splats in depth order, exactly the ones the reference's
//                     tile test keeps
//   k_legacy_composite one workgroup per tile, one thread per pixel (looping
//                     for tiles over 256 pixels), splats staged through LDS;
//                     each pixel runs the reference's per-pixel update on its
//                     framebuffer values in registers.
// exp is legacy_expf (the C library's expf restated, legacy_expf.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>

#define LEGACY_HD __host__ __device__ __forceinline__ static
#include "legacy_expf.h"
#include "legacy.hpp"
#include "scan.hpp"

namespace lumen {

struct LegacyView {
  float m[12];             // view matrix rows (right, up, -forward) with translations
  float focal;             // H / (2 tan(fov_y / 2)), both axes
  float half_w, half_h;    // (float)W * 0.5f, (float)H * 0.5f
  float near_plane, far_plane, eta, sigma;
  uint32_t width, height, ts, tiles_x, tiles_y;
};

// Splat record: (sx, sy, depth, radius), (inv_var, r, g, b), opacity.
struct LegacyRec {
  float4 a, b;
  float op;
  float pad[3];
};

// (uint8_t)(float) as x86-64 code does it: a truncating conversion to a
// 32-bit integer (NaN and out-of-range values give INT32_MIN), then its low
// byte.  Values in [0, 256) convert as usual.
__device__ __forceinline__ uint8_t legacy_u8(float v) {
  int32_t i;
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) i = INT32_MIN;
  else i = (int32_t)v;
  return (uint8_t)(uint32_t)i;
}

// An ascending integer key of a float depth (+0 and -0 compare equal, as
// the reference's comparator has it).
__device__ __forceinline__ uint64_t legacy_depth_key(float d) {
  const uint32_t u = __float_as_uint(d + 0.0f);
  return (uint64_t)((u & 0x80000000u) ? ~u : (u | 0x80000000u));
}

// lumen_c.c:407-468 for Gaussian i.
__global__ __launch_bounds__(256) void k_legacy_project(const float* __restrict__ aos, uint32_t n,
                                                        LegacyView v,
                                                        LegacyRec* __restrict__ proj,
                                                        uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float* g = aos + (size_t)i * 59u;
  const float x = g[0], y = g[1], z = g[2];
  const float vx = v.m[0] * x + v.m[1] * y + v.m[2] * z + v.m[3];
  const float vy = v.m[4] * x + v.m[5] * y + v.m[6] * z + v.m[7];
  const float vz = -(v.m[8] * x + v.m[9] * y + v.m[10] * z + v.m[11]);
  uint32_t vis = 0;
  if (!(vz <= v.near_plane || vz >= v.far_plane)) {
    const float sx = (vx / vz) * v.focal + v.half_w;
    const float sy = (vy / vz) * v.focal + v.half_h;
    float smax = g[3];
    if (g[4] > smax) smax = g[4];
    if (g[5] > smax) smax = g[5];
    float radius = (smax * v.focal / vz) * v.sigma;
    if (radius < 1.0f) radius = 1.0f;
    if (!(sx + radius < 0.0f || sx - radius >= (float)v.width || sy + radius < 0.0f ||
          sy - radius >= (float)v.height)) {
      float var = (radius / v.sigma) * (radius / v.sigma);
      if (var < 0.5f) var = 0.5f;
      const float c0 = 0.2820947917f;
      float c[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        c[k] = 0.5f + c0 * g[11 + k];
        if (c[k] < 0.0f) c[k] = 0.0f;
        if (c[k] > 1.0f) c[k] = 1.0f;
      }
      LegacyRec r;
      r.a = make_float4(sx, sy, vz, radius);
      r.b = make_float4(1.0f / var, c[0], c[1], c[2]);
      r.op = g[10];
      r.pad[0] = r.pad[1] = r.pad[2] = 0.0f;
      proj[i] = r;
      vis = 1;
    }
  }
  flag[i] = vis;
}

// Visible splats in index order, with their depth keys and indices.
__global__ __launch_bounds__(256) void k_legacy_compact(const LegacyRec* __restrict__ proj,
                                                        const uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ off, uint32_t n,
                                                        LegacyRec* __restrict__ rec,
                                                        uint64_t* __restrict__ dkeys,
                                                        uint32_t* __restrict__ dvals) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint32_t j = off[i];
  const LegacyRec r = proj[i];
  rec[j] = r;
  dkeys[j] = legacy_depth_key(r.a.z);
  dvals[j] = j;
}

// The run of tiles along one axis whose [start, end) pixel span the box
// [lo, hi] meets by the reference's test: !(hi < start) && !(lo >= end),
// start = t ts, end = min(start + ts, size).  Both conditions are monotone
// in t (the first holds up to some t, the second from some t; NaN keeps
// both), so two binary searches give the run.  Returns false when empty.
__device__ __forceinline__ bool legacy_axis(float lo, float hi, uint32_t ts, uint32_t size,
                                            uint32_t ntiles, uint32_t& first, uint32_t& last) {
  auto meets_start = [&](uint32_t t) { return !(hi < (float)(t * ts)); };
  auto meets_end = [&](uint32_t t) {
    const uint32_t e = t * ts + ts < size ? t * ts + ts : size;
    return !(lo >= (float)e);
  };
  if (!meets_start(0) || !meets_end(ntiles - 1)) return false;
  uint32_t a = 0, b = ntiles - 1;  // last t with meets_start
  while (a < b) {
    const uint32_t mid = (a + b + 1) >> 1;
    if (meets_start(mid)) a = mid;
    else b = mid - 1;
  }
  last = a;
  a = 0;
  b = ntiles - 1;  // first t with meets_end
  while (a < b) {
    const uint32_t mid = (a + b) >> 1;
    if (meets_end(mid)) b = mid;
    else a = mid + 1;
  }
  first = a;
  return first <= last;
}

// Per splat in depth order (rank j): its record, its tile rectangle and the
// number of tiles in it.
__global__ __launch_bounds__(256) void k_legacy_tiles(const LegacyRec* __restrict__ rec,
                                                      const uint32_t* __restrict__ order, uint32_t nv,
                                                      LegacyView v, LegacyRec* __restrict__ sorted,
                                                      uint4* __restrict__ rect,
                                                      uint64_t* __restrict__ cnt) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= nv) return;
  const LegacyRec r = rec[order[j]];
  sorted[j] = r;
  const float sx = r.a.x, sy = r.a.y, rad = r.a.w;
  uint32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0;
  const bool hit = legacy_axis(sx - rad, sx + rad, v.ts, v.width, v.tiles_x, x0, x1) &&
                   legacy_axis(sy - rad, sy + rad, v.ts, v.height, v.tiles_y, y0, y1);
  rect[j] = make_uint4(x0, y0, x1, y1);
  cnt[j] = hit ? (uint64_t)(x1 - x0 + 1) * (uint64_t)(y1 - y0 + 1) : 0ull;
}

// One thread per (tile, rank) pair: the pair's rank by a search of the
// scanned counts, its tile from the rectangle, ty-major.
__global__ __launch_bounds__(256) void k_legacy_emit(const uint64_t* __restrict__ koff, uint32_t nv,
                                                     uint64_t m, const uint4* __restrict__ rect,
                                                     uint32_t tiles_x, uint64_t* __restrict__ tkeys,
                                                     uint32_t* __restrict__ tvals) {
  const uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (k >= m) return;
  uint32_t a = 0, b = nv - 1;  // last rank with koff[rank] <= k
  while (a < b) {
    const uint32_t mid = (a + b + 1) >> 1;
    if (koff[mid] <= k) a = mid;
    else b = mid - 1;
  }
  const uint4 r = rect[a];
  const uint64_t local = k - koff[a];
  const uint32_t w = r.z - r.x + 1u;
  const uint32_t ty = r.y + (uint32_t)(local / w);
  const uint32_t tx = r.x + (uint32_t)(local % w);
  const uint64_t tile = (uint64_t)ty * tiles_x + tx;
  tkeys[k] = (tile << 32) | (uint64_t)a;
  tvals[k] = a;
}

// lumen_c.c:474-550 for the pixels of one tile.
__global__ __launch_bounds__(256) void k_legacy_composite(const int2* __restrict__ ranges,
                                                          const uint32_t* __restrict__ tvals,
                                                          const LegacyRec* __restrict__ sorted,
                                                          LegacyView v, uint8_t* __restrict__ rgb,
                                                          float* __restrict__ depth,
                                                          float* __restrict__ alpha,
                                                          uint32_t* __restrict__ touched) {
  __shared__ float4 s_a[256], s_b[256];
  __shared__ float s_op[256];
  const uint32_t tile = blockIdx.x;
  const int2 rg = ranges[tile];
  if (rg.x < 0 || rg.x >= rg.y) return;  // no splat meets the tile
  if (threadIdx.x == 0) atomicAdd(touched, 1u);
  const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
  const uint32_t x0 = tx * v.ts, y0 = ty * v.ts;
  const uint32_t x1 = x0 + v.ts < v.width ? x0 + v.ts : v.width;
  const uint32_t y1 = y0 + v.ts < v.height ? y0 + v.ts : v.height;
  const uint32_t w = x1 - x0;
  const uint64_t npix = (uint64_t)w * (y1 - y0);
  const float inv255 = 1.0f / 255.0f;
  for (uint64_t pb = 0; pb < npix; pb += 256u) {
    const uint64_t p = pb + threadIdx.x;
    const bool active = p < npix;
    uint32_t px = 0, py = 0;
    size_t pi = 0;
    uint8_t c8[3] = {0, 0, 0};
    float a = 0.0f, d = 0.0f;
    if (active) {
      px = x0 + (uint32_t)(p % w);
      py = y0 + (uint32_t)(p / w);
      pi = (size_t)py * v.width + px;
      c8[0] = rgb[3 * pi + 0];
      c8[1] = rgb[3 * pi + 1];
      c8[2] = rgb[3 * pi + 2];
      a = alpha[pi];
      d = depth[pi];
    }
    const float fx = (float)px, fy = (float)py;
    for (int base = rg.x; base < rg.y; base += 256) {
      const int cnt = min(256, rg.y - base);
      __syncthreads();
      if ((int)threadIdx.x < cnt) {
        const LegacyRec r = sorted[tvals[base + threadIdx.x]];
        s_a[threadIdx.x] = r.a;
        s_b[threadIdx.x] = r.b;
        s_op[threadIdx.x] = r.op;
      }
      __syncthreads();
      if (!active) continue;
      for (int j = 0; j < cnt; ++j) {
        if (a >= v.eta) continue;  // early ray termination
        const float4 sa = s_a[j];  // sx, sy, depth, radius
        const float dy = fy - sa.y;
        const float dx = fx - sa.x;
        if (dx * dx + dy * dy > sa.w * sa.w) continue;
        const float4 sb = s_b[j];  // inv_var, r, g, b
        float power = -0.5f * (sb.x * dx * dx + sb.x * dy * dy);
        if (power > 0.0f) power = 0.0f;
        const float al = s_op[j] * legacy_expf(power);
        if (al < inv255) continue;
        const float wgt = al * (1.0f - a);
        const float col[3] = {sb.y, sb.z, sb.w};
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          float val = (float)c8[ch] / 255.0f + wgt * col[ch];
          if (val > 1.0f) val = 1.0f;
          c8[ch] = legacy_u8(val * 255.0f);
        }
        if (a < 0.5f && sa.z < d) d = sa.z;
        a = a + wgt;
      }
    }
    if (active) {
      rgb[3 * pi + 0] = c8[0];
      rgb[3 * pi + 1] = c8[1];
      rgb[3 * pi + 2] = c8[2];
      alpha[pi] = a;
      depth[pi] = d;
    }
  }
}

void legacy_release(LegacyBuffers& b) {
  for (DevBuf* d : {&b.proj, &b.flag, &b.off, &b.part, &b.totals, &b.rec, &b.rec_sorted, &b.dkeys,
                    &b.dvals, &b.rect, &b.cnt, &b.koff, &b.kpart, &b.tkeys, &b.tvals, &b.ranges,
                    &b.rgb, &b.depth, &b.alpha, &b.touched})
    release(*d);
}

// The view of lumen_c.c:391-401 (look_at :354-369, vector helpers :330-352),
// on the host in the C expression trees.
static void legacy_view(const LumenConfig& cfg, const LumenCamera& cam, uint32_t W, uint32_t H,
                        LegacyView& v) {
  auto dot = [](const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  auto cross = [](const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto normalize = [&](float x[3]) {
    const float len = sqrtf(dot(x, x));
    if (len < 1e-6f) return;
    const float inv = 1.0f / len;
    x[0] = x[0] * inv;
    x[1] = x[1] * inv;
    x[2] = x[2] * inv;
  };
  const float eye[3] = {cam.position.x, cam.position.y, cam.position.z};
  const float up[3] = {cam.up.x, cam.up.y, cam.up.z};
  float f[3] = {cam.target.x - eye[0], cam.target.y - eye[1], cam.target.z - eye[2]};
  normalize(f);
  float r[3], u[3];
  cross(f, up, r);
  normalize(r);
  cross(r, f, u);
  v.m[0] = r[0]; v.m[1] = r[1]; v.m[2] = r[2]; v.m[3] = -dot(r, eye);
  v.m[4] = u[0]; v.m[5] = u[1]; v.m[6] = u[2]; v.m[7] = -dot(u, eye);
  v.m[8] = -f[0]; v.m[9] = -f[1]; v.m[10] = -f[2]; v.m[11] = dot(f, eye);
  const float fov_y = cam.fov_y_degrees * (float)3.14159265358979323846 / 180.0f;
  v.focal = (float)H / (2.0f * tanf(fov_y * 0.5f));
  v.half_w = (float)W * 0.5f;
  v.half_h = (float)H * 0.5f;
  v.near_plane = cfg.near_plane;
  v.far_plane = cfg.far_plane;
  v.eta = cfg.early_termination_alpha;
  v.sigma = cfg.gaussian_radius_sigma;
  v.width = W;
  v.height = H;
  v.ts = cfg.tile_size;
  v.tiles_x = (W + cfg.tile_size - 1) / cfg.tile_size;
  v.tiles_y = (H + cfg.tile_size - 1) / cfg.tile_size;
}

LumenStatus legacy_render(LumenExContext* ctx, LegacyBuffers& b, const float* aos, uint64_t n,
                          const LumenConfig& cfg, const LumenCamera& cam, LumenFrameBuffer* fb,
                          uint64_t stats[5]) {
  const uint32_t W = fb->width, H = fb->height;
  if (cfg.tile_size == 0 || cfg.tile_size > 65535u || W > 65535u || H > 65535u)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                "legacy mode: tile size %u, image %ux%u (1..65535 each)", cfg.tile_size, W, H);
  if (n >= 0xFFFFFFFFull) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "legacy mode: too many Gaussians");
  LegacyView v;
  legacy_view(cfg, cam, W, H, v);
  const uint64_t tiles = (uint64_t)v.tiles_x * v.tiles_y;
  if (tiles >= 0x7FFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "legacy mode: %llu tiles (at most 2^31 - 2)",
                (unsigned long long)tiles);
  const uint64_t pix = (uint64_t)W * H;
  hipStream_t st = nullptr;
  // the framebuffer's planes, read and updated in place
  LUMEN_TRY(ensure(b.rgb, pix * 3 ? pix * 3 : 1));
  LUMEN_TRY(ensure(b.depth, (pix ? pix : 1) * 4));
  LUMEN_TRY(ensure(b.alpha, (pix ? pix : 1) * 4));
  LUMEN_TRY(ensure(b.touched, 4));
  LUMEN_TRY(ensure(b.totals, 16));
  uint64_t V = 0, M = 0;
  if (n > 0) {
    const uint32_t nb = div_up(n, 256);
    LUMEN_TRY(ensure(b.proj, n * sizeof(LegacyRec)));
    LUMEN_TRY(ensure(b.flag, n * 4));
    LUMEN_TRY(ensure(b.off, n * 4));
    LUMEN_TRY(ensure(b.part, (uint64_t)div_up(n, kScanChunk) * 8));
    hipLaunchKernelGGL(k_legacy_project, dim3(nb), dim3(256), 0, st, aos, (uint32_t)n, v,
                       b.proj.as<LegacyRec>(), b.flag.as<uint32_t>());
    LUMEN_LAUNCHED();
    uint32_t* tot32 = b.totals.as<uint32_t>();
    LUMEN_TRY(exclusive_scan<uint32_t>(b.flag.as<uint32_t>(), b.off.as<uint32_t>(), n,
                                       b.part.as<uint32_t>(), tot32, st));
    uint32_t v32 = 0;
    LUMEN_HIP(hipMemcpy(&v32, tot32, 4, hipMemcpyDeviceToHost));
    V = v32;
  }
  if (V > 0 && tiles > 0) {
    LUMEN_TRY(ensure(b.rec, V * sizeof(LegacyRec)));
    LUMEN_TRY(ensure(b.rec_sorted, V * sizeof(LegacyRec)));
    LUMEN_TRY(ensure(b.dkeys, V * 8));
    LUMEN_TRY(ensure(b.dvals, V * 4));
    hipLaunchKernelGGL(k_legacy_compact, dim3(div_up(n, 256)), dim3(256), 0, st,
                       b.proj.as<LegacyRec>(), b.flag.as<uint32_t>(), b.off.as<uint32_t>(),
                       (uint32_t)n, b.rec.as<LegacyRec>(), b.dkeys.as<uint64_t>(),
                       b.dvals.as<uint32_t>());
    LUMEN_LAUNCHED();
    // ascending depth, equal depths in index order
    LUMEN_TRY(lumen_ex_radix_sort(ctx, b.dkeys.as<uint64_t>(), b.dvals.as<uint32_t>(), V, st));
    LUMEN_TRY(ensure(b.rect, V * 16));
    LUMEN_TRY(ensure(b.cnt, V * 8));
    LUMEN_TRY(ensure(b.koff, V * 8));
    LUMEN_TRY(ensure(b.kpart, (uint64_t)div_up(V, kScanChunk) * 8));
    hipLaunchKernelGGL(k_legacy_tiles, dim3(div_up(V, 256)), dim3(256), 0, st,
                       b.rec.as<LegacyRec>(), b.dvals.as<uint32_t>(), (uint32_t)V, v,
                       b.rec_sorted.as<LegacyRec>(), b.rect.as<uint4>(), b.cnt.as<uint64_t>());
    LUMEN_LAUNCHED();
    uint64_t* tot64 = reinterpret_cast<uint64_t*>(b.totals.as<uint8_t>() + 8);
    LUMEN_TRY(exclusive_scan<uint64_t>(b.cnt.as<uint64_t>(), b.koff.as<uint64_t>(), V,
                                       b.kpart.as<uint64_t>(), tot64, st));
    LUMEN_HIP(hipMemcpy(&M, tot64, 8, hipMemcpyDeviceToHost));
    if (M >= 0x7FFFFFFFull)
      return fail(LUMEN_ERROR_OUT_OF_MEMORY, "legacy mode: %llu (tile, splat) pairs exceed 2^31-1",
                  (unsigned long long)M);
  }
  LUMEN_TRY(ensure(b.ranges, (tiles ? tiles : 1) * 8));
  if (M > 0) {
    LUMEN_TRY(ensure(b.tkeys, M * 8));
    LUMEN_TRY(ensure(b.tvals, M * 4));
    hipLaunchKernelGGL(k_legacy_emit, dim3(div_up(M, 256)), dim3(256), 0, st,
                       b.koff.as<uint64_t>(), (uint32_t)V, M, b.rect.as<uint4>(), v.tiles_x,
                       b.tkeys.as<uint64_t>(), b.tvals.as<uint32_t>());
    LUMEN_LAUNCHED();
    // by tile; each tile's splats in depth order (the pairs' keys are unique)
    LUMEN_TRY(lumen_ex_radix_sort(ctx, b.tkeys.as<uint64_t>(), b.tvals.as<uint32_t>(), M, st));
  }
  LUMEN_TRY(lumen_ex_build_ranges(b.tkeys.as<uint64_t>(), M, (uint32_t)tiles, b.ranges.as<int32_t>(),
                                  st));
  LUMEN_TRY(fill<uint32_t>(b.touched.as<uint32_t>(), 0u, 1, st));
  uint32_t touched = 0;
  if (pix > 0) {
    LUMEN_HIP(hipMemcpy(b.rgb.ptr, fb->rgb_data, pix * 3, hipMemcpyHostToDevice));
    LUMEN_HIP(hipMemcpy(b.depth.ptr, fb->depth_data, pix * 4, hipMemcpyHostToDevice));
    LUMEN_HIP(hipMemcpy(b.alpha.ptr, fb->alpha_data, pix * 4, hipMemcpyHostToDevice));
    if (M > 0) {
      hipLaunchKernelGGL(k_legacy_composite, dim3((uint32_t)tiles), dim3(256), 0, st,
                         b.ranges.as<int2>(), b.tvals.as<uint32_t>(), b.rec_sorted.as<LegacyRec>(),
                         v, b.rgb.as<uint8_t>(), b.depth.as<float>(), b.alpha.as<float>(),
                         b.touched.as<uint32_t>());
      LUMEN_LAUNCHED();
    }
    LUMEN_HIP(hipMemcpy(fb->rgb_data, b.rgb.ptr, pix * 3, hipMemcpyDeviceToHost));
    LUMEN_HIP(hipMemcpy(fb->depth_data, b.depth.ptr, pix * 4, hipMemcpyDeviceToHost));
    LUMEN_HIP(hipMemcpy(fb->alpha_data, b.alpha.ptr, pix * 4, hipMemcpyDeviceToHost));
    LUMEN_HIP(hipMemcpy(&touched, b.touched.ptr, 4, hipMemcpyDeviceToHost));
  }
  stats[0] = n;
  stats[1] = V;
  stats[2] = n - V;
  stats[3] = tiles;
  stats[4] = touched;
  return LUMEN_SUCCESS;
}

}  // namespace lumen
