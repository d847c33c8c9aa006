// liblumen.hip -- forward splat rasterization on MI355X (gfx950) behind the
// lumen_ex.h C ABI.  The stages and the reference code they replace:
//
//   k_count      culling + covariance + EWA + 3-sigma radius + tile rectangle
//                per Gaussian; writes (tile count << 32 | visible) for the
//                compaction scan      src/raster/pipeline.mojo:84-130
//   scan         visible index and key offset per Gaussian (stable, so the
//                visible order is the reference's append order) :141-142
//   k_emit       same geometry again (bit-identical), SH colour, inverse
//                covariance, the 48-byte splat record at its visible index and
//                the 64-bit tile|depth keys      :132-157
//   sort         stable LSD radix sort over the varying digits    :166
//   k_ranges     per-tile [start, end)                            :169
//   k_composite  one workgroup per tile, one lane per pixel, splat records
//                staged through LDS in batches, per-pixel front-to-back blend
//                with T <= 1e-4 termination and a block-wide exit :172-229
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "common.hpp"
#include "lumen_ex.h"
#include "scan.hpp"
#include "sort.hpp"
#include "splat_math.hpp"
#include "context.hpp"

namespace lumen {

static thread_local std::string t_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
}
const char* last_error() { return t_last_error.c_str(); }
void clear_error() { t_last_error.clear(); }

__global__ void k_init_bounds(uint32_t* b) {
  b[0] = 0xFFFFFFFFu;
  b[1] = 0u;
  b[2] = 0xFFFFFFFFu;
  b[3] = 0u;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, d, 64));
  return x;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, d, 64));
  return x;
}

// Pass 1: visibility and tile count per Gaussian, plus the depth-key and
// tile-id bounds that decide which sort digits can vary.
__global__ __launch_bounds__(256) void k_count(SceneDev s, ViewParams v,
                                               uint64_t* __restrict__ packed,
                                               uint32_t* __restrict__ bounds) {
  __shared__ uint32_t sb[4][4];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  uint32_t dmin = 0xFFFFFFFFu, dmax = 0u, tmin = 0xFFFFFFFFu, tmax = 0u;
  if (i < s.n) {
    SplatGeom g;
    uint64_t val = 0;
    if (load_geometry(s, v, i, g)) {
      const uint32_t cnt =
          (uint32_t)(g.max_tx - g.min_tx + 1) * (uint32_t)(g.max_ty - g.min_ty + 1);
      val = ((uint64_t)cnt << 32) | 1ull;
      dmin = dmax = depth_key(g.depth);
      tmin = (uint32_t)(g.min_ty * v.tiles_x + g.min_tx);
      tmax = (uint32_t)(g.max_ty * v.tiles_x + g.max_tx);
    }
    packed[i] = val;
  }
  dmin = wave_min_u32(dmin);
  dmax = wave_max_u32(dmax);
  tmin = wave_min_u32(tmin);
  tmax = wave_max_u32(tmax);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sb[w][0] = dmin;
    sb[w][1] = dmax;
    sb[w][2] = tmin;
    sb[w][3] = tmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      dmin = min(dmin, sb[k][0]);
      dmax = max(dmax, sb[k][1]);
      tmin = min(tmin, sb[k][2]);
      tmax = max(tmax, sb[k][3]);
    }
    if (dmin != 0xFFFFFFFFu) {
      atomicMin(&bounds[0], dmin);
      atomicMax(&bounds[1], dmax);
      atomicMin(&bounds[2], tmin);
      atomicMax(&bounds[3], tmax);
    }
  }
}

// Pass 2: records and keys of visible Gaussians at their scanned offsets.
__global__ __launch_bounds__(256) void k_emit(SceneDev s, ViewParams v,
                                              const uint64_t* __restrict__ offsets,
                                              float4* __restrict__ rec,
                                              int32_t* __restrict__ vis2orig,
                                              int32_t* __restrict__ radius,
                                              uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= s.n) return;
  SplatGeom g;
  if (!load_geometry(s, v, i, g)) return;
  const uint64_t off = offsets[i];
  const uint32_t vi = (uint32_t)off;
  uint32_t ko = (uint32_t)(off >> 32);

  // evaluate_sh_color(sh, i*48, normalize(pos - eye), degree)
  // src/math/spherical_harmonics.mojo:133-172, pipeline.mojo:133-139
  float dx = s.px[i] - v.eye[0], dy = s.py[i] - v.eye[1], dz = s.pz[i] - v.eye[2];
  normalize3(dx, dy, dz);
  float Y[16];
  sh_basis(dx, dy, dz, Y);
  const float4* sh4 = reinterpret_cast<const float4*>(s.sh) + (size_t)i * 12u;
  float c[48];
  const int nfloat = 3 * v.sh_coeffs;
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    float4 x = (4 * q < nfloat) ? sh4[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    c[4 * q + 0] = x.x;
    c[4 * q + 1] = x.y;
    c[4 * q + 2] = x.z;
    c[4 * q + 3] = x.w;
  }
  float cr = 0.5f, cg = 0.5f, cb = 0.5f;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (k < v.sh_coeffs) {
      cr += c[3 * k + 0] * Y[k];
      cg += c[3 * k + 1] * Y[k];
      cb += c[3 * k + 2] * Y[k];
    }
  }
  if (cr < 0.0f) cr = 0.0f; else if (cr > 1.0f) cr = 1.0f;
  if (cg < 0.0f) cg = 0.0f; else if (cg > 1.0f) cg = 1.0f;
  if (cb < 0.0f) cb = 0.0f; else if (cb > 1.0f) cb = 1.0f;

  float i00, i01, i11;
  inverse2d(g.c00, g.c01, g.c11, i00, i01, i11);
  rec[3u * vi + 0u] = make_float4(g.sx, g.sy, i00, 2.0f * i01);
  rec[3u * vi + 1u] = make_float4(i11, s.op[i], cr, cg);
  rec[3u * vi + 2u] = make_float4(cb, g.depth, 0.0f, 0.0f);
  vis2orig[vi] = (int32_t)i;
  radius[vi] = g.radius;

  // Emission order: ty-major then tx (pipeline.mojo:152-157).
  const uint64_t dk = depth_key(g.depth);
  for (int ty = g.min_ty; ty <= g.max_ty; ++ty) {
    for (int tx = g.min_tx; tx <= g.max_tx; ++tx) {
      keys[ko] = ((uint64_t)(uint32_t)(ty * v.tiles_x + tx) << 32) | dk;
      vals[ko] = vi;
      ++ko;
    }
  }
}

__global__ __launch_bounds__(256) void k_ranges(const uint64_t* __restrict__ keys, uint32_t m,
                                                int2* __restrict__ ranges) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= m) return;
  const uint32_t tile = (uint32_t)(keys[i] >> 32);
  if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != tile) ranges[tile].x = (int32_t)i;
  if (i == m - 1 || (uint32_t)(keys[i + 1] >> 32) != tile) ranges[tile].y = (int32_t)(i + 1);
}

// One workgroup per tile, one lane per pixel (lane t -> (t % tile_w, t /
// tile_w), so a wave covers 64/tile_w full rows).  Splat records of the
// tile's sorted range are staged through LDS BATCH at a time; every lane
// walks the batch front to back exactly as PixelAccumulator does
// (src/raster/compositor.mojo:25-89).  A lane stops at T <= 1e-4; the block
// stops when all its lanes have.
template <int BATCH>
__global__ __launch_bounds__(256) void k_composite(const int2* __restrict__ ranges,
                                                   const uint32_t* __restrict__ vals,
                                                   const float4* __restrict__ rec, ViewParams v,
                                                   float* __restrict__ rgb,
                                                   float* __restrict__ depth,
                                                   float* __restrict__ t_final,
                                                   int32_t* __restrict__ n_eval) {
  __shared__ float4 s0[BATCH], s1[BATCH], s2[BATCH];
  const int nthreads = v.tile_w * v.tile_h;
  const int tile = blockIdx.x;
  const int tx = tile % v.tiles_x, ty = tile / v.tiles_x;
  const int t = threadIdx.x;
  const int px = tx * v.tile_w + t % v.tile_w;
  const int py = ty * v.tile_h + t / v.tile_w;
  const bool inside = px < v.width && py < v.height;
  const int2 rg = ranges[tile];
  float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, T = 1.0f, D = 0.0f;
  int32_t ne = 0;
  if (rg.x != -1 && rg.x < rg.y) {
    const float pxc = (float)px + 0.5f;
    const float pyc = (float)py + 0.5f;
    bool done = !inside;
    for (int base = rg.x; base < rg.y; base += BATCH) {
      if (__syncthreads_count(!done) == 0) break;
      const int n = min(BATCH, rg.y - base);
      for (int j = t; j < n; j += nthreads) {
        const uint32_t vi = vals[base + j];
        s0[j] = rec[3u * vi + 0u];
        s1[j] = rec[3u * vi + 1u];
        s2[j] = rec[3u * vi + 2u];
      }
      __syncthreads();
      if (!done) {
        for (int j = 0; j < n; ++j) {
          const float4 a = s0[j];
          const float4 b = s1[j];
          // compute_gaussian_pixel_alpha, compositor.mojo:62-89
          const float dx = pxc - a.x;
          const float dy = pyc - a.y;
          float power = 0.5f * (a.z * dx * dx + a.w * dx * dy + b.x * dy * dy);
          ++ne;
          if (power < 0.0f) power = 0.0f;
          if (power > 4.5f) continue;
          float alpha = b.y * splat_exp(-power);
          if (alpha > 0.99f) alpha = 0.99f;
          // PixelAccumulator.blend, compositor.mojo:25-53
          if (alpha < kMinAlpha) continue;
          const float4 c = s2[j];
          const float wgt = alpha * T;
          C0 += b.z * wgt;
          C1 += b.w * wgt;
          C2 += c.x * wgt;
          D += c.y * wgt;
          T *= (1.0f - alpha);
          if (T <= kTransmittanceEps) {
            done = true;
            break;
          }
        }
      }
    }
  }
  if (!inside) return;
  const size_t pix = (size_t)py * v.width + px;
  // PixelAccumulator.finalize + RenderOutput writes, pipeline.mojo:182-229
  const float bg = v.white_background ? T : 0.0f;
  rgb[3 * pix + 0] = C0 + bg;
  rgb[3 * pix + 1] = C1 + bg;
  rgb[3 * pix + 2] = C2 + bg;
  if (depth) depth[pix] = v.compute_depth ? D : 0.0f;
  if (t_final) t_final[pix] = T;
  if (n_eval) n_eval[pix] = ne;
}

__global__ void k_minmax_u64(const uint64_t* __restrict__ keys, uint64_t n,
                             unsigned long long* __restrict__ mm) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long lo = ~0ull, hi = 0ull;
  for (uint64_t j = i; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = keys[j];
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long a = __shfl_xor(lo, d, 64);
    const unsigned long long b = __shfl_xor(hi, d, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&mm[0], lo);
    atomicMax(&mm[1], hi);
  }
}

}  // namespace lumen

using namespace lumen;

static bool plane_ok(const float* p) { return p != nullptr; }

static LumenStatus composite_launch(uint32_t batch, uint32_t tiles, uint32_t threads,
                                    hipStream_t st, const int2* ranges, const uint32_t* vals,
                                    const float4* rec, const ViewParams& v, float* rgb,
                                    float* depth, float* t_final, int32_t* n_eval) {
  switch (batch) {
    case 64:
      hipLaunchKernelGGL(k_composite<64>, dim3(tiles), dim3(threads), 0, st, ranges, vals, rec, v,
                         rgb, depth, t_final, n_eval);
      break;
    case 128:
      hipLaunchKernelGGL(k_composite<128>, dim3(tiles), dim3(threads), 0, st, ranges, vals, rec, v,
                         rgb, depth, t_final, n_eval);
      break;
    default:
      hipLaunchKernelGGL(k_composite<256>, dim3(tiles), dim3(threads), 0, st, ranges, vals, rec, v,
                         rgb, depth, t_final, n_eval);
      break;
  }
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

static LumenStatus render_impl(LumenExContext* c, const LumenExScene* sc, const LumenExCamera* cam,
                               const LumenExRenderConfig* cfg, float* out_rgb, float* out_depth,
                               LumenExRenderStats* stats, hipStream_t st) {
  if (!c || !sc || !cam || !cfg || !out_rgb)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: null argument");
  if (cam->width == 0 || cam->height == 0 || cam->width > 16384 || cam->height > 16384)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: bad image size %ux%u",
                cam->width, cam->height);
  const uint32_t tw = cfg->tile_width, th = cfg->tile_height;
  const uint32_t threads = tw * th;
  if (tw == 0 || th == 0 || threads < 64 || threads > 256 || threads % 64 != 0)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                "lumen_ex_render: tile %ux%u unsupported (pixels per tile must be 64, 128, 192 "
                "or 256)",
                tw, th);
  const uint32_t batch = cfg->batch_depth ? cfg->batch_depth : 256u;
  if (batch != 64 && batch != 128 && batch != 256)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: batch_depth %u not in {64,128,256}",
                batch);
  if (cfg->max_sh_degree < 0 || cfg->max_sh_degree > 3)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: max_sh_degree %d not in 0..3",
                cfg->max_sh_degree);
  if (sc->count >= 0x7FFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: %llu Gaussians exceed 2^31-1",
                (unsigned long long)sc->count);
  if (sc->count > 0 &&
      !(plane_ok(sc->pos_x) && plane_ok(sc->pos_y) && plane_ok(sc->pos_z) &&
        plane_ok(sc->scale_x) && plane_ok(sc->scale_y) && plane_ok(sc->scale_z) &&
        plane_ok(sc->rot_w) && plane_ok(sc->rot_x) && plane_ok(sc->rot_y) && plane_ok(sc->rot_z) &&
        plane_ok(sc->opacity) && plane_ok(sc->sh)))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: null scene plane");
  LUMEN_HIP(hipSetDevice(c->device));

  ViewParams v;
  fill_view(v, cam, cfg);

  const uint32_t N = (uint32_t)sc->count;
  const uint32_t T = (uint32_t)v.tiles_x * (uint32_t)v.tiles_y;
  const uint64_t P = (uint64_t)cam->width * cam->height;
  for (bool& r : c->stage_ran) r = false;
  c->kept = false;
  c->pixels = false;

  SceneDev sd{sc->pos_x, sc->pos_y, sc->pos_z, sc->scale_x, sc->scale_y, sc->scale_z, sc->rot_w,
              sc->rot_x, sc->rot_y, sc->rot_z, sc->opacity, sc->sh, N};

  uint64_t V = 0, M = 0;
  uint32_t hb[4] = {0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u};
  if (N > 0) {
    LUMEN_TRY(ensure(c->packed, (uint64_t)N * 8));
    LUMEN_TRY(ensure(c->part, (uint64_t)div_up(N, kScanChunk) * 8));
    LUMEN_TRY(ensure(c->totals, 16));
    LUMEN_TRY(ensure(c->bounds, 16));
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_COUNT, false, st));
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(1), 0, st, c->bounds.as<uint32_t>());
    LUMEN_LAUNCHED();
    hipLaunchKernelGGL(k_count, dim3(div_up(N, 256)), dim3(256), 0, st, sd, v,
                       c->packed.as<uint64_t>(), c->bounds.as<uint32_t>());
    LUMEN_LAUNCHED();
    LUMEN_TRY(exclusive_scan<uint64_t>(c->packed.as<uint64_t>(), c->packed.as<uint64_t>(), N,
                                       c->part.as<uint64_t>(), c->totals.as<uint64_t>(), st));
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_COUNT, true, st));
    uint8_t* h = static_cast<uint8_t*>(c->pinned);
    LUMEN_HIP(hipMemcpyAsync(h, c->totals.ptr, 8, hipMemcpyDeviceToHost, st));
    LUMEN_HIP(hipMemcpyAsync(h + 8, c->bounds.ptr, 16, hipMemcpyDeviceToHost, st));
    LUMEN_HIP(hipStreamSynchronize(st));
    uint64_t tot;
    memcpy(&tot, h, 8);
    memcpy(hb, h + 8, 16);
    V = tot & 0xFFFFFFFFull;
    M = tot >> 32;
    if (M >= 0x7FFFFFFFull)
      return fail(LUMEN_EX_ERROR_CAPACITY, "lumen_ex_render: %llu tile instances exceed 2^31-1",
                  (unsigned long long)M);
  }

  LUMEN_TRY(ensure(c->ranges, (uint64_t)T * 8));
  int2* ranges = c->ranges.as<int2>();
  LUMEN_TRY(fill<int32_t>(c->ranges.as<int32_t>(), -1, 2ull * T, st));
  int nd = 0;
  int digits[8];
  int final_buf = 0;
  if (M > 0) {
    LUMEN_TRY(ensure(c->rec, V * 48));
    LUMEN_TRY(ensure(c->vis2orig, V * 4));
    LUMEN_TRY(ensure(c->radius, V * 4));
    for (int b = 0; b < 2; ++b) {
      LUMEN_TRY(ensure(c->keys[b], M * 8));
      LUMEN_TRY(ensure(c->vals[b], M * 4));
    }
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_EMIT, false, st));
    hipLaunchKernelGGL(k_emit, dim3(div_up(N, 256)), dim3(256), 0, st, sd, v,
                       c->packed.as<uint64_t>(), c->rec.as<float4>(), c->vis2orig.as<int32_t>(),
                       c->radius.as<int32_t>(), c->keys[0].as<uint64_t>(),
                       c->vals[0].as<uint32_t>());
    LUMEN_LAUNCHED();
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_EMIT, true, st));
    if (cfg->flags & LUMEN_EX_KEEP_STAGES) {
      LUMEN_TRY(ensure(c->keys_keep, M * 8));
      LUMEN_TRY(ensure(c->vals_keep, M * 4));
      LUMEN_HIP(hipMemcpyAsync(c->keys_keep.ptr, c->keys[0].ptr, M * 8, hipMemcpyDeviceToDevice, st));
      LUMEN_HIP(hipMemcpyAsync(c->vals_keep.ptr, c->vals[0].ptr, M * 4, hipMemcpyDeviceToDevice, st));
      c->kept = true;
    }
    nd = varying_digits(hb[0], hb[1], 0, 4, digits);
    nd += varying_digits(hb[2], hb[3], 4, 4, digits + nd);
    uint64_t* kb[2] = {c->keys[0].as<uint64_t>(), c->keys[1].as<uint64_t>()};
    uint32_t* vb[2] = {c->vals[0].as<uint32_t>(), c->vals[1].as<uint32_t>()};
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_SORT, false, st));
    LUMEN_TRY(radix_sort_digits(kb, vb, (uint32_t)M, digits, nd, SortScratch{&c->hist, &c->hist_part},
                                &final_buf, st));
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_SORT, true, st));
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_RANGES, false, st));
    hipLaunchKernelGGL(k_ranges, dim3(div_up(M, 256)), dim3(256), 0, st, kb[final_buf], (uint32_t)M,
                       ranges);
    LUMEN_LAUNCHED();
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_RANGES, true, st));
  }

  float* tf = nullptr;
  int32_t* ne = nullptr;
  if (cfg->flags & LUMEN_EX_SAVE_PIXELS) {
    LUMEN_TRY(ensure(c->t_final, P * 4));
    LUMEN_TRY(ensure(c->n_eval, P * 4));
    tf = c->t_final.as<float>();
    ne = c->n_eval.as<int32_t>();
    c->pixels = true;
  }
  // rasterize_scene returns its zero-filled RenderOutput before the
  // background fill when there is no tile instance (pipeline.mojo:63-65,
  // 162-163), so an empty frame is black even with white_background.
  if (M == 0) v.white_background = 0;
  const uint32_t* vals_sorted = M > 0 ? c->vals[final_buf].as<uint32_t>() : nullptr;
  const float4* rec = M > 0 ? c->rec.as<float4>() : nullptr;
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_COMPOSITE, false, st));
  LUMEN_TRY(composite_launch(batch, T, threads, st, ranges, vals_sorted, rec, v, out_rgb, out_depth,
                             tf, ne));
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_COMPOSITE, true, st));

  c->V = V;
  c->M = M;
  c->T = T;
  c->P = P;
  c->N = N;
  c->width = cam->width;
  c->height = cam->height;
  c->tile_w = tw;
  c->tile_h = th;
  c->final_buf = final_buf;
  c->has_grad = false;
  if (stats) {
    stats->visible_splats = V;
    stats->total_splat_instances = M;
    stats->tiles_x = (uint64_t)v.tiles_x;
    stats->tiles_y = (uint64_t)v.tiles_y;
    stats->sort_passes = (uint32_t)nd;
    stats->reserved = 0;
  }
  return LUMEN_SUCCESS;
}

extern "C" {

int lumen_ex_abi_version(void) { return LUMEN_EX_ABI_VERSION; }

const char* lumen_ex_last_error(void) { return lumen::last_error(); }

LumenExContext* lumen_ex_context_create(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    set_error("lumen_ex_context_create: device %d not available (%d devices)", device, count);
    return nullptr;
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_error("lumen_ex_context_create: hipSetDevice(%d) failed", device);
    return nullptr;
  }
  LumenExContext* c = new LumenExContext();
  c->device = device;
  if (hipHostMalloc(&c->pinned, 256, hipHostMallocDefault) != hipSuccess) {
    set_error("lumen_ex_context_create: hipHostMalloc failed");
    delete c;
    return nullptr;
  }
  return c;
}

void lumen_ex_context_destroy(LumenExContext* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  if (c->events_ready)
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  DevBuf* all[] = {&c->packed, &c->part, &c->totals, &c->bounds, &c->rec, &c->vis2orig,
                   &c->radius, &c->keys[0], &c->keys[1], &c->vals[0], &c->vals[1],
                   &c->keys_keep, &c->vals_keep, &c->hist, &c->hist_part, &c->ranges,
                   &c->t_final, &c->n_eval, &c->minmax, &c->grad_vis};
  for (DevBuf* b : all) release(*b);
  if (c->pinned) (void)hipHostFree(c->pinned);
  delete c;
}

LumenStatus lumen_ex_render(LumenExContext* ctx, const LumenExScene* scene,
                            const LumenExCamera* camera, const LumenExRenderConfig* config,
                            float* out_rgb, float* out_depth, LumenExRenderStats* out_stats,
                            void* stream) {
  clear_error();
  return render_impl(ctx, scene, camera, config, out_rgb, out_depth, out_stats,
                     (hipStream_t)stream);
}

LumenStatus lumen_ex_radix_sort(LumenExContext* c, uint64_t* keys, uint32_t* values, uint64_t n,
                                void* stream) {
  clear_error();
  if (!c || (n > 0 && (!keys || !values)))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_radix_sort: null argument");
  if (n >= 0xFFFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_radix_sort: n too large");
  if (n <= 1) return LUMEN_SUCCESS;
  hipStream_t st = (hipStream_t)stream;
  LUMEN_HIP(hipSetDevice(c->device));
  LUMEN_TRY(ensure(c->minmax, 16));
  unsigned long long* mm = c->minmax.as<unsigned long long>();
  LUMEN_TRY(fill<unsigned long long>(mm, ~0ull, 1, st));
  LUMEN_TRY(fill<unsigned long long>(mm + 1, 0ull, 1, st));
  const uint32_t blocks = div_up(n, 256) < 2048 ? div_up(n, 256) : 2048;
  hipLaunchKernelGGL(k_minmax_u64, dim3(blocks), dim3(256), 0, st, keys, n, mm);
  LUMEN_LAUNCHED();
  uint64_t hmm[2];
  LUMEN_HIP(hipMemcpyAsync(c->pinned, mm, 16, hipMemcpyDeviceToHost, st));
  LUMEN_HIP(hipStreamSynchronize(st));
  memcpy(hmm, c->pinned, 16);
  int digits[8];
  const int nd = varying_digits(hmm[0], hmm[1], 0, 8, digits);
  LUMEN_TRY(ensure(c->keys[1], n * 8));
  LUMEN_TRY(ensure(c->vals[1], n * 4));
  uint64_t* kb[2] = {keys, c->keys[1].as<uint64_t>()};
  uint32_t* vb[2] = {values, c->vals[1].as<uint32_t>()};
  int fb = 0;
  LUMEN_TRY(radix_sort_digits(kb, vb, (uint32_t)n, digits, nd, SortScratch{&c->hist, &c->hist_part},
                              &fb, st));
  if (fb == 1) {
    LUMEN_HIP(hipMemcpyAsync(keys, kb[1], n * 8, hipMemcpyDeviceToDevice, st));
    LUMEN_HIP(hipMemcpyAsync(values, vb[1], n * 4, hipMemcpyDeviceToDevice, st));
  }
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_build_ranges(const uint64_t* sorted_keys, uint64_t n, uint32_t tiles,
                                  int32_t* ranges, void* stream) {
  clear_error();
  if (!ranges || (n > 0 && !sorted_keys))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_build_ranges: null argument");
  if (n >= 0x7FFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_build_ranges: n too large");
  hipStream_t st = (hipStream_t)stream;
  LUMEN_TRY(fill<int32_t>(ranges, -1, 2ull * tiles, st));
  if (n > 0) {
    hipLaunchKernelGGL(k_ranges, dim3(div_up(n, 256)), dim3(256), 0, st, sorted_keys, (uint32_t)n,
                       reinterpret_cast<int2*>(ranges));
    LUMEN_LAUNCHED();
  }
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_copy_stage(LumenExContext* c, int stage, void* dst, size_t bytes) {
  clear_error();
  if (!c || !dst) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_copy_stage: null argument");
  const void* src = nullptr;
  size_t want = 0;
  switch (stage) {
    case LUMEN_EX_STAGE_VIS_TO_ORIG: src = c->vis2orig.ptr; want = c->V * 4; break;
    case LUMEN_EX_STAGE_RECORDS: src = c->rec.ptr; want = c->V * 48; break;
    case LUMEN_EX_STAGE_RADIUS: src = c->radius.ptr; want = c->V * 4; break;
    case LUMEN_EX_STAGE_KEYS_UNSORTED:
      if (!c->kept && c->M) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_KEEP_STAGES)");
      src = c->keys_keep.ptr; want = c->M * 8; break;
    case LUMEN_EX_STAGE_VALS_UNSORTED:
      if (!c->kept && c->M) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_KEEP_STAGES)");
      src = c->vals_keep.ptr; want = c->M * 4; break;
    case LUMEN_EX_STAGE_KEYS: src = c->keys[c->final_buf].ptr; want = c->M * 8; break;
    case LUMEN_EX_STAGE_VALS: src = c->vals[c->final_buf].ptr; want = c->M * 4; break;
    case LUMEN_EX_STAGE_RANGES: src = c->ranges.ptr; want = c->T * 8; break;
    case LUMEN_EX_STAGE_T_FINAL:
      if (!c->pixels) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_SAVE_PIXELS)");
      src = c->t_final.ptr; want = c->P * 4; break;
    case LUMEN_EX_STAGE_N_EVAL:
      if (!c->pixels) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_SAVE_PIXELS)");
      src = c->n_eval.ptr; want = c->P * 4; break;
    case LUMEN_EX_STAGE_GRAD_VIS:
      if (!c->has_grad) return fail(LUMEN_ERROR_UNINITIALIZED, "no backward since the last render");
      src = c->grad_vis.ptr; want = c->V * 40; break;
    default: return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_copy_stage: unknown stage %d", stage);
  }
  if (bytes != want)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_copy_stage: stage %d holds %zu bytes, got %zu",
                stage, want, bytes);
  if (want == 0) return LUMEN_SUCCESS;
  LUMEN_HIP(hipSetDevice(c->device));
  LUMEN_HIP(hipDeviceSynchronize());
  LUMEN_HIP(hipMemcpy(dst, src, want, hipMemcpyDeviceToHost));
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_set_profiling(LumenExContext* c, int enable) {
  clear_error();
  if (!c) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_set_profiling: null context");
  LUMEN_HIP(hipSetDevice(c->device));
  if (enable && !c->events_ready) {
    for (hipEvent_t& e : c->ev) LUMEN_HIP(hipEventCreate(&e));
    c->events_ready = true;
  }
  c->profile = enable != 0;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_stage_times(LumenExContext* c, float* ms, int n) {
  clear_error();
  if (!c || !ms) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_stage_times: null argument");
  if (!c->events_ready) return fail(LUMEN_ERROR_UNINITIALIZED, "profiling was never enabled");
  for (int s = 0; s < n && s < LUMEN_EX_TIME_STAGES; ++s) {
    ms[s] = 0.0f;
    if (!c->stage_ran[s]) continue;
    LUMEN_HIP(hipEventSynchronize(c->ev[2 * s + 1]));
    LUMEN_HIP(hipEventElapsedTime(&ms[s], c->ev[2 * s], c->ev[2 * s + 1]));
  }
  return LUMEN_SUCCESS;
}

}  // extern "C"
