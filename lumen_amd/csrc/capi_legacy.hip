// capi_legacy.hip -- the 22 functions of lumen.h (c_api/include/lumen.h:112-159).
//
// Cloud, framebuffer and rasterizer handles keep the reference's observable
// behaviour (c_api/src/lumen_c.c): NULL-safe destroy, doubling growth with a
// minimum of 16, framebuffer cleared on create to RGB 0 / depth 1000 /
// alpha 0, getters on NULL returning 0.  lumen_rasterize renders with the GPU
// splat pipeline of lumen_ex_render and composites the frame into the
// caller's 8-bit framebuffer, which it accumulates into (as the reference
// does, lumen_c.c:511-550).  Handles touch the GPU only inside
// lumen_rasterize.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "common.hpp"
#include "legacy.hpp"
#include "lumen.h"
#include "lumen_ex.h"

static_assert(sizeof(LumenConfig) == 28, "LumenConfig layout");
static_assert(sizeof(LumenVec3) == 12, "LumenVec3 layout");
static_assert(sizeof(LumenQuat) == 16, "LumenQuat layout");
static_assert(sizeof(LumenCamera) == 52, "LumenCamera layout");
static_assert(sizeof(LumenGaussian) == 236, "LumenGaussian layout");
static_assert(offsetof(LumenGaussian, sh_coeffs) == 44, "LumenGaussian layout");
static_assert(sizeof(LumenFrameBuffer) == 32, "LumenFrameBuffer layout");
static_assert(offsetof(LumenFrameBuffer, rgb_data) == 8, "LumenFrameBuffer layout");
static_assert(sizeof(LumenStats) == 48, "LumenStats layout");

struct LumenCloud {
  LumenGaussian* data;
  size_t count;
  size_t capacity;
};

struct LumenRasterizer {
  LumenConfig config;
  LumenExContext* ctx;
  lumen::DevBuf aos, planes, rgb, depth;
  int mode;                     // LUMEN_EX_RASTER_PIPELINE or LUMEN_EX_RASTER_LEGACY
  lumen::LegacyBuffers legacy;  // the legacy mode's device buffers
};

namespace {

constexpr float kPi = 3.14159265358979323846f;

// AoS LumenGaussian (59 floats) -> 11 SoA planes + SH block.
__global__ void k_aos_to_soa(const float* __restrict__ aos, uint64_t n, float* __restrict__ planes) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* g = aos + i * 59;
  // position(3) scale(3) rotation w,x,y,z(4) opacity(1) -> planes 0..10
#pragma unroll
  for (int k = 0; k < 11; ++k) planes[(uint64_t)k * n + i] = g[k];
  float* sh = planes + 11 * n + i * 48;
#pragma unroll 8
  for (int k = 0; k < 48; ++k) sh[k] = g[11 + k];
}

// Camera-to-world rotation with columns (right, down, forward) for a camera
// at `eye` looking at `target`; Shepperd's method for the quaternion.
void look_at_quat(const LumenVec3& eye, const LumenVec3& target, const LumenVec3& up, float q[4]) {
  double f[3] = {(double)target.x - eye.x, (double)target.y - eye.y, (double)target.z - eye.z};
  double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  if (fl < 1e-12) fl = 1.0;
  for (double& x : f) x /= fl;
  double u[3] = {up.x, up.y, up.z};
  double r[3] = {f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]};
  double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (rl < 1e-12) {  // up parallel to the view direction: pick any perpendicular axis
    double a[3] = {fabs(f[0]) < 0.9 ? 1.0 : 0.0, fabs(f[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    r[0] = f[1] * a[2] - f[2] * a[1];
    r[1] = f[2] * a[0] - f[0] * a[2];
    r[2] = f[0] * a[1] - f[1] * a[0];
    rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  }
  for (double& x : r) x /= rl;
  double d[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
  // R = [r | d | f] (columns)
  const double m00 = r[0], m01 = d[0], m02 = f[0];
  const double m10 = r[1], m11 = d[1], m12 = f[1];
  const double m20 = r[2], m21 = d[2], m22 = f[2];
  const double tr = m00 + m11 + m22;
  double w, x, y, z;
  if (tr > 0.0) {
    double s = 2.0 * sqrt(tr + 1.0);
    w = 0.25 * s; x = (m21 - m12) / s; y = (m02 - m20) / s; z = (m10 - m01) / s;
  } else if (m00 > m11 && m00 > m22) {
    double s = 2.0 * sqrt(1.0 + m00 - m11 - m22);
    w = (m21 - m12) / s; x = 0.25 * s; y = (m01 + m10) / s; z = (m02 + m20) / s;
  } else if (m11 > m22) {
    double s = 2.0 * sqrt(1.0 + m11 - m00 - m22);
    w = (m02 - m20) / s; x = (m01 + m10) / s; y = 0.25 * s; z = (m12 + m21) / s;
  } else {
    double s = 2.0 * sqrt(1.0 + m22 - m00 - m11);
    w = (m10 - m01) / s; x = (m02 + m20) / s; y = (m12 + m21) / s; z = 0.25 * s;
  }
  q[0] = (float)w; q[1] = (float)x; q[2] = (float)y; q[3] = (float)z;
}

}  // namespace

extern "C" {

const char* lumen_get_version(void) { return LUMEN_VERSION_STRING; }
int lumen_get_version_major(void) { return LUMEN_VERSION_MAJOR; }
int lumen_get_version_minor(void) { return LUMEN_VERSION_MINOR; }
int lumen_get_version_patch(void) { return LUMEN_VERSION_PATCH; }

// c_api/src/lumen_c.c:67-77
LumenConfig lumen_default_config(void) {
  LumenConfig c;
  c.image_width = 800;
  c.image_height = 600;
  c.tile_size = 16;
  c.near_plane = 0.1f;
  c.far_plane = 100.0f;
  c.early_termination_alpha = 0.99f;
  c.gaussian_radius_sigma = 3.0f;
  return c;
}

LumenCloud* lumen_cloud_create(size_t initial_capacity) {
  LumenCloud* c = (LumenCloud*)malloc(sizeof(LumenCloud));
  if (!c) return nullptr;
  c->capacity = initial_capacity > 0 ? initial_capacity : 16;
  c->data = (LumenGaussian*)malloc(c->capacity * sizeof(LumenGaussian));
  if (!c->data) {
    free(c);
    return nullptr;
  }
  c->count = 0;
  return c;
}

void lumen_cloud_destroy(LumenCloud* c) {
  if (!c) return;
  free(c->data);
  free(c);
}

LumenStatus lumen_cloud_reserve(LumenCloud* c, size_t cap) {
  if (!c) return LUMEN_ERROR_INVALID_ARGUMENT;
  if (cap <= c->capacity) return LUMEN_SUCCESS;
  LumenGaussian* p = (LumenGaussian*)realloc(c->data, cap * sizeof(LumenGaussian));
  if (!p) return LUMEN_ERROR_OUT_OF_MEMORY;
  c->data = p;
  c->capacity = cap;
  return LUMEN_SUCCESS;
}

size_t lumen_cloud_get_count(const LumenCloud* c) { return c ? c->count : 0; }
size_t lumen_cloud_get_capacity(const LumenCloud* c) { return c ? c->capacity : 0; }

LumenStatus lumen_cloud_add_gaussian(LumenCloud* c, const LumenGaussian* g) {
  if (!c || !g) return LUMEN_ERROR_INVALID_ARGUMENT;
  if (c->count >= c->capacity) {
    size_t next = c->capacity * 2;
    if (next < 16) next = 16;
    LumenStatus s = lumen_cloud_reserve(c, next);
    if (s != LUMEN_SUCCESS) return s;
  }
  c->data[c->count++] = *g;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_cloud_get_gaussian(const LumenCloud* c, size_t index, LumenGaussian* out) {
  if (!c || !out || index >= c->count) return LUMEN_ERROR_INVALID_ARGUMENT;
  *out = c->data[index];
  return LUMEN_SUCCESS;
}

void lumen_cloud_clear(LumenCloud* c) {
  if (c) c->count = 0;
}

// Lattice of points_per_axis^3 splats in x-major order, colour from the
// position (c_api/src/lumen_c.c:156-208).
LumenCloud* lumen_create_synthetic_cube(size_t n_axis, float size) {
  if (n_axis < 1) n_axis = 1;
  LumenCloud* c = lumen_cloud_create(n_axis * n_axis * n_axis);
  if (!c) return nullptr;
  const float step = n_axis > 1 ? size / (float)(n_axis - 1) : 0.0f;
  const float half = size * 0.5f;
  float s = step * 0.6f;
  if (s < 0.01f) s = 0.05f;
  const float c0 = 0.2820947917f;
  for (size_t ix = 0; ix < n_axis; ++ix)
    for (size_t iy = 0; iy < n_axis; ++iy)
      for (size_t iz = 0; iz < n_axis; ++iz) {
        LumenGaussian g;
        memset(&g, 0, sizeof(g));
        g.position = {-half + (float)ix * step, -half + (float)iy * step, -half + (float)iz * step};
        g.scale = {s, s, s};
        g.rotation = {1.0f, 0.0f, 0.0f, 0.0f};
        g.opacity = 0.85f;
        g.sh_coeffs[0] = ((g.position.x + half) / size - 0.5f) / c0;
        g.sh_coeffs[1] = ((g.position.y + half) / size - 0.5f) / c0;
        g.sh_coeffs[2] = ((g.position.z + half) / size - 0.5f) / c0;
        lumen_cloud_add_gaussian(c, &g);
      }
  return c;
}

// Fibonacci sphere (c_api/src/lumen_c.c:210-256).
LumenCloud* lumen_create_synthetic_sphere(size_t n, float radius) {
  if (n < 1) n = 1;
  LumenCloud* c = lumen_cloud_create(n);
  if (!c) return nullptr;
  const float golden = 1.6180339887f;
  const float c0 = 0.2820947917f;
  for (size_t i = 0; i < n; ++i) {
    const float theta = 2.0f * kPi * (float)i / golden;
    const float phi = acosf(1.0f - 2.0f * ((float)i + 0.5f) / (float)n);
    LumenGaussian g;
    memset(&g, 0, sizeof(g));
    g.position = {radius * sinf(phi) * cosf(theta), radius * sinf(phi) * sinf(theta),
                  radius * cosf(phi)};
    g.scale = {radius * 0.12f, radius * 0.12f, radius * 0.12f};
    g.rotation = {1.0f, 0.0f, 0.0f, 0.0f};
    g.opacity = 0.90f;
    g.sh_coeffs[0] = (0.5f + 0.5f * (g.position.x / radius) - 0.5f) / c0;
    g.sh_coeffs[1] = (0.5f + 0.5f * (g.position.y / radius) - 0.5f) / c0;
    g.sh_coeffs[2] = (0.5f + 0.5f * (g.position.z / radius) - 0.5f) / c0;
    lumen_cloud_add_gaussian(c, &g);
  }
  return c;
}

LumenFrameBuffer* lumen_framebuffer_create(uint32_t w, uint32_t h) {
  if (w == 0 || h == 0) return nullptr;
  LumenFrameBuffer* fb = (LumenFrameBuffer*)malloc(sizeof(LumenFrameBuffer));
  if (!fb) return nullptr;
  fb->width = w;
  fb->height = h;
  const size_t n = (size_t)w * h;
  fb->rgb_data = (uint8_t*)malloc(n * 3);
  fb->depth_data = (float*)malloc(n * sizeof(float));
  fb->alpha_data = (float*)malloc(n * sizeof(float));
  if (!fb->rgb_data || !fb->depth_data || !fb->alpha_data) {
    lumen_framebuffer_destroy(fb);
    return nullptr;
  }
  lumen_framebuffer_clear(fb, 0, 0, 0);
  return fb;
}

void lumen_framebuffer_destroy(LumenFrameBuffer* fb) {
  if (!fb) return;
  free(fb->rgb_data);
  free(fb->depth_data);
  free(fb->alpha_data);
  free(fb);
}

void lumen_framebuffer_clear(LumenFrameBuffer* fb, uint8_t r, uint8_t g, uint8_t b) {
  if (!fb) return;
  const size_t n = (size_t)fb->width * fb->height;
  for (size_t i = 0; i < n; ++i) {
    fb->rgb_data[3 * i + 0] = r;
    fb->rgb_data[3 * i + 1] = g;
    fb->rgb_data[3 * i + 2] = b;
    fb->depth_data[i] = 1000.0f;
    fb->alpha_data[i] = 0.0f;
  }
}

LumenRasterizer* lumen_rasterizer_create(const LumenConfig* config) {
  LumenRasterizer* r = new (std::nothrow) LumenRasterizer();
  if (!r) return nullptr;
  r->config = config ? *config : lumen_default_config();
  r->ctx = nullptr;
  r->mode = LUMEN_EX_RASTER_PIPELINE;
  return r;
}

void lumen_rasterizer_destroy(LumenRasterizer* r) {
  if (!r) return;
  if (r->ctx) {
    lumen::release(r->aos);
    lumen::release(r->planes);
    lumen::release(r->rgb);
    lumen::release(r->depth);
    lumen::legacy_release(r->legacy);
    lumen_ex_context_destroy(r->ctx);
  }
  delete r;
}

LumenStatus lumen_rasterizer_set_config(LumenRasterizer* r, const LumenConfig* config) {
  if (!r || !config) return LUMEN_ERROR_INVALID_ARGUMENT;
  r->config = *config;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_rasterizer_set_mode(LumenRasterizer* r, int mode) {
  if (!r || (mode != LUMEN_EX_RASTER_PIPELINE && mode != LUMEN_EX_RASTER_LEGACY))
    return lumen::fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_rasterizer_set_mode: bad argument");
  r->mode = mode;
  return LUMEN_SUCCESS;
}

int lumen_ex_rasterizer_get_mode(const LumenRasterizer* r) {
  return r ? r->mode : LUMEN_EX_RASTER_PIPELINE;
}

// lumen_rasterize in LUMEN_EX_RASTER_LEGACY mode: the reference C library's
// own renderer, bit for bit (legacy.hip).
static LumenStatus rasterize_legacy(LumenRasterizer* r, const LumenCloud* cloud,
                                    const LumenCamera* cam, LumenFrameBuffer* fb, LumenStats* stats,
                                    std::chrono::steady_clock::time_point t0) {
  const uint64_t n = cloud->count;
  LUMEN_TRY(lumen::ensure(r->aos, (n ? n : 1) * sizeof(LumenGaussian)));
  if (n) LUMEN_HIP(hipMemcpy(r->aos.ptr, cloud->data, n * sizeof(LumenGaussian), hipMemcpyHostToDevice));
  uint64_t st[5];
  LUMEN_TRY(lumen::legacy_render(r->ctx, r->legacy, r->aos.as<float>(), n, r->config, *cam, fb, st));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats->total_splats = st[0];
    stats->visible_splats = st[1];
    stats->culled_splats = st[2];
    stats->total_tiles = st[3];
    stats->touched_tiles = st[4];
    stats->rasterize_time_ms = (float)ms;
    stats->frames_per_second = ms > 0.001 ? (float)(1000.0 / ms) : 10000.0f;
  }
  return LUMEN_SUCCESS;
}

LumenStatus lumen_rasterize(LumenRasterizer* r, const LumenCloud* cloud, const LumenCamera* cam,
                            LumenFrameBuffer* fb, LumenStats* stats) {
  if (!r || !cloud || !cam || !fb) return LUMEN_ERROR_INVALID_ARGUMENT;
  const auto t0 = std::chrono::steady_clock::now();
  if (r->mode == LUMEN_EX_RASTER_LEGACY) {
    if (!r->ctx) {
      r->ctx = lumen_ex_context_create(0);
      if (!r->ctx) return LUMEN_ERROR_UNINITIALIZED;
    }
    LUMEN_HIP(hipSetDevice(0));
    return rasterize_legacy(r, cloud, cam, fb, stats, t0);
  }
  const uint32_t W = fb->width, H = fb->height;
  uint32_t ts = r->config.tile_size;
  if (ts * ts < 64 || ts * ts > 256) ts = 16;  // the GPU compositor needs 64..256 lanes per tile
  if (!r->ctx) {
    r->ctx = lumen_ex_context_create(0);
    if (!r->ctx) return LUMEN_ERROR_UNINITIALIZED;
  }
  const uint64_t n = cloud->count;
  const uint64_t pix = (uint64_t)W * H;
  LUMEN_TRY(lumen::ensure(r->aos, (n ? n : 1) * sizeof(LumenGaussian)));
  LUMEN_TRY(lumen::ensure(r->planes, (n ? n : 1) * 59 * sizeof(float)));
  LUMEN_TRY(lumen::ensure(r->rgb, pix * 3 * sizeof(float)));
  LUMEN_TRY(lumen::ensure(r->depth, pix * sizeof(float)));
  float* planes = r->planes.as<float>();
  if (n) {
    LUMEN_HIP(hipMemcpy(r->aos.ptr, cloud->data, n * sizeof(LumenGaussian), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(lumen::div_up(n, 256)), dim3(256), 0, nullptr,
                       r->aos.as<float>(), n, planes);
    LUMEN_LAUNCHED();
  }
  LumenExScene sc;
  sc.count = n;
  const float* p[12];
  for (int k = 0; k < 12; ++k) p[k] = planes + (uint64_t)k * n;
  sc.pos_x = p[0]; sc.pos_y = p[1]; sc.pos_z = p[2];
  sc.scale_x = p[3]; sc.scale_y = p[4]; sc.scale_z = p[5];
  sc.rot_w = p[6]; sc.rot_x = p[7]; sc.rot_y = p[8]; sc.rot_z = p[9];
  sc.opacity = p[10];
  sc.sh = p[11];
  // Focal length from the vertical field of view, principal point at the
  // image centre, clip planes from the rasterizer config (lumen_c.c:391-416).
  LumenExCamera ec;
  ec.width = W;
  ec.height = H;
  const float fov_y = cam->fov_y_degrees * kPi / 180.0f;
  ec.focal_y = (float)H / (2.0f * tanf(fov_y * 0.5f));
  ec.focal_x = ec.focal_y;
  ec.principal_x = (float)W * 0.5f;
  ec.principal_y = (float)H * 0.5f;
  ec.near_plane = r->config.near_plane;
  ec.far_plane = r->config.far_plane;
  ec.position[0] = cam->position.x;
  ec.position[1] = cam->position.y;
  ec.position[2] = cam->position.z;
  look_at_quat(cam->position, cam->target, cam->up, ec.rotation);
  LumenExRenderConfig cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.tile_width = ts;
  cfg.tile_height = ts;
  cfg.lowpass_filter = 0.3f;
  cfg.max_sh_degree = 3;
  cfg.white_background = 0;
  cfg.compute_depth = 1;
  cfg.flags = LUMEN_EX_SAVE_PIXELS;
  LumenExRenderStats es;
  LUMEN_TRY(lumen_ex_render(r->ctx, &sc, &ec, &cfg, r->rgb.as<float>(), r->depth.as<float>(), &es,
                            nullptr));
  std::vector<float> rgb(pix * 3), depth(pix), tfin(pix);
  std::vector<int32_t> ranges(2 * es.tiles_x * es.tiles_y);
  LUMEN_HIP(hipDeviceSynchronize());
  LUMEN_HIP(hipMemcpy(rgb.data(), r->rgb.ptr, pix * 3 * sizeof(float), hipMemcpyDeviceToHost));
  LUMEN_HIP(hipMemcpy(depth.data(), r->depth.ptr, pix * sizeof(float), hipMemcpyDeviceToHost));
  LUMEN_TRY(lumen_ex_copy_stage(r->ctx, LUMEN_EX_STAGE_T_FINAL, tfin.data(), pix * sizeof(float)));
  LUMEN_TRY(lumen_ex_copy_stage(r->ctx, LUMEN_EX_STAGE_RANGES, ranges.data(),
                                ranges.size() * sizeof(int32_t)));
  // Composite the new frame behind what the framebuffer already holds, in
  // the framebuffer's 8-bit colour and float alpha/depth planes.
  for (uint64_t i = 0; i < pix; ++i) {
    const float a_new = 1.0f - tfin[i];
    if (a_new <= 0.0f) continue;
    const float cur = fb->alpha_data[i];
    const float trans = 1.0f - cur;
    if (trans <= 0.0f) continue;
    for (int ch = 0; ch < 3; ++ch) {
      float v = (float)fb->rgb_data[3 * i + ch] / 255.0f + trans * rgb[3 * i + ch];
      if (v > 1.0f) v = 1.0f;
      fb->rgb_data[3 * i + ch] = (uint8_t)(v * 255.0f);
    }
    const float z = depth[i] / a_new;
    if (cur < 0.5f && z < fb->depth_data[i]) fb->depth_data[i] = z;
    fb->alpha_data[i] = cur + trans * a_new;
  }
  size_t touched = 0;
  for (size_t t = 0; t < ranges.size() / 2; ++t)
    if (ranges[2 * t] != -1 && ranges[2 * t] < ranges[2 * t + 1]) ++touched;
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats->total_splats = n;
    stats->visible_splats = es.visible_splats;
    stats->culled_splats = n - es.visible_splats;
    stats->total_tiles = es.tiles_x * es.tiles_y;
    stats->touched_tiles = touched;
    stats->rasterize_time_ms = (float)ms;
    stats->frames_per_second = ms > 0.001 ? (float)(1000.0 / ms) : 10000.0f;
  }
  return LUMEN_SUCCESS;
}

}  // extern "C"
