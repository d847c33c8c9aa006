// splat_math.hpp -- per-splat arithmetic of the rasterization path, shared by
// the device kernels and the host-side camera set-up.
//
// Every expression keeps the reference's operand order (left-associative, no
// FMA: the library is built with -ffp-contract=off) and uses correctly
// rounded '/' and sqrt, so integer decisions derived from floats (culling,
// radii, tile rectangles, depth keys) are bit-identical to the reference's
// Float32 arithmetic.  Sources are cited per function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lumen {

#define LUMEN_HD __host__ __device__ __forceinline__

// src/core/types.mojo:34-36
constexpr float kMinAlpha = 0.00392156862f;
constexpr float kTransmittanceEps = 0.0001f;

// src/math/spherical_harmonics.mojo:17-26
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f;
constexpr float kShC2_1 = 0.31539156525252005f;
constexpr float kShC2_2 = 0.5462742152960396f;
constexpr float kShC3_0 = 0.5900435899266435f;
constexpr float kShC3_1 = 2.890611442640554f;
constexpr float kShC3_2 = 0.4570457994644658f;
constexpr float kShC3_3 = 0.3731763325901154f;
constexpr float kShC3_4 = 1.445305721320277f;

struct Quat {
  float w, x, y, z;
};

// Int(Float32) made total: NaN -> 0, saturation at the int64 range.
LUMEN_HD int64_t trunc_to_i64(float f) {
  if (f != f) return 0;
  if (f >= 9.2233720368547758e18f) return INT64_MAX;
  if (f <= -9.2233720368547758e18f) return INT64_MIN;
  return (int64_t)f;
}

// src/math/quaternion.mojo:46-57
LUMEN_HD float quat_norm2(Quat q) { return q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z; }

LUMEN_HD Quat quat_normalized(Quat q) {
  float n = sqrtf(quat_norm2(q));
  if (n > 0.000001f) {
    float inv_n = 1.0f / n;
    return Quat{q.w * inv_n, q.x * inv_n, q.y * inv_n, q.z * inv_n};
  }
  return Quat{1.0f, 0.0f, 0.0f, 0.0f};
}

// src/math/quaternion.mojo:62-67
LUMEN_HD Quat quat_inverse(Quat q) {
  float ns = quat_norm2(q);
  if (ns > 0.000001f) {
    float inv_ns = 1.0f / ns;
    return Quat{q.w * inv_ns, -q.x * inv_ns, -q.y * inv_ns, -q.z * inv_ns};
  }
  return Quat{1.0f, 0.0f, 0.0f, 0.0f};
}

// Hamilton product, src/math/quaternion.mojo:69-78
LUMEN_HD Quat quat_mul(Quat a, Quat b) {
  return Quat{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
              a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}

// src/math/quaternion.mojo:89-117 (normalizes first).  Row-major 3x3.
LUMEN_HD void quat_matrix(Quat q0, float m[9]) {
  Quat q = quat_normalized(q0);
  float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  m[0] = 1.0f - (yy + zz);
  m[1] = xy - wz;
  m[2] = xz + wy;
  m[3] = xy + wz;
  m[4] = 1.0f - (xx + zz);
  m[5] = yz - wx;
  m[6] = xz - wy;
  m[7] = yz + wx;
  m[8] = 1.0f - (xx + yy);
}

// Per-view constants.  world_to_cam (src/scene/camera.mojo:69-71) rotates by
// q_view = inverse(normalized(camera rotation)) as q_view * (0,v) * inverse(q_view);
// both quaternions and the view rotation matrix (camera.mojo:66-67) are
// computed once per view on the host with the same functions.
struct ViewParams {
  int32_t width, height;
  int32_t tile_w, tile_h, tiles_x, tiles_y;
  float fx, fy, cx, cy, near_plane, far_plane, lowpass;
  float eye[3];
  Quat q_view, q_view_inv;
  float rot[9];
  int32_t sh_coeffs;  // 1, 4, 9 or 16 (src/math/spherical_harmonics.mojo:144-150)
  int32_t white_background, compute_depth;
  // Position of this view inside a multi-view frame: its tiles are global
  // tiles [tile_base, tile_base + tiles_x * tiles_y) and its pixels start at
  // pix_base in the batch's output planes.
  uint32_t tile_base;
  uint32_t view;
  uint64_t pix_base;
};

LUMEN_HD void make_view(ViewParams& v, const float eye[3], const float rot_wxyz[4]) {
  Quat qc = quat_normalized(Quat{rot_wxyz[0], rot_wxyz[1], rot_wxyz[2], rot_wxyz[3]});
  v.q_view = quat_inverse(qc);
  v.q_view_inv = quat_inverse(v.q_view);
  quat_matrix(v.q_view, v.rot);
  v.eye[0] = eye[0];
  v.eye[1] = eye[1];
  v.eye[2] = eye[2];
}

// Geometry of one Gaussian in one view: everything rasterize_scene decides
// before the SH colour (src/raster/pipeline.mojo:85-130).
struct SplatGeom {
  float sx, sy, depth;
  float c00, c01, c11;  // EWA screen covariance
  int32_t radius;
  int32_t min_tx, min_ty, max_tx, max_ty;
};

// Returns true when the Gaussian survives every cull and touches a tile.
LUMEN_HD bool splat_geometry(const ViewParams& v, float px, float py, float pz, float s0, float s1,
                             float s2, Quat graw, float opacity, SplatGeom& g) {
  // FrustumCuller.is_opacity_visible, src/raster/culling.mojo:47-51
  if (!(opacity >= kMinAlpha)) return false;
  // world_to_cam, src/scene/camera.mojo:69-71 + quaternion.mojo:80-87
  Quat p{0.0f, px - v.eye[0], py - v.eye[1], pz - v.eye[2]};
  Quat r = quat_mul(quat_mul(v.q_view, p), v.q_view_inv);
  const float tx = r.x, ty = r.y, tz = r.z;
  // is_depth_visible, culling.mojo:41-45
  if (!(tz >= v.near_plane && tz <= v.far_plane)) return false;
  // cam_to_screen, camera.mojo:73-79
  float z = tz;
  if (z < 0.0001f && z > -0.0001f) z = 0.0001f;
  const float sx = v.fx * (tx / z) + v.cx;
  const float sy = v.fy * (ty / z) + v.cy;
  // Covariance3D.from_scale_and_rotation, src/math/covariance.mojo:42-83;
  // the Gaussian's rotation is normalized at fetch (gaussian.mojo:36) and
  // again inside to_rotation_matrix.
  float R[9];
  quat_matrix(quat_normalized(graw), R);
  if (s0 < 0.00001f) s0 = 0.00001f;
  if (s1 < 0.00001f) s1 = 0.00001f;
  if (s2 < 0.00001f) s2 = 0.00001f;
  const float m00 = R[0] * s0, m01 = R[1] * s1, m02 = R[2] * s2;
  const float m10 = R[3] * s0, m11 = R[4] * s1, m12 = R[5] * s2;
  const float m20 = R[6] * s0, m21 = R[7] * s1, m22 = R[8] * s2;
  const float S00 = m00 * m00 + m01 * m01 + m02 * m02;
  const float S01 = m00 * m10 + m01 * m11 + m02 * m12;
  const float S02 = m00 * m20 + m01 * m21 + m02 * m22;
  const float S11 = m10 * m10 + m11 * m11 + m12 * m12;
  const float S12 = m10 * m20 + m11 * m21 + m12 * m22;
  const float S22 = m20 * m20 + m21 * m21 + m22 * m22;
  // ProjectiveJacobian.compute + compute_ewa_covariance, projection.mojo:80-148
  float jz = tz;
  if (jz < 0.0001f) jz = 0.0001f;
  const float itz = 1.0f / jz;
  const float itz2 = itz * itz;
  const float j00 = v.fx * itz, j02 = -v.fx * tx * itz2;
  const float j11 = v.fy * itz, j12 = -v.fy * ty * itz2;
  const float* W = v.rot;
  const float t00 = j00 * W[0] + j02 * W[6];
  const float t01 = j00 * W[1] + j02 * W[7];
  const float t02 = j00 * W[2] + j02 * W[8];
  const float t10 = j11 * W[3] + j12 * W[6];
  const float t11 = j11 * W[4] + j12 * W[7];
  const float t12 = j11 * W[5] + j12 * W[8];
  const float q00 = t00 * S00 + t01 * S01 + t02 * S02;
  const float q01 = t00 * S01 + t01 * S11 + t02 * S12;
  const float q02 = t00 * S02 + t01 * S12 + t02 * S22;
  const float q10 = t10 * S00 + t11 * S01 + t12 * S02;
  const float q11 = t10 * S01 + t11 * S11 + t12 * S12;
  const float q12 = t10 * S02 + t11 * S12 + t12 * S22;
  const float c00 = q00 * t00 + q01 * t01 + q02 * t02 + v.lowpass;
  const float c01 = q00 * t10 + q01 * t11 + q02 * t12;
  const float c11 = q10 * t10 + q11 * t11 + q12 * t12 + v.lowpass;
  // solve_eigenvalues_2x2, src/math/cholesky.mojo:28-51
  const float mid = (c00 + c11) * 0.5f;
  const float diff = (c00 - c11) * 0.5f;
  const float root = sqrtf(diff * diff + c01 * c01);
  const float l_max = mid + root;
  const int64_t rad = trunc_to_i64(ceilf(3.0f * sqrtf(l_max)));
  if (rad < 1) return false;  // pipeline.mojo:108
  // is_screen_box_visible, culling.mojo:53-71
  const float rf = (float)rad;
  if (sx + rf < 0.0f) return false;
  if (sx - rf >= (float)v.width) return false;
  if (sy + rf < 0.0f) return false;
  if (sy - rf >= (float)v.height) return false;
  // get_gaussian_tile_overlap, src/raster/tile.mojo:62-97
  int64_t min_px = trunc_to_i64(sx - rf);
  int64_t max_px = trunc_to_i64(sx + rf);
  int64_t min_py = trunc_to_i64(sy - rf);
  int64_t max_py = trunc_to_i64(sy + rf);
  if (max_px < 0 || min_px >= v.width || max_py < 0 || min_py >= v.height) return false;
  if (min_px < 0) min_px = 0;
  if (max_px >= v.width) max_px = v.width - 1;
  if (min_py < 0) min_py = 0;
  if (max_py >= v.height) max_py = v.height - 1;
  g.sx = sx;
  g.sy = sy;
  g.depth = tz;
  g.c00 = c00;
  g.c01 = c01;
  g.c11 = c11;
  g.radius = rad > INT32_MAX ? INT32_MAX : (int32_t)rad;
  g.min_tx = (int32_t)min_px / v.tile_w;
  g.max_tx = (int32_t)max_px / v.tile_w;
  g.min_ty = (int32_t)min_py / v.tile_h;
  g.max_ty = (int32_t)max_py / v.tile_h;
  return true;
}

// float_to_ordered_uint32, src/raster/radix_sort.mojo:8-19
LUMEN_HD uint32_t depth_key(float depth) {
  int64_t scaled = trunc_to_i64(depth * 1000.0f);
  if (scaled < 0) scaled = 0;
  else if (scaled > 0x7FFFFFFF) scaled = 0x7FFFFFFF;
  return (uint32_t)scaled;
}

// Covariance2D.inverse, src/math/projection.mojo:30-42
LUMEN_HD void inverse2d(float c00, float c01, float c11, float& i00, float& i01, float& i11) {
  const float det = c00 * c11 - c01 * c01;
  if (det > -0.0000001f && det < 0.0000001f) {
    i00 = 0.0f;
    i01 = 0.0f;
    i11 = 0.0f;
    return;
  }
  const float inv_det = 1.0f / det;
  i00 = c11 * inv_det;
  i01 = -c01 * inv_det;
  i11 = c00 * inv_det;
}

// Vec3f.normalized, src/math/vector.mojo:100-110
LUMEN_HD void normalize3(float& x, float& y, float& z) {
  const float l = sqrtf(x * x + y * y + z * z);
  if (l > 0.000001f) {
    x = x / l;
    y = y / l;
    z = z / l;
  } else {
    x = 0.0f;
    y = 0.0f;
    z = 0.0f;
  }
}

// SHBasis.evaluate, src/math/spherical_harmonics.mojo:68-111.  dir is the
// already-normalized view direction; it is normalized once more here as the
// reference does.
LUMEN_HD void sh_basis(float dx, float dy, float dz, float Y[16]) {
  normalize3(dx, dy, dz);
  const float x = dx, y = dy, z = dz;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, yz = y * z, xz = x * z;
  Y[0] = kShC0;
  Y[1] = -kShC1 * y;
  Y[2] = kShC1 * z;
  Y[3] = -kShC1 * x;
  Y[4] = kShC2_0 * xy;
  Y[5] = -kShC2_0 * yz;
  Y[6] = kShC2_1 * (2.0f * zz - xx - yy);
  Y[7] = -kShC2_0 * xz;
  Y[8] = kShC2_2 * (xx - yy);
  Y[9] = -kShC3_0 * y * (3.0f * xx - yy);
  Y[10] = kShC3_1 * xy * z;
  Y[11] = -kShC3_2 * y * (4.0f * zz - xx - yy);
  Y[12] = kShC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  Y[13] = -kShC3_2 * x * (4.0f * zz - xx - yy);
  Y[14] = kShC3_4 * z * (xx - yy);
  Y[15] = -kShC3_0 * x * (xx - 3.0f * yy);
}

// exp(x) by Cody-Waite reduction x = k*ln2 + r, a degree-6 polynomial for
// e^r on |r| <= ln2/2 (constant term 1; the others a minimax fit of the
// relative error, rounded to float32, max 2.0e-9 before rounding) evaluated
// in Horner form with fused multiply-adds, and scaling by 2^k.  Each step is
// one IEEE operation with one defined rounding (rint, fma, ldexp), so host
// and device builds return identical bits.  Against exp in float64 it is
// within 0.994 ulp at every float of the compositor's domain [-4.5, 0] and
// within 1.02 ulp over the normal range (six fused multiply-adds where a
// degree-7 Taylor polynomial needs seven).
LUMEN_HD float splat_exp(float x) {
  if (!(x >= -104.0f)) return (x != x) ? x : 0.0f;
  if (x > 88.8f) return __builtin_inff();
  const float k = rintf(x * 1.44269504f);
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.42860677e-06f, r);
  float p = 1.384365372e-3f;
  p = fmaf(p, r, 8.374155499e-3f);
  p = fmaf(p, r, 4.166800156e-2f);
  p = fmaf(p, r, 1.666643173e-1f);
  p = fmaf(p, r, 4.999999404e-1f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// splat_exp on the compositor's domain x = -power, power in [0, 4.5] (or
// NaN): the same operations without the range branches, so identical bits
// there (NaN in gives NaN out).
LUMEN_HD float splat_exp_neg(float x) {
  const float k = rintf(x * 1.44269504f);
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.42860677e-06f, r);
  float p = 1.384365372e-3f;
  p = fmaf(p, r, 8.374155499e-3f);
  p = fmaf(p, r, 4.166800156e-2f);
  p = fmaf(p, r, 1.666643173e-1f);
  p = fmaf(p, r, 4.999999404e-1f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// The largest power at which alpha = op * exp(-power), as the compositor
// computes it, can still reach 1/255: ln(255 op) plus a 1e-3 margin (four
// orders of magnitude above the rounding of splat_exp, the product and
// logf), capped at the reference's 4.5.  At a larger power alpha is below
// 1/255 (the blend skips it) or the power test zeroes it, so the compositor
// may skip the pixel before the exp with the same outcome.  NaN opacity
// gives 4.5 (the pixel is evaluated as before).
__device__ __forceinline__ float power_limit(float op) {
  const float l = logf(255.0f * op) + 0.001f;
  return l < 4.5f ? l : 4.5f;
}

// The splat's box for a power limit below 4.5: the record's half-widths are
// the 3-sigma (power 4.5) extents times 1.02 plus 0.02 px; the extents of
// power <= lim are sqrt(lim / 4.5) times the 3-sigma ones, with the same
// margins.  Infinite (unbounded) boxes stay infinite.
__device__ __forceinline__ void limit_box(float lim, float& hx, float& hy) {
  if (!(lim < 4.5f)) return;
  const float s = sqrtf(lim * (1.0f / 4.5f));
  hx = (hx - 0.02f) * s + 0.02f;
  hy = (hy - 0.02f) * s + 0.02f;
}

// The smallest power 0.5 (i00 dx^2 + 2 i01 dx dy + i11 dy^2), d = p - (sx,
// sy), over the rectangle [x0, x1] x [y0, y1]: 0 with the centre inside,
// else the least of the four edges' minima (the power is convex, so on an
// edge it is least at its stationary point clamped to the edge).
__device__ __forceinline__ float rect_min_power(float sx, float sy, float i00, float i01, float i11,
                                                float x0, float x1, float y0, float y1) {
  if (sx >= x0 && sx <= x1 && sy >= y0 && sy <= y1) return 0.0f;
  auto q = [&](float dx, float dy) {
    return 0.5f * (i00 * dx * dx + 2.0f * i01 * dx * dy + i11 * dy * dy);
  };
  const float r00 = 1.0f / i00, r11 = 1.0f / i11;
  float m = __builtin_inff();
  const float ys[2] = {y0, y1}, xs[2] = {x0, x1};
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float dy = ys[e] - sy;  // edge y = ys[e]: least at dx = -i01 dy / i00
    const float dx = fminf(fmaxf(-i01 * dy * r00, x0 - sx), x1 - sx);
    m = fminf(m, q(dx, dy));
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float dx = xs[e] - sx;  // edge x = xs[e]: least at dy = -i01 dx / i11
    const float dy = fminf(fmaxf(-i01 * dx * r11, y0 - sy), y1 - sy);
    m = fminf(m, q(dx, dy));
  }
  return m;
}

// Bit w of the result: splat (centre (sx, sy), inverse covariance i00, i01,
// i11, power limit lim) can reach lim at a pixel centre of wave w's strip of
// the tile (64 / tile_w full rows).  A wave whose bit is clear computes
// power > lim (no blend) for the splat at each of its pixels, so skipping
// the splat there changes nothing.  First the splat's box (centre +- (hx,
// hy), outside of which its power exceeds the limit at every pixel centre);
// then, for conics with condition number up to 1e4 (the box's own bound),
// the power's minimum over the strip's rectangle of pixel centres, which
// must exceed 1.02 lim + 1e-3 to clear the bit: a margin far above the
// rounding of either evaluation.  A rotated, elongated splat's box meets
// strips its ellipse misses (14% of the box's (splat, strip) pairs at
// 1M/1080p).
__device__ __forceinline__ uint32_t strip_mask(float sx, float sy, float hx, float hy, float i00,
                                               float i01, float i11, float lim, int tile_x0,
                                               int tile_y0, int tile_w, int rows, int nwaves) {
  const float x0 = (float)tile_x0 + 0.5f;
  const float x1 = (float)(tile_x0 + tile_w) - 0.5f;
  if (sx + hx < x0 || sx - hx > x1) return 0u;
  const float det = i00 * i11 - i01 * i01;
  const float tr = i00 + i11;
  // (a NaN centre or limit keeps the box's answer: such a splat blends NaN)
  const bool exact = i00 > 0.0f && i11 > 0.0f && det > 0.0f && tr * tr <= 1.0e4f * det &&
                     sx == sx && sy == sy && lim == lim;
  const float cut = 1.02f * lim + 1e-3f;
  uint32_t m = 0u;
  for (int w = 0; w < nwaves; ++w) {
    const float y0 = (float)(tile_y0 + w * rows) + 0.5f;
    const float y1 = (float)(tile_y0 + w * rows + rows) - 0.5f;
    if (sy + hy < y0 || sy - hy > y1) continue;
    if (exact && rect_min_power(sx, sy, i00, i01, i11, x0, x1, y0, y1) > cut) continue;
    m |= 1u << w;
  }
  return m;
}

// ---- the splat record -------------------------------------------------------
// k_emit writes one 64-byte record per visible splat; the compositors stage
// it per tile.  Everything that depends on the splat alone is computed there
// once, not per (splat, tile):
//   q0 = (x, y, i00 / 2, i01)        centre and the halved conic (the power
//   q1 = (i11 / 2, opacity, lim, cut)  0.5 (i00 dx dx + 2 i01 dx dy + i11 dy dy)
//   q2 = (r, g, b, depth)              with its factor 0.5 folded in: halving
//   q3 = (hx, hy, 1 / i00, 1 / i11)    commutes with each rounding of normal
//                                      floats, so the power keeps its bits)
// lim is power_limit(opacity), (hx, hy) the splat's box at that limit
// (limit_box), cut = 1.02 lim + 1e-3 the strip test's bound on the power's
// minimum over a wave's pixels, or +inf where that test does not apply
// (conics with condition number above 1e4, singular or indefinite ones, NaN
// centres or limits), and 1 / i00, 1 / i11 the stationary-point factors of
// rect_min_power.
struct SplatRecord {
  float4 q0, q1, q2, q3;
};

__device__ __forceinline__ SplatRecord make_record(float sx, float sy, float i00, float i01,
                                                   float i11, float op, float r, float g,
                                                   float b, float depth, float hx, float hy) {
  const float lim = power_limit(op);
  limit_box(lim, hx, hy);
  const float det = i00 * i11 - i01 * i01;
  const float tr = i00 + i11;
  const bool exact = i00 > 0.0f && i11 > 0.0f && det > 0.0f && tr * tr <= 1.0e4f * det &&
                     sx == sx && sy == sy && lim == lim;
  SplatRecord rec;
  rec.q0 = make_float4(sx, sy, 0.5f * i00, i01);
  rec.q1 = make_float4(0.5f * i11, op, lim, exact ? 1.02f * lim + 1e-3f : __builtin_inff());
  rec.q2 = make_float4(r, g, b, depth);
  rec.q3 = make_float4(hx, hy, 1.0f / i00, 1.0f / i11);
  return rec;
}

// rect_min_power with the record's reciprocals (i00 = 2 q0.z, i11 = 2 q1.x).
__device__ __forceinline__ float rect_min_power_rec(float sx, float sy, float i00, float i01,
                                                    float i11, float r00, float r11, float x0,
                                                    float x1, float y0, float y1) {
  if (sx >= x0 && sx <= x1 && sy >= y0 && sy <= y1) return 0.0f;
  auto q = [&](float dx, float dy) {
    return 0.5f * (i00 * dx * dx + 2.0f * i01 * dx * dy + i11 * dy * dy);
  };
  float m = __builtin_inff();
  const float ys[2] = {y0, y1}, xs[2] = {x0, x1};
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float dy = ys[e] - sy;  // edge y = ys[e]: least at dx = -i01 dy / i00
    const float dx = fminf(fmaxf(-i01 * dy * r00, x0 - sx), x1 - sx);
    m = fminf(m, q(dx, dy));
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const float dx = xs[e] - sx;  // edge x = xs[e]: least at dy = -i01 dx / i11
    const float dy = fminf(fmaxf(-i01 * dx * r11, y0 - sy), y1 - sy);
    m = fminf(m, q(dx, dy));
  }
  return m;
}

// The waves whose pixel centres a record's splat can reach at its power
// limit: bit w set unless the record's box misses wave w's block of qw x qh
// pixels (column w % nwx, row w / nwx of the tile) or the power's minimum
// over the block's pixel centres exceeds the record's cut.  The same answer
// strip_mask / block_mask give from the splat's raw quantities (full-width
// strips are blocks of qw = tile width).
// Whether a record's splat can reach the block of qw x qh pixels at (bx, by)
// (one wave's bit of record_mask; q0, q1 and q3 of the record are read).
__device__ __forceinline__ bool record_hits_block(const SplatRecord& r, int bx, int by, int qw,
                                                  int qh) {
  const float sx = r.q0.x, sy = r.q0.y, hx = r.q3.x, hy = r.q3.y, cut = r.q1.w;
  const float i00 = 2.0f * r.q0.z, i01 = r.q0.w, i11 = 2.0f * r.q1.x;
  const float x0 = (float)bx + 0.5f, x1 = (float)(bx + qw) - 0.5f;
  const float y0 = (float)by + 0.5f, y1 = (float)(by + qh) - 0.5f;
  if (sx + hx < x0 || sx - hx > x1 || sy + hy < y0 || sy - hy > y1) return false;
  return !(cut < __builtin_inff() &&
           rect_min_power_rec(sx, sy, i00, i01, i11, r.q3.z, r.q3.w, x0, x1, y0, y1) > cut);
}

__device__ __forceinline__ uint32_t record_mask(const SplatRecord& r, int tile_x0, int tile_y0,
                                                int qw, int qh, int nwx, int nwaves) {
  uint32_t m = 0u;
  for (int w = 0; w < nwaves; ++w)
    if (record_hits_block(r, tile_x0 + (w % nwx) * qw, tile_y0 + (w / nwx) * qh, qw, qh))
      m |= 1u << w;
  return m;
}

// strip_mask for waves that cover qw x qh blocks of the tile (wave w: block
// column w % nwx, row w / nwx): bit w set unless the splat's box misses the
// block's pixel centres or, for conics with condition number up to 1e4, its
// power's minimum over them exceeds 1.02 lim + 1e-3 (the same margins).
__device__ __forceinline__ uint32_t block_mask(float sx, float sy, float hx, float hy, float i00,
                                               float i01, float i11, float lim, int tile_x0,
                                               int tile_y0, int qw, int qh, int nwx, int nwaves) {
  const float det = i00 * i11 - i01 * i01;
  const float tr = i00 + i11;
  const bool exact = i00 > 0.0f && i11 > 0.0f && det > 0.0f && tr * tr <= 1.0e4f * det &&
                     sx == sx && sy == sy && lim == lim;
  const float cut = 1.02f * lim + 1e-3f;
  uint32_t m = 0u;
  for (int w = 0; w < nwaves; ++w) {
    const int bx = tile_x0 + (w % nwx) * qw, by = tile_y0 + (w / nwx) * qh;
    const float x0 = (float)bx + 0.5f, x1 = (float)(bx + qw) - 0.5f;
    const float y0 = (float)by + 0.5f, y1 = (float)(by + qh) - 0.5f;
    if (sx + hx < x0 || sx - hx > x1 || sy + hy < y0 || sy - hy > y1) continue;
    if (exact && rect_min_power(sx, sy, i00, i01, i11, x0, x1, y0, y1) > cut) continue;
    m |= 1u << w;
  }
  return m;
}

}  // namespace lumen
