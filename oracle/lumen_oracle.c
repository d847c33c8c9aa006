/*
 * lumen_oracle.c -- CPU oracle for the splat rasterization path.
 *
 * TEST INFRASTRUCTURE ONLY (see lumen_oracle.h).  Scalar, single-threaded
 * restatement of the reference's Mojo path.  Every function names the Mojo
 * source it restates.  Expressions are written with the reference's operand
 * order (left-associative, no FMA) so that every integer decision (culling,
 * radii, tile rectangles, depth keys) and every float output comes out bit
 * for bit as the reference's expression trees evaluate in IEEE binary32.
 *
 * Compile with -ffp-contract=off -fno-fast-math (oracle/Makefile).
 */
#include "lumen_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/core/types.mojo:34-36 */
#define MIN_ALPHA_THRESHOLD 0.00392156862f
#define TRANSMITTANCE_EPSILON 0.0001f

/* src/math/spherical_harmonics.mojo:17-26 */
#define SH_C0 0.28209479177387814f
#define SH_C1 0.4886025119029199f
#define SH_C2_0 1.0925484305920792f
#define SH_C2_1 0.31539156525252005f
#define SH_C2_2 0.5462742152960396f
#define SH_C3_0 0.5900435899266435f
#define SH_C3_1 2.890611442640554f
#define SH_C3_2 0.4570457994644658f
#define SH_C3_3 0.3731763325901154f
#define SH_C3_4 1.445305721320277f

/* Int(Float32) of the reference, made total: NaN -> 0 and saturation at the
 * int64 range (the reference leaves both undefined; the HIP kernels use the
 * same rule). */
static int64_t f2i_trunc(float f) {
    if (f != f) return 0;
    if (f >= 9.2233720368547758e18f) return INT64_MAX;
    if (f <= -9.2233720368547758e18f) return INT64_MIN;
    return (int64_t)f;
}

/* exp(x): Cody-Waite reduction x = k*ln2 + r (ln2 split in a 15-bit head and
 * a tail), a degree-6 polynomial of e^r on |r| <= ln2/2 in Horner form with
 * fused multiply-adds (constant term 1, the others a minimax fit of the
 * relative error rounded to float32: oracle/exp_poly_fit.py), then scaling
 * by 2^k.  Every step is an IEEE
 * operation with a single defined rounding (rint, fma, ldexp), so this
 * function returns identical bits on any IEEE machine; the HIP compositor
 * implements the same steps.  Max error vs exp in float64 is checked in
 * tests/test_oracle_kats.py. */
float orc_expf(float x) {
    if (!(x >= -104.0f)) return (x != x) ? x : 0.0f;
    if (x > 88.8f) return INFINITY;
    float k = rintf(x * 1.44269504f);
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

/* ------------------------------------------------------------------------ */
/* Quaternions: src/math/quaternion.mojo                                     */
/* ------------------------------------------------------------------------ */
typedef struct { float w, x, y, z; } quat;
typedef struct { float x, y, z; } vec3;

static float q_norm_squared(quat q) {                    /* :46-47 */
    return q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
}

static quat q_normalized(quat q) {                       /* :52-57 */
    float n = sqrtf(q_norm_squared(q));
    if (n > 0.000001f) {
        float inv_n = 1.0f / n;
        quat r = {q.w * inv_n, q.x * inv_n, q.y * inv_n, q.z * inv_n};
        return r;
    }
    quat id = {1.0f, 0.0f, 0.0f, 0.0f};
    return id;
}

static quat q_inverse(quat q) {                          /* :62-67 */
    float ns = q_norm_squared(q);
    if (ns > 0.000001f) {
        float inv_ns = 1.0f / ns;
        quat r = {q.w * inv_ns, -q.x * inv_ns, -q.y * inv_ns, -q.z * inv_ns};
        return r;
    }
    quat id = {1.0f, 0.0f, 0.0f, 0.0f};
    return id;
}

static quat q_mul(quat a, quat b) {                      /* :69-78 */
    quat r = {
        a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
        a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
        a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
        a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
    return r;
}

static vec3 q_rotate_vec(quat q, vec3 v) {               /* :80-87 */
    quat p = {0.0f, v.x, v.y, v.z};
    quat q_inv = q_inverse(q);
    quat rot = q_mul(q_mul(q, p), q_inv);
    vec3 r = {rot.x, rot.y, rot.z};
    return r;
}

static void q_to_matrix(quat q0, float m[9]) {           /* :89-117 */
    quat q = q_normalized(q0);
    float w = q.w, x = q.x, y = q.y, z = q.z;
    float x2 = x + x, y2 = y + y, z2 = z + z;
    float xx = x * x2, xy = x * y2, xz = x * z2;
    float yy = y * y2, yz = y * z2, zz = z * z2;
    float wx = w * x2, wy = w * y2, wz = w * z2;
    m[0] = 1.0f - (yy + zz); m[1] = xy - wz;          m[2] = xz + wy;
    m[3] = xy + wz;          m[4] = 1.0f - (xx + zz); m[5] = yz - wx;
    m[6] = xz - wy;          m[7] = yz + wx;          m[8] = 1.0f - (xx + yy);
}

static vec3 v_sub(vec3 a, vec3 b) {                      /* vector.mojo:78-79 */
    vec3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}

static vec3 v_normalized(vec3 v) {                       /* vector.mojo:100-110 */
    float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (l > 0.000001f) {
        vec3 r = {v.x / l, v.y / l, v.z / l};
        return r;
    }
    vec3 z = {0.0f, 0.0f, 0.0f};
    return z;
}

void orc_quat_normalized(const float q[4], float out[4]) {
    quat a = {q[0], q[1], q[2], q[3]};
    quat r = q_normalized(a);
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}

void orc_quat_to_matrix(const float q[4], float out[9]) {
    quat a = {q[0], q[1], q[2], q[3]};
    q_to_matrix(a, out);
}

void orc_quat_inverse(const float q[4], float out[4]) {
    quat a = {q[0], q[1], q[2], q[3]};
    quat r = q_inverse(a);
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}

void orc_quat_mul(const float a[4], const float b[4], float out[4]) {
    quat qa = {a[0], a[1], a[2], a[3]}, qb = {b[0], b[1], b[2], b[3]};
    quat r = q_mul(qa, qb);
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}

void orc_quat_rotate_vec(const float q[4], const float v[3], float out[3]) {
    quat a = {q[0], q[1], q[2], q[3]};
    vec3 b = {v[0], v[1], v[2]};
    vec3 r = q_rotate_vec(a, b);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* ------------------------------------------------------------------------ */
/* Camera: src/scene/camera.mojo                                             */
/* ------------------------------------------------------------------------ */
static quat cam_rotation(const orc_camera *cam) {        /* ctor :37-40 */
    quat q = {cam->rot[0], cam->rot[1], cam->rot[2], cam->rot[3]};
    return q_normalized(q);
}

static vec3 cam_world_to_cam(const orc_camera *cam, quat rot, vec3 p) { /* :69-71 */
    vec3 eye = {cam->pos[0], cam->pos[1], cam->pos[2]};
    vec3 rel = v_sub(p, eye);
    return q_rotate_vec(q_inverse(rot), rel);
}

static void cam_to_screen(const orc_camera *cam, vec3 pc, float *u, float *v) { /* :73-79 */
    float tz = pc.z;
    if (tz < 0.0001f && tz > -0.0001f) tz = 0.0001f;
    *u = cam->focal_x * (pc.x / tz) + cam->principal_x;
    *v = cam->focal_y * (pc.y / tz) + cam->principal_y;
}

void orc_world_to_cam(const orc_camera *cam, const float p[3], float out[3]) {
    vec3 pp = {p[0], p[1], p[2]};
    vec3 r = cam_world_to_cam(cam, cam_rotation(cam), pp);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_cam_to_screen(const orc_camera *cam, const float pc[3], float out[2]) {
    vec3 p = {pc[0], pc[1], pc[2]};
    cam_to_screen(cam, p, &out[0], &out[1]);
}

/* ------------------------------------------------------------------------ */
/* Covariance / EWA / eigen: src/math/{covariance,projection,cholesky}.mojo  */
/* ------------------------------------------------------------------------ */
void orc_cov3d(const float scale[3], const float qin[4], float c[6]) { /* covariance.mojo:42-83 */
    float sx = scale[0]; if (sx < 0.00001f) sx = 0.00001f;
    float sy = scale[1]; if (sy < 0.00001f) sy = 0.00001f;
    float sz = scale[2]; if (sz < 0.00001f) sz = 0.00001f;
    quat q = {qin[0], qin[1], qin[2], qin[3]};
    float R[9];
    q_to_matrix(q, R);
    float m00 = R[0] * sx, m01 = R[1] * sy, m02 = R[2] * sz;
    float m10 = R[3] * sx, m11 = R[4] * sy, m12 = R[5] * sz;
    float m20 = R[6] * sx, m21 = R[7] * sy, m22 = R[8] * sz;
    c[0] = m00 * m00 + m01 * m01 + m02 * m02;
    c[1] = m00 * m10 + m01 * m11 + m02 * m12;
    c[2] = m00 * m20 + m01 * m21 + m02 * m22;
    c[3] = m10 * m10 + m11 * m11 + m12 * m12;
    c[4] = m10 * m20 + m11 * m21 + m12 * m22;
    c[5] = m20 * m20 + m21 * m21 + m22 * m22;
}

void orc_jacobian(const float t[3], float fx, float fy, float j[6]) { /* projection.mojo:80-96 */
    float tz = t[2];
    if (tz < 0.0001f) tz = 0.0001f;
    float inv_tz = 1.0f / tz;
    float inv_tz2 = inv_tz * inv_tz;
    j[0] = fx * inv_tz;
    j[1] = 0.0f;
    j[2] = -fx * t[0] * inv_tz2;
    j[3] = 0.0f;
    j[4] = fy * inv_tz;
    j[5] = -fy * t[1] * inv_tz2;
}

void orc_ewa(const float s[6], const float pc[3], const float W[9], float fx,
             float fy, float lowpass, float out[3]) {     /* projection.mojo:99-148 */
    float J[6];
    orc_jacobian(pc, fx, fy, J);
    float t00 = J[0] * W[0] + J[2] * W[6];
    float t01 = J[0] * W[1] + J[2] * W[7];
    float t02 = J[0] * W[2] + J[2] * W[8];
    float t10 = J[4] * W[3] + J[5] * W[6];
    float t11 = J[4] * W[4] + J[5] * W[7];
    float t12 = J[4] * W[5] + J[5] * W[8];
    float s00 = s[0], s01 = s[1], s02 = s[2], s11 = s[3], s12 = s[4], s22 = s[5];
    float q00 = t00 * s00 + t01 * s01 + t02 * s02;
    float q01 = t00 * s01 + t01 * s11 + t02 * s12;
    float q02 = t00 * s02 + t01 * s12 + t02 * s22;
    float q10 = t10 * s00 + t11 * s01 + t12 * s02;
    float q11 = t10 * s01 + t11 * s11 + t12 * s12;
    float q12 = t10 * s02 + t11 * s12 + t12 * s22;
    out[0] = q00 * t00 + q01 * t01 + q02 * t02 + lowpass;
    out[1] = q00 * t10 + q01 * t11 + q02 * t12;
    out[2] = q10 * t10 + q11 * t11 + q12 * t12 + lowpass;
}

int64_t orc_eigen(const float cov[3], float lam[2]) {   /* cholesky.mojo:28-51 */
    float a = cov[0], b = cov[1], c = cov[2];
    float mid = (a + c) * 0.5f;
    float diff = (a - c) * 0.5f;
    float disc = diff * diff + b * b;
    float root = sqrtf(disc);
    float l_max = mid + root;
    float l_min = mid - root;
    if (l_min < 0.0f) l_min = 0.0f;
    if (lam) { lam[0] = l_max; lam[1] = l_min; }
    float r_float = 3.0f * sqrtf(l_max);
    return f2i_trunc(ceilf(r_float));
}

void orc_inverse2d(const float c[3], float out[3]) {    /* projection.mojo:30-42 */
    float det = c[0] * c[2] - c[1] * c[1];
    if (det > -0.0000001f && det < 0.0000001f) {
        out[0] = 0.0f; out[1] = 0.0f; out[2] = 0.0f;
        return;
    }
    float inv_det = 1.0f / det;
    out[0] = c[2] * inv_det;
    out[1] = -c[1] * inv_det;
    out[2] = c[0] * inv_det;
}

/* ------------------------------------------------------------------------ */
/* Spherical harmonics: src/math/spherical_harmonics.mojo:68-172             */
/* ------------------------------------------------------------------------ */
void orc_sh_basis(const float dir[3], float y[16]) {
    vec3 dv = {dir[0], dir[1], dir[2]};
    vec3 d = v_normalized(dv);
    float x = d.x, yy_ = d.y, z = d.z;
    float xx = x * x, yy = yy_ * yy_, zz = z * z;
    float xy = x * yy_, yz = yy_ * z, xz = x * z;
    y[0] = SH_C0;
    y[1] = -SH_C1 * yy_;
    y[2] = SH_C1 * z;
    y[3] = -SH_C1 * x;
    y[4] = SH_C2_0 * xy;
    y[5] = -SH_C2_0 * yz;
    y[6] = SH_C2_1 * (2.0f * zz - xx - yy);
    y[7] = -SH_C2_0 * xz;
    y[8] = SH_C2_2 * (xx - yy);
    y[9] = -SH_C3_0 * yy_ * (3.0f * xx - yy);
    y[10] = SH_C3_1 * xy * z;
    y[11] = -SH_C3_2 * yy_ * (4.0f * zz - xx - yy);
    y[12] = SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    y[13] = -SH_C3_2 * x * (4.0f * zz - xx - yy);
    y[14] = SH_C3_4 * z * (xx - yy);
    y[15] = -SH_C3_0 * x * (xx - 3.0f * yy);
}

void orc_sh_color(const float *sh, const float dir[3], int64_t degree, float out[3]) {
    float y[16];
    orc_sh_basis(dir, y);
    int64_t max_coeffs = 16;
    if (degree == 0) max_coeffs = 1;
    else if (degree == 1) max_coeffs = 4;
    else if (degree == 2) max_coeffs = 9;
    float r = 0.5f, g = 0.5f, b = 0.5f;
    for (int64_t i = 0; i < max_coeffs; ++i) {
        float yv = y[i];
        r += sh[i * 3 + 0] * yv;
        g += sh[i * 3 + 1] * yv;
        b += sh[i * 3 + 2] * yv;
    }
    if (r < 0.0f) r = 0.0f; else if (r > 1.0f) r = 1.0f;
    if (g < 0.0f) g = 0.0f; else if (g > 1.0f) g = 1.0f;
    if (b < 0.0f) b = 0.0f; else if (b > 1.0f) b = 1.0f;
    out[0] = r; out[1] = g; out[2] = b;
}

/* ------------------------------------------------------------------------ */
/* Culling and tiles: src/raster/{culling,tile}.mojo                         */
/* ------------------------------------------------------------------------ */
int orc_depth_visible(float near_plane, float far_plane, float z) { /* culling.mojo:41-45 */
    return z >= near_plane && z <= far_plane;
}

int orc_opacity_visible(float op) {                      /* culling.mojo:47-51 */
    return op >= MIN_ALPHA_THRESHOLD;
}

int orc_screen_box_visible(int64_t w, int64_t h, float cx, float cy, int64_t r) { /* culling.mojo:53-71 */
    float rf = (float)r;
    if (cx + rf < 0.0f) return 0;
    if (cx - rf >= (float)w) return 0;
    if (cy + rf < 0.0f) return 0;
    if (cy - rf >= (float)h) return 0;
    return 1;
}

/* Python floor division on Int (tile.mojo:56-60); operands are clamped to
 * be non-negative before it is used, so C division agrees. */
int orc_tile_overlap(int64_t w, int64_t h, int64_t tw, int64_t th, float cx,
                     float cy, int64_t r, int64_t out[4]) { /* tile.mojo:62-97 */
    float rf = (float)r;
    int64_t min_px = f2i_trunc(cx - rf);
    int64_t max_px = f2i_trunc(cx + rf);
    int64_t min_py = f2i_trunc(cy - rf);
    int64_t max_py = f2i_trunc(cy + rf);
    if (max_px < 0 || min_px >= w || max_py < 0 || min_py >= h) return 0;
    if (min_px < 0) min_px = 0;
    if (max_px >= w) max_px = w - 1;
    if (min_py < 0) min_py = 0;
    if (max_py >= h) max_py = h - 1;
    out[0] = min_px / tw;
    out[1] = min_py / th;
    out[2] = max_px / tw;
    out[3] = max_py / th;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Keys, sort, ranges: src/raster/{radix_sort,binning}.mojo                  */
/* ------------------------------------------------------------------------ */
uint64_t orc_encode_key(uint32_t tile, float depth) {    /* radix_sort.mojo:8-27 */
    int64_t scaled = f2i_trunc(depth * 1000.0f);
    if (scaled < 0) scaled = 0;
    else if (scaled > 0x7FFFFFFF) scaled = 0x7FFFFFFF;
    return ((uint64_t)tile << 32) | (uint64_t)scaled;
}

void orc_radix_sort(uint64_t *keys, int64_t *vals, int64_t n) { /* radix_sort.mojo:39-87 */
    if (n <= 1) return;
    uint64_t *kt = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    int64_t *vt = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int pass = 0; pass < 8; ++pass) {
        unsigned shift = (unsigned)(pass * 8);
        int64_t counts[256];
        memset(counts, 0, sizeof(counts));
        for (int64_t i = 0; i < n; ++i) counts[(keys[i] >> shift) & 0xFF]++;
        int64_t offsets[256];
        int64_t total = 0;
        for (int b = 0; b < 256; ++b) { offsets[b] = total; total += counts[b]; }
        for (int64_t i = 0; i < n; ++i) {
            uint64_t k = keys[i];
            int64_t dst = offsets[(k >> shift) & 0xFF]++;
            kt[dst] = k;
            vt[dst] = vals[i];
        }
        memcpy(keys, kt, (size_t)n * sizeof(uint64_t));
        memcpy(vals, vt, (size_t)n * sizeof(int64_t));
    }
    free(kt);
    free(vt);
}

void orc_build_ranges(const uint64_t *keys, int64_t n, int64_t tiles, int64_t *ranges) { /* binning.mojo:51-73 */
    for (int64_t t = 0; t < tiles; ++t) { ranges[2 * t] = -1; ranges[2 * t + 1] = -1; }
    if (n == 0) return;
    int64_t prev = -1;
    for (int64_t i = 0; i < n; ++i) {
        int64_t cur = (int64_t)((keys[i] >> 32) & 0xFFFFFFFFu);
        if (cur != prev) {
            if (prev >= 0 && prev < tiles) ranges[2 * prev + 1] = i;
            if (cur < tiles) { ranges[2 * cur] = i; ranges[2 * cur + 1] = n; }
            prev = cur;
        }
    }
    if (prev >= 0 && prev < tiles) ranges[2 * prev + 1] = n;
}

/* ------------------------------------------------------------------------ */
/* Compositor: src/raster/compositor.mojo                                   */
/* ------------------------------------------------------------------------ */
static float alpha_at(float px, float py, float cx, float cy, float i00, float i01,
                      float i11, float op, int32_t exp_mode) { /* :62-89 */
    float dx = px - cx;
    float dy = py - cy;
    float power = 0.5f * (i00 * dx * dx + 2.0f * i01 * dx * dy + i11 * dy * dy);
    if (power < 0.0f) power = 0.0f;
    else if (power > 4.5f) return 0.0f;
    float g = exp_mode ? expf(-power) : orc_expf(-power);
    float alpha = op * g;
    if (alpha > 0.99f) alpha = 0.99f;
    return alpha;
}

float orc_pixel_alpha(float px, float py, float cx, float cy, const float inv[3],
                      float op, int32_t exp_mode) {
    return alpha_at(px, py, cx, cy, inv[0], inv[1], inv[2], op, exp_mode);
}

/* PixelAccumulator.blend (:25-53).  s = {r, g, b, T, depth}. */
static int blend(float s[5], float cr, float cg, float cb, float a, float z) {
    if (a < MIN_ALPHA_THRESHOLD) return 0;
    float w = a * s[3];
    s[0] += cr * w;
    s[1] += cg * w;
    s[2] += cb * w;
    s[4] += z * w;
    s[3] *= (1.0f - a);
    return s[3] <= TRANSMITTANCE_EPSILON;
}

int64_t orc_blend_sequence(const float *colors, const float *alphas, const float *depths,
                           int64_t n, float s[5]) {
    for (int64_t i = 0; i < n; ++i)
        if (blend(s, colors[3 * i], colors[3 * i + 1], colors[3 * i + 2], alphas[i], depths[i]))
            return i;
    return -1;
}

/* ------------------------------------------------------------------------ */
/* rasterize_scene: src/raster/pipeline.mojo:48-231                          */
/* ------------------------------------------------------------------------ */
void orc_result_free(orc_result *r) {
    if (!r) return;
    free(r->vis_to_orig); free(r->proj); free(r->radius); free(r->rect);
    free(r->keys_unsorted); free(r->vals_unsorted); free(r->keys); free(r->vals);
    free(r->ranges); free(r->rgb); free(r->depth); free(r->t_final); free(r->n_eval);
    free(r);
}

typedef struct { uint64_t *k; int64_t *v; int64_t n, cap; } keylist;

static void kl_push(keylist *l, uint64_t k, int64_t v) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 1024;
        l->k = (uint64_t *)realloc(l->k, (size_t)l->cap * sizeof(uint64_t));
        l->v = (int64_t *)realloc(l->v, (size_t)l->cap * sizeof(int64_t));
    }
    l->k[l->n] = k;
    l->v[l->n] = v;
    l->n++;
}

orc_result *orc_rasterize(const orc_scene *sc, const orc_camera *cam, const orc_config *cfg) {
    orc_result *out = (orc_result *)calloc(1, sizeof(orc_result));
    const int64_t W = cam->width, H = cam->height;
    const int64_t tw = cfg->tile_width, th = cfg->tile_height;
    const int64_t tiles_x = (W + tw - 1) / tw, tiles_y = (H + th - 1) / th;
    const int64_t tiles = tiles_x * tiles_y;
    const int64_t N = sc->n;
    out->n = N; out->width = W; out->height = H;
    out->tiles = tiles; out->tiles_x = tiles_x; out->tiles_y = tiles_y;
    out->rgb = (float *)calloc((size_t)(W * H * 3), sizeof(float));    /* RenderOutput :29-35 */
    out->depth = (float *)calloc((size_t)(W * H), sizeof(float));
    out->t_final = (float *)malloc((size_t)(W * H) * sizeof(float));
    out->n_eval = (int32_t *)calloc((size_t)(W * H), sizeof(int32_t));
    out->ranges = (int64_t *)malloc((size_t)(2 * tiles) * sizeof(int64_t));
    for (int64_t i = 0; i < W * H; ++i) out->t_final[i] = 1.0f;
    for (int64_t t = 0; t < tiles; ++t) { out->ranges[2 * t] = -1; out->ranges[2 * t + 1] = -1; }
    out->vis_to_orig = (int64_t *)malloc((size_t)(N > 0 ? N : 1) * sizeof(int64_t));
    out->proj = (float *)malloc((size_t)(N > 0 ? N : 1) * 10 * sizeof(float));
    out->radius = (int64_t *)malloc((size_t)(N > 0 ? N : 1) * sizeof(int64_t));
    out->rect = (int64_t *)malloc((size_t)(N > 0 ? N : 1) * 4 * sizeof(int64_t));
    if (N == 0) return out;                                  /* :63-65 */

    quat crot = cam_rotation(cam);
    float view_rot[9];
    q_to_matrix(q_inverse(crot), view_rot);                 /* get_view_rotation camera.mojo:66-67 */
    vec3 eye = {cam->pos[0], cam->pos[1], cam->pos[2]};

    keylist kl = {0, 0, 0, 0};
    int64_t V = 0;
    for (int64_t i = 0; i < N; ++i) {                        /* :84-157 */
        vec3 pos = {sc->pos_x[i], sc->pos_y[i], sc->pos_z[i]};
        quat graw = {sc->rot_w[i], sc->rot_x[i], sc->rot_y[i], sc->rot_z[i]};
        quat grot = q_normalized(graw);                     /* gaussian.mojo:36 via get_gaussian */
        float op = sc->opacity[i];
        if (!orc_opacity_visible(op)) continue;
        vec3 pc = cam_world_to_cam(cam, crot, pos);
        if (!orc_depth_visible(cam->near_plane, cam->far_plane, pc.z)) continue;
        float sx, sy;
        cam_to_screen(cam, pc, &sx, &sy);
        float scale[3] = {sc->scale_x[i], sc->scale_y[i], sc->scale_z[i]};
        float qv[4] = {grot.w, grot.x, grot.y, grot.z};
        float c3[6], c2[3], pcv[3] = {pc.x, pc.y, pc.z};
        orc_cov3d(scale, qv, c3);
        orc_ewa(c3, pcv, view_rot, cam->focal_x, cam->focal_y, cfg->lowpass_filter, c2);
        int64_t r = orc_eigen(c2, NULL);
        if (r < 1) continue;
        if (!orc_screen_box_visible(W, H, sx, sy, r)) continue;
        int64_t rect[4];
        if (!orc_tile_overlap(W, H, tw, th, sx, sy, r, rect)) continue;
        vec3 vd = v_normalized(v_sub(pos, eye));
        float dir[3] = {vd.x, vd.y, vd.z}, col[3], inv[3];
        orc_sh_color(sc->sh + i * 48, dir, cfg->max_sh_degree, col);
        orc_inverse2d(c2, inv);
        float *p = out->proj + V * 10;
        p[0] = sx; p[1] = sy; p[2] = pc.z;
        p[3] = inv[0]; p[4] = inv[1]; p[5] = inv[2];
        p[6] = col[0]; p[7] = col[1]; p[8] = col[2]; p[9] = op;
        out->vis_to_orig[V] = i;
        out->radius[V] = r;
        memcpy(out->rect + 4 * V, rect, sizeof(rect));
        for (int64_t ty = rect[1]; ty <= rect[3]; ++ty)
            for (int64_t tx = rect[0]; tx <= rect[2]; ++tx)
                kl_push(&kl, orc_encode_key((uint32_t)(ty * tiles_x + tx), pc.z), V);
        V++;
    }
    const int64_t M = kl.n;
    out->visible = V;
    out->instances = M;
    size_t mcap = (size_t)(M > 0 ? M : 1);
    out->keys_unsorted = (uint64_t *)malloc(mcap * sizeof(uint64_t));
    out->vals_unsorted = (int64_t *)malloc(mcap * sizeof(int64_t));
    out->keys = (uint64_t *)malloc(mcap * sizeof(uint64_t));
    out->vals = (int64_t *)malloc(mcap * sizeof(int64_t));
    if (M > 0) {
        memcpy(out->keys_unsorted, kl.k, (size_t)M * sizeof(uint64_t));
        memcpy(out->vals_unsorted, kl.v, (size_t)M * sizeof(int64_t));
        memcpy(out->keys, kl.k, (size_t)M * sizeof(uint64_t));
        memcpy(out->vals, kl.v, (size_t)M * sizeof(int64_t));
    }
    free(kl.k);
    free(kl.v);
    if (M == 0) return out;                                  /* :162-163 */

    orc_radix_sort(out->keys, out->vals, M);                 /* :166 */
    orc_build_ranges(out->keys, M, tiles, out->ranges);      /* :169 */

    int64_t E = 0;
    for (int64_t ty = 0; ty < tiles_y; ++ty) {               /* :172-229 */
        for (int64_t tx = 0; tx < tiles_x; ++tx) {
            int64_t tile = ty * tiles_x + tx;
            int64_t rs = out->ranges[2 * tile], re = out->ranges[2 * tile + 1];
            int64_t px0 = tx * tw, py0 = ty * th;
            int64_t px1 = (tx + 1) * tw; if (px1 > W) px1 = W;
            int64_t py1 = (ty + 1) * th; if (py1 > H) py1 = H;
            if (rs == -1 || rs >= re) {
                if (cfg->white_background)
                    for (int64_t py = py0; py < py1; ++py)
                        for (int64_t px = px0; px < px1; ++px) {
                            float *c = out->rgb + (py * W + px) * 3;
                            c[0] = 1.0f; c[1] = 1.0f; c[2] = 1.0f;
                        }
                continue;
            }
            for (int64_t py = py0; py < py1; ++py) {
                float pyc = (float)py + 0.5f;
                for (int64_t px = px0; px < px1; ++px) {
                    float pxc = (float)px + 0.5f;
                    float s[5] = {0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
                    int32_t ne = 0;
                    for (int64_t k = rs; k < re; ++k) {
                        const float *p = out->proj + out->vals[k] * 10;
                        float a = alpha_at(pxc, pyc, p[0], p[1], p[3], p[4], p[5], p[9], cfg->exp_mode);
                        ne++;
                        if (blend(s, p[6], p[7], p[8], a, p[2])) break;
                    }
                    E += ne;
                    if (cfg->white_background) { s[0] += s[3]; s[1] += s[3]; s[2] += s[3]; }
                    float *c = out->rgb + (py * W + px) * 3;
                    c[0] = s[0]; c[1] = s[1]; c[2] = s[2];
                    if (cfg->compute_depth) out->depth[py * W + px] = s[4];
                    out->t_final[py * W + px] = s[3];
                    out->n_eval[py * W + px] = ne;
                }
            }
        }
    }
    out->evaluations = E;
    return out;
}

/* ------------------------------------------------------------------------ */
/* Losses: src/optim/loss.mojo                                               */
/* ------------------------------------------------------------------------ */
float orc_l1_loss(const float *pred, const float *target, int64_t n, float *grad) { /* :11-35 */
    if (n == 0) return 0.0f;
    float sum = 0.0f;
    float inv_n = 1.0f / (float)n;
    for (int64_t i = 0; i < n; ++i) {
        float diff = pred[i] - target[i];
        if (diff >= 0.0f) { sum += diff; if (grad) grad[i] = inv_n; }
        else { sum -= diff; if (grad) grad[i] = -inv_n; }
    }
    return sum * inv_n;
}

float orc_psnr(const float *pred, const float *target, int64_t n) { /* :38-62 */
    if (n == 0) return 99.99f;
    float s = 0.0f;
    for (int64_t i = 0; i < n; ++i) { float d = pred[i] - target[i]; s += d * d; }
    float mse = s / (float)n;
    if (mse < 1e-10f) return 99.99f;
    return -4.3429448f * logf(mse);
}

float orc_ssim(const float *pred, const float *target, int64_t width, int64_t height,
               int64_t win) {                            /* :65-133 */
    const float C1 = 0.0001f, C2 = 0.0009f;
    float total = 0.0f;
    int64_t count = 0;
    float inv_win = 1.0f / (float)(win * win * 3);
    for (int64_t y = 0; y + win <= height; y += win) {
        for (int64_t x = 0; x + win <= width; x += win) {
            float mx = 0.0f, my = 0.0f;
            for (int64_t wy = 0; wy < win; ++wy)
                for (int64_t wx = 0; wx < win; ++wx) {
                    int64_t b = ((y + wy) * width + (x + wx)) * 3;
                    mx += pred[b] + pred[b + 1] + pred[b + 2];
                    my += target[b] + target[b + 1] + target[b + 2];
                }
            mx *= inv_win;
            my *= inv_win;
            float vx = 0.0f, vy = 0.0f, cxy = 0.0f;
            for (int64_t wy = 0; wy < win; ++wy)
                for (int64_t wx = 0; wx < win; ++wx) {
                    int64_t b = ((y + wy) * width + (x + wx)) * 3;
                    for (int c = 0; c < 3; ++c) {
                        float dx = pred[b + c] - mx;
                        float dy = target[b + c] - my;
                        vx += dx * dx;
                        vy += dy * dy;
                        cxy += dx * dy;
                    }
                }
            vx *= inv_win;
            vy *= inv_win;
            cxy *= inv_win;
            float num = (2.0f * mx * my + C1) * (2.0f * cxy + C2);
            float den = (mx * mx + my * my + C1) * (vx + vy + C2);
            total += num / den;
            count++;
        }
    }
    if (count == 0) return 1.0f;
    return total / (float)count;
}

float orc_combined_loss(float l1, float ssim, float lambda_dssim) { /* :136-144 */
    return (1.0f - lambda_dssim) * l1 + lambda_dssim * (1.0f - ssim);
}

/* Vectorised orc_expf for the numpy brute-force checks in tests/. */
void orc_expf_array(const float *x, float *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = orc_expf(x[i]);
}
