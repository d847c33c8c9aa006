/*
 * lumen_oracle_legacy.c -- CPU oracle of the legacy C renderer.
 *
 * TEST INFRASTRUCTURE ONLY (see lumen_oracle.h).  A scalar restatement of
 * lumen_rasterize in the reference's C library (c_api/src/lumen_c.c:380-575),
 * the isotropic, DC-only renderer behind lumen.h:
 *   - a look-at view matrix with (right, up, -forward) rows (:350-369),
 *     focal length H / (2 tan(fov_y / 2)) for both axes (:393-395);
 *   - per Gaussian: view-space position, culling on the rasterizer config's
 *     near/far planes (exclusive) and on the screen box, radius
 *     sigma * max(scale) * f / z clamped to >= 1, variance (radius /
 *     sigma)^2 clamped to >= 0.5, colour 0.5 + C0 * sh[0..2] clamped to
 *     [0, 1] (:407-468);
 *   - the visible splats in ascending depth (the C library's qsort, a stable
 *     merge sort in glibc; restated here as a stable merge sort) (:470-472);
 *   - per tile, every splat whose box meets the tile, in depth order, and per
 *     pixel of the tile the saturating 8-bit accumulation into the
 *     framebuffer's existing contents: early termination on its alpha plane,
 *     the radius test, alpha = opacity exp(-(dx^2 + dy^2) / (2 var)), the
 *     1/255 skip, colour += alpha (1 - a) c clamped at 1 and truncated to
 *     8 bits, depth taken while a < 0.5, a += alpha (1 - a) (:474-550);
 *   - the stats (:559-567).
 * Same float expression trees as the C source, -ffp-contract=off, libm's
 * expf and tanf.  Validated bit for bit against the reference's own C
 * library compiled from its source (oracle/_ref, tests/test_legacy_oracle.py).
 */
#include "lumen_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float sx, sy, depth, radius, inv_var, cr, cg, cb, opacity;
} legacy_splat;

static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

/* length below 1e-6: returned unchanged (lumen_c.c:342-352) */
static void normalize3(float v[3]) {
    float len = sqrtf(dot3(v, v));
    if (len < 1e-6f) return;
    float inv = 1.0f / len;
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

/* rows: right, up, -forward; translations -r.eye, -u.eye, f.eye */
static void view_matrix(const orc_legacy_camera *cam, float m[12]) {
    float f[3] = {cam->target[0] - cam->pos[0], cam->target[1] - cam->pos[1],
                  cam->target[2] - cam->pos[2]};
    normalize3(f);
    float r[3], u[3];
    cross3(f, cam->up, r);
    normalize3(r);
    cross3(r, f, u);
    m[0] = r[0]; m[1] = r[1]; m[2] = r[2]; m[3] = -dot3(r, cam->pos);
    m[4] = u[0]; m[5] = u[1]; m[6] = u[2]; m[7] = -dot3(u, cam->pos);
    m[8] = -f[0]; m[9] = -f[1]; m[10] = -f[2]; m[11] = dot3(f, cam->pos);
}

/* stable merge sort of idx[0..n) by splats[idx].depth ascending */
static void sort_by_depth(const legacy_splat *s, int64_t *idx, int64_t *tmp, int64_t n) {
    for (int64_t width = 1; width < n; width *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * width) {
            int64_t mid = lo + width < n ? lo + width : n;
            int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            int64_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                /* the right run's element goes first only if strictly less */
                if (s[idx[b]].depth < s[idx[a]].depth) tmp[o++] = idx[b++];
                else tmp[o++] = idx[a++];
            }
            while (a < mid) tmp[o++] = idx[a++];
            while (b < hi) tmp[o++] = idx[b++];
        }
        memcpy(idx, tmp, (size_t)n * sizeof(int64_t));
    }
}

int orc_legacy_rasterize(const float *aos, int64_t n, const orc_legacy_camera *cam,
                         const orc_legacy_config *cfg, uint32_t width, uint32_t height,
                         uint8_t *rgb, float *depth, float *alpha, orc_legacy_stats *stats) {
    if (cfg->tile_size == 0) return -1;
    const float fov_y = cam->fov_y_deg * (float)M_PI / 180.0f;
    const float focal = (float)height / (2.0f * tanf(fov_y * 0.5f));
    float m[12];
    view_matrix(cam, m);
    legacy_splat *sp = (legacy_splat *)malloc((size_t)(n > 0 ? n : 1) * sizeof(legacy_splat));
    int64_t visible = 0, culled = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float *g = aos + i * 59;  /* position 0-2, scale 3-5, rotation 6-9, opacity 10, sh 11- */
        const float x = g[0], y = g[1], z = g[2];
        const float vx = m[0] * x + m[1] * y + m[2] * z + m[3];
        const float vy = m[4] * x + m[5] * y + m[6] * z + m[7];
        const float vz = -(m[8] * x + m[9] * y + m[10] * z + m[11]);
        if (vz <= cfg->near_plane || vz >= cfg->far_plane) { ++culled; continue; }
        const float sx = (vx / vz) * focal + (float)width * 0.5f;
        const float sy = (vy / vz) * focal + (float)height * 0.5f;
        float smax = g[3];
        if (g[4] > smax) smax = g[4];
        if (g[5] > smax) smax = g[5];
        float radius = (smax * focal / vz) * cfg->gaussian_radius_sigma;
        if (radius < 1.0f) radius = 1.0f;
        if (sx + radius < 0.0f || sx - radius >= (float)width || sy + radius < 0.0f ||
            sy - radius >= (float)height) { ++culled; continue; }
        const float sig = radius / cfg->gaussian_radius_sigma;
        float var = sig * sig;
        if (var < 0.5f) var = 0.5f;
        const float c0 = 0.2820947917f;
        float c[3];
        for (int k = 0; k < 3; ++k) {
            c[k] = 0.5f + c0 * g[11 + k];
            if (c[k] < 0.0f) c[k] = 0.0f;
            if (c[k] > 1.0f) c[k] = 1.0f;
        }
        legacy_splat s = {sx, sy, vz, radius, 1.0f / var, c[0], c[1], c[2], g[10]};
        sp[visible++] = s;
    }
    int64_t *order = (int64_t *)malloc((size_t)(visible > 0 ? visible : 1) * sizeof(int64_t));
    int64_t *tmp = (int64_t *)malloc((size_t)(visible > 0 ? visible : 1) * sizeof(int64_t));
    for (int64_t k = 0; k < visible; ++k) order[k] = k;
    sort_by_depth(sp, order, tmp, visible);

    const uint32_t ts = cfg->tile_size;
    const uint32_t tiles_x = (width + ts - 1) / ts, tiles_y = (height + ts - 1) / ts;
    uint64_t touched = 0;
    for (uint32_t ty = 0; ty < tiles_y; ++ty) {
        const uint32_t y0 = ty * ts, y1 = y0 + ts < height ? y0 + ts : height;
        for (uint32_t tx = 0; tx < tiles_x; ++tx) {
            const uint32_t x0 = tx * ts, x1 = x0 + ts < width ? x0 + ts : width;
            int hit = 0;
            for (int64_t k = 0; k < visible; ++k) {
                const legacy_splat *p = &sp[order[k]];
                if (p->sx + p->radius < (float)x0 || p->sx - p->radius >= (float)x1 ||
                    p->sy + p->radius < (float)y0 || p->sy - p->radius >= (float)y1)
                    continue;
                hit = 1;
                for (uint32_t py = y0; py < y1; ++py) {
                    const float dy = (float)py - p->sy;
                    for (uint32_t px = x0; px < x1; ++px) {
                        const size_t pi = (size_t)py * width + px;
                        const float a = alpha[pi];
                        if (a >= cfg->early_termination_alpha) continue;
                        const float dx = (float)px - p->sx;
                        if (dx * dx + dy * dy > p->radius * p->radius) continue;
                        float power = -0.5f * (p->inv_var * dx * dx + p->inv_var * dy * dy);
                        if (power > 0.0f) power = 0.0f;
                        const float w_alpha = p->opacity * expf(power);
                        if (w_alpha < 1.0f / 255.0f) continue;
                        const float wgt = w_alpha * (1.0f - a);
                        const float col[3] = {p->cr, p->cg, p->cb};
                        for (int ch = 0; ch < 3; ++ch) {
                            float v = (float)rgb[pi * 3 + ch] / 255.0f + wgt * col[ch];
                            if (v > 1.0f) v = 1.0f;
                            rgb[pi * 3 + ch] = (uint8_t)(v * 255.0f);
                        }
                        if (a < 0.5f && p->depth < depth[pi]) depth[pi] = p->depth;
                        alpha[pi] = a + wgt;
                    }
                }
            }
            touched += (uint64_t)hit;
        }
    }
    if (stats) {
        stats->total = (uint64_t)n;
        stats->visible = (uint64_t)visible;
        stats->culled = (uint64_t)culled;
        stats->total_tiles = (uint64_t)tiles_x * tiles_y;
        stats->touched = touched;
    }
    free(sp);
    free(order);
    free(tmp);
    return 0;
}
