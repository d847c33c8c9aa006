/*
 * lumen_oracle_optim.c -- CPU oracle for the optimizer side of training.
 *
 * TEST INFRASTRUCTURE ONLY (see lumen_oracle.h).  Scalar restatements of
 *   AdamOptimizer.get_position_lr / step   src/optim/optimizer.mojo:103-170
 *   densify_and_prune                      src/optim/density_control.mojo:25-98
 *   reset_opacities                        src/optim/density_control.mojo:101-109
 * in the reference's Float32 expression order (no contraction).  The scene
 * and every per-Gaussian state are in the GaussianSoA plane layout: 11
 * planes of n floats (pos xyz, scale xyz, rot wxyz, opacity) and n*48 SH.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "lumen_oracle.h"

/* get_position_lr (optimizer.mojo:103-110) and the bias corrections of
 * step (optimizer.mojo:124-130); pow is the C library's powf. */
void orc_adam_schedule(const orc_train_config *cfg, int64_t step, orc_adam_scalars *out) {
    float progress = (float)step / (float)cfg->iterations;
    if (progress > 1.0f) progress = 1.0f;
    const float ratio = cfg->position_lr_final / cfg->position_lr_init;
    out->pos_lr = cfg->position_lr_init * powf(ratio, progress);
    const float t = (float)step;
    const float b1t = powf(cfg->beta1, t);
    const float b2t = powf(cfg->beta2, t);
    out->inv_b1 = 1.0f / (1.0f - b1t);
    out->inv_b2 = 1.0f / (1.0f - b2t);
}

/* One parameter: m, v and the update (optimizer.mojo:135-141). */
static float adam_one(float p, float g, float *m, float *v, float lr, const orc_train_config *c,
                      const orc_adam_scalars *s) {
    *m = c->beta1 * *m + (1.0f - c->beta1) * g;
    *v = c->beta2 * *v + (1.0f - c->beta2) * g * g;
    const float mh = *m * s->inv_b1;
    const float vh = *v * s->inv_b2;
    return p - lr * mh / (sqrtf(vh) + c->epsilon);
}

/* step (optimizer.mojo:112-170): pos x, y, z with the scheduled rate and
 * opacity with opacity_lr, clamped to [0.01, 0.99].  all_params also updates
 * scale (scaling_lr), rotation (rotation_lr) and SH (feature_lr) the same
 * way (no clamp), which the reference declares but does not apply. */
void orc_adam_step(float *planes, float *sh, const float *gplanes, const float *gsh, float *mplanes,
                   float *msh, float *vplanes, float *vsh, int64_t n, const orc_train_config *c,
                   const orc_adam_scalars *s, int all_params) {
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            const int64_t j = k * n + i;
            planes[j] = adam_one(planes[j], gplanes[j], &mplanes[j], &vplanes[j], s->pos_lr, c, s);
        }
        const int64_t jo = 10 * n + i;
        float op = adam_one(planes[jo], gplanes[jo], &mplanes[jo], &vplanes[jo], c->opacity_lr, c, s);
        if (op < 0.01f) op = 0.01f;
        else if (op > 0.99f) op = 0.99f;
        planes[jo] = op;
        if (!all_params) continue;
        for (int k = 3; k < 10; ++k) {
            const int64_t j = k * n + i;
            const float lr = k < 6 ? c->scaling_lr : c->rotation_lr;
            planes[j] = adam_one(planes[j], gplanes[j], &mplanes[j], &vplanes[j], lr, c, s);
        }
        for (int k = 0; k < 48; ++k) {
            const int64_t j = i * 48 + k;
            sh[j] = adam_one(sh[j], gsh[j], &msh[j], &vsh[j], c->feature_lr, c, s);
        }
    }
}

void orc_reset_opacities(float *opacity, int64_t n, float max_val) {
    for (int64_t i = 0; i < n; ++i)
        if (opacity[i] > max_val) opacity[i] = max_val;
}

/* Quatf.normalized (src/math/quaternion.mojo:52-57), which the Gaussian3D
 * constructor applies to every Gaussian densify_and_prune adds. */
static void quat_normalized(const float q[4], float out[4]) {
    const float ns = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    const float nn = sqrtf(ns);
    if (nn > 0.000001f) {
        const float inv = 1.0f / nn;
        out[0] = q[0] * inv; out[1] = q[1] * inv; out[2] = q[2] * inv; out[3] = q[3] * inv;
    } else {
        out[0] = 1.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 0.0f;
    }
}

static void put(float *op, float *osh, int64_t cap, int64_t j, const float pos[3], const float sc[3],
                const float rot[4], float opacity, const float *sh48) {
    for (int k = 0; k < 3; ++k) op[k * cap + j] = pos[k];
    for (int k = 0; k < 3; ++k) op[(3 + k) * cap + j] = sc[k];
    for (int k = 0; k < 4; ++k) op[(6 + k) * cap + j] = rot[k];
    op[10 * cap + j] = opacity;
    memcpy(osh + j * 48, sh48, 48 * sizeof(float));
}

/* densify_and_prune (density_control.mojo:25-98).  Output planes have a
 * plane stride of `cap`; origin[j] = source index * 2 + (1 if the output is
 * a new Gaussian: the clone's offset copy or a split child).  Returns the
 * new count, or -1 when it exceeds cap (nothing is then written past cap). */
int64_t orc_densify_and_prune(const float *planes, const float *sh, const float *vs_grad,
                              const int32_t *vs_count, int64_t n, const orc_train_config *c,
                              float scene_extent, float *out_planes, float *out_sh, int64_t cap,
                              int64_t *origin, int64_t stats[3]) {
    stats[0] = stats[1] = stats[2] = 0;
    const float scale_threshold = scene_extent * 0.01f;
    int64_t j = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float op = planes[10 * n + i];
        if (op < c->prune_opacity_threshold) {
            stats[2] += 1;
            continue;
        }
        const float pos[3] = {planes[i], planes[n + i], planes[2 * n + i]};
        const float sc[3] = {planes[3 * n + i], planes[4 * n + i], planes[5 * n + i]};
        const float q[4] = {planes[6 * n + i], planes[7 * n + i], planes[8 * n + i], planes[9 * n + i]};
        float max_scale = sc[0];
        if (sc[1] > max_scale) max_scale = sc[1];
        if (sc[2] > max_scale) max_scale = sc[2];
        if (max_scale > scene_extent * 0.5f) {
            stats[2] += 1;
            continue;
        }
        float avg = 0.0f;  /* get_average_viewspace_grad (gradients.mojo:101-106) */
        if (vs_count[i] > 0) avg = vs_grad[i] / (float)vs_count[i];
        float rot[4];
        quat_normalized(q, rot);
        const float *s48 = sh + i * 48;
        if (avg > c->densify_grad_threshold) {
            if (j + 2 > cap) return -1;
            if (max_scale <= scale_threshold) {
                stats[0] += 1;
                const float off[3] = {pos[0] + 0.01f * sc[0], pos[1] + 0.01f * sc[1],
                                      pos[2] + 0.01f * sc[2]};
                put(out_planes, out_sh, cap, j, pos, sc, rot, op, s48);
                origin[j++] = 2 * i;
                put(out_planes, out_sh, cap, j, off, sc, rot, op, s48);
                origin[j++] = 2 * i + 1;
            } else {
                stats[1] += 1;
                const float ns[3] = {sc[0] / 1.6f, sc[1] / 1.6f, sc[2] / 1.6f};
                const float p1[3] = {pos[0] + ns[0] * 0.5f, pos[1] + ns[1] * 0.5f,
                                     pos[2] + ns[2] * 0.5f};
                const float p2[3] = {pos[0] - ns[0] * 0.5f, pos[1] - ns[1] * 0.5f,
                                     pos[2] - ns[2] * 0.5f};
                put(out_planes, out_sh, cap, j, p1, ns, rot, op, s48);
                origin[j++] = 2 * i + 1;
                put(out_planes, out_sh, cap, j, p2, ns, rot, op, s48);
                origin[j++] = 2 * i + 1;
            }
        } else {
            if (j + 1 > cap) return -1;
            put(out_planes, out_sh, cap, j, pos, sc, rot, op, s48);
            origin[j++] = 2 * i;
        }
    }
    return j;
}
