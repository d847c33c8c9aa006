/*
 * lumen_oracle.h -- CPU oracle for the splat rasterization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar, single-threaded C restatement
 * of the reference's Mojo rasterizer (src/raster/pipeline.mojo:48-231 and the
 * math it calls).  It is the checker for the HIP implementation in lumen_amd/
 * and the CPU baseline that bench.py times.  Only tests/, the smoke check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may load it; the product
 * path (liblumen.so and the lumen_amd Python modules) never links, loads or calls it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math).  No FMA
 * contraction, correctly rounded '/' and sqrt, IEEE binary32 everywhere the
 * reference uses Float32 and int64 where it uses Int.
 */
#ifndef LUMEN_ORACLE_H
#define LUMEN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CameraConfig (src/core/config.mojo:17-42) + PinholeCamera pose
 * (src/scene/camera.mojo:37-40).  rot = (w, x, y, z) camera-to-world. */
typedef struct {
    int64_t width, height;
    float focal_x, focal_y, principal_x, principal_y;
    float near_plane, far_plane;
    float pos[3];
    float rot[4];
} orc_camera;

/* RenderConfig fields the reference reads (src/core/config.mojo:65-102,
 * src/raster/pipeline.mojo:60,104,138,184,221,228). */
typedef struct {
    int64_t tile_width, tile_height;
    float lowpass_filter;
    int64_t max_sh_degree;
    int32_t white_background;
    int32_t compute_depth;
    /* 0: exp from orc_expf (the Cody-Waite + degree-6 polynomial exp the HIP
     *    kernel also implements); 1: libm expf.  Mojo's std.math.exp is neither;
     *    both agree with it to ~1 ulp. */
    int32_t exp_mode;
} orc_config;

/* Scene in the GaussianSoA layout (src/scene/gaussian.mojo:47-107):
 * 11 planes of n floats, sh = n*48 floats with index g*48 + k*3 + c. */
typedef struct {
    int64_t n;
    const float *pos_x, *pos_y, *pos_z;
    const float *scale_x, *scale_y, *scale_z;
    const float *rot_w, *rot_x, *rot_y, *rot_z;
    const float *opacity;
    const float *sh;
} orc_scene;

/* Everything rasterize_scene computes, stage by stage. */
typedef struct {
    int64_t n, visible, instances, tiles, tiles_x, tiles_y, evaluations;
    int64_t width, height;
    /* per visible splat (visible index order) */
    int64_t *vis_to_orig;  /* visible -> original Gaussian index */
    float *proj;           /* V x 10: cx, cy, depth, inv00, inv01, inv11, r, g, b, opacity */
    int64_t *radius;       /* V */
    int64_t *rect;         /* V x 4: min_tx, min_ty, max_tx, max_ty */
    /* per instance */
    uint64_t *keys_unsorted;
    int64_t *vals_unsorted;
    uint64_t *keys;        /* sorted */
    int64_t *vals;         /* sorted */
    /* per tile: start, end ([-1,-1] when empty) */
    int64_t *ranges;
    /* per pixel */
    float *rgb;            /* H x W x 3 */
    float *depth;          /* H x W */
    float *t_final;        /* H x W transmittance after the last blend */
    int32_t *n_eval;       /* H x W alpha evaluations incl. the terminating one */
} orc_result;

/* ---- whole path ------------------------------------------------------- */
orc_result *orc_rasterize(const orc_scene *scene, const orc_camera *cam,
                          const orc_config *cfg);
void orc_result_free(orc_result *r);

/* ---- single stages (known-answer tests and sort/range parity) ---------- */
float orc_expf(float x);
void orc_expf_array(const float *x, float *y, int64_t n);
void orc_quat_normalized(const float q[4], float out[4]);
void orc_quat_to_matrix(const float q[4], float out[9]);
void orc_quat_inverse(const float q[4], float out[4]);
void orc_quat_mul(const float a[4], const float b[4], float out[4]);
void orc_quat_rotate_vec(const float q[4], const float v[3], float out[3]);
void orc_world_to_cam(const orc_camera *cam, const float p[3], float out[3]);
void orc_cam_to_screen(const orc_camera *cam, const float pc[3], float out[2]);
void orc_cov3d(const float scale[3], const float q[4], float out[6]);
void orc_jacobian(const float t[3], float fx, float fy, float out[6]);
void orc_ewa(const float cov3d[6], const float pc[3], const float view_rot[9],
             float fx, float fy, float lowpass, float out[3]);
int64_t orc_eigen(const float cov2d[3], float out_lambda[2]);
void orc_inverse2d(const float cov2d[3], float out[3]);
void orc_sh_basis(const float dir[3], float out[16]);
void orc_sh_color(const float *sh48, const float dir[3], int64_t degree, float out[3]);
int orc_depth_visible(float near_plane, float far_plane, float z);
int orc_opacity_visible(float op);
int orc_screen_box_visible(int64_t w, int64_t h, float cx, float cy, int64_t r);
int orc_tile_overlap(int64_t w, int64_t h, int64_t tw, int64_t th, float cx,
                     float cy, int64_t r, int64_t out[4]);
uint64_t orc_encode_key(uint32_t tile, float depth);
void orc_radix_sort(uint64_t *keys, int64_t *vals, int64_t n);
void orc_build_ranges(const uint64_t *keys, int64_t n, int64_t tiles, int64_t *ranges);
float orc_pixel_alpha(float px, float py, float cx, float cy, const float inv[3],
                      float opacity, int32_t exp_mode);
/* PixelAccumulator.blend over a list; state = {r, g, b, T, depth}. Returns
 * the index of the terminating splat or -1. */
int64_t orc_blend_sequence(const float *colors, const float *alphas,
                           const float *depths, int64_t n, float state[5]);

/* ---- FP64 forward and analytic backward (lumen_oracle_bwd.c) ----------- */
typedef struct {
    int64_t n;
    const double *planes; /* 11 x n, GaussianSoA plane order */
    const double *sh;     /* n x 48 */
} orc_scene64;

/* GaussianGradients (src/optim/gradients.mojo:10-106), accumulated into. */
typedef struct {
    double *planes;       /* 11 x n: pos xyz, scale xyz, rot wxyz, opacity */
    double *sh;           /* n x 48 */
    double *viewspace;    /* n: sum over views of |dL/d mean2d| (pixels) */
    int64_t *count;       /* n: views in which the Gaussian was visible */
} orc_grads64;

int orc64_forward(const orc_scene64 *sc, const orc_camera *cam, const orc_config *cfg,
                  double *rgb, double *depth, uint64_t *signature);
/* g_vis (optional, V x 10): per visible splat dL/d(mean x, mean y, inv00,
 * inv01, inv11, opacity, r, g, b, depth).  Returns V. */
int orc64_backward(const orc_scene64 *sc, const orc_camera *cam, const orc_config *cfg,
                   const double *g_rgb, const double *g_depth, orc_grads64 *out,
                   double *g_vis);

/* ---- losses (src/optim/loss.mojo) -------------------------------------- */
/* ---- optimizer and density control (lumen_oracle_optim.c) ------------- */

/* TrainingConfig fields these read (src/core/config.mojo:113-166) and the
 * Adam constants (src/optim/optimizer.mojo:80-83). */
typedef struct {
    int64_t iterations;
    float position_lr_init, position_lr_final, feature_lr, opacity_lr, scaling_lr, rotation_lr;
    float beta1, beta2, epsilon;
    float densify_grad_threshold, prune_opacity_threshold;
} orc_train_config;

typedef struct {
    float pos_lr, inv_b1, inv_b2;
} orc_adam_scalars;

void orc_adam_schedule(const orc_train_config *cfg, int64_t step, orc_adam_scalars *out);
void orc_adam_step(float *planes, float *sh, const float *gplanes, const float *gsh, float *mplanes,
                   float *msh, float *vplanes, float *vsh, int64_t n, const orc_train_config *c,
                   const orc_adam_scalars *s, int all_params);
void orc_reset_opacities(float *opacity, int64_t n, float max_val);
int64_t orc_densify_and_prune(const float *planes, const float *sh, const float *vs_grad,
                              const int32_t *vs_count, int64_t n, const orc_train_config *c,
                              float scene_extent, float *out_planes, float *out_sh, int64_t cap,
                              int64_t *origin, int64_t stats[3]);

float orc_l1_loss(const float *pred, const float *target, int64_t n, float *grad);
float orc_ssim(const float *pred, const float *target, int64_t w, int64_t h, int64_t win);
float orc_psnr(const float *pred, const float *target, int64_t n);
float orc_combined_loss(float l1, float ssim, float lambda_dssim);

/* ---- the legacy C renderer (lumen_oracle_legacy.c) --------------------- */
/* LumenCamera (c_api/include/lumen.h:65-73): position, target, up, fov x/y
 * in degrees, near, far (the renderer reads the config's planes instead). */
typedef struct {
    float pos[3], target[3], up[3];
    float fov_x_deg, fov_y_deg, near_plane, far_plane;
} orc_legacy_camera;
/* the LumenConfig fields lumen_rasterize reads (c_api/include/lumen.h:39-47) */
typedef struct {
    float near_plane, far_plane, early_termination_alpha, gaussian_radius_sigma;
    uint32_t tile_size;
} orc_legacy_config;
typedef struct {
    uint64_t total, visible, culled, total_tiles, touched;
} orc_legacy_stats;
/* lumen_rasterize (c_api/src/lumen_c.c:380-575) over n AoS LumenGaussians
 * (59 floats each) into a width x height framebuffer whose planes it reads
 * and updates in place.  Returns -1 for a zero tile size (the C library
 * divides by it), else 0. */
int orc_legacy_rasterize(const float *aos, int64_t n, const orc_legacy_camera *cam,
                         const orc_legacy_config *cfg, uint32_t width, uint32_t height,
                         uint8_t *rgb, float *depth, float *alpha, orc_legacy_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
