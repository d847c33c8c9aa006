/*
 * lumen_ex.h -- liblumen extension ABI: the splat rasterization path of the
 * reference's Mojo core (src/raster/pipeline.mojo:48-231) on MI355X.
 *
 * Plain C, caller-owned device pointers, LumenStatus returns, the caller's
 * HIP stream (hipStream_t passed as void*, NULL = default stream).  No torch
 * types.  A context owns the per-frame workspace (grown on demand, never
 * shrunk), so steady-state calls allocate nothing.  One context serves one
 * stream at a time; distinct contexts are independent.
 *
 * Callers: the Python SDK re-pointed through ctypes (lumen_amd/_lib.py), and
 * the Mojo core through external_call (see INTEGRATION.md for both bindings).
 */
#ifndef LUMEN_EX_H
#define LUMEN_EX_H

#include "lumen.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LUMEN_EX_ABI_VERSION 2

/* Status values beyond LumenStatus (c_api/include/lumen.h:29-36). */
#define LUMEN_EX_ERROR_DEVICE ((LumenStatus)-100)   /* a HIP call failed */
#define LUMEN_EX_ERROR_CAPACITY ((LumenStatus)-101) /* a reserved capacity was exceeded */

typedef struct LumenExContext LumenExContext;

/* Gaussian scene as device planes; replaces GaussianSoA
 * (src/scene/gaussian.mojo:47-107): 11 planes of `count` floats and
 * sh = count*48 floats indexed g*48 + k*3 + c.  opacity is activated (not a
 * logit), rotation is (w, x, y, z) and need not be normalized. */
typedef struct {
    uint64_t count;
    const float* pos_x;
    const float* pos_y;
    const float* pos_z;
    const float* scale_x;
    const float* scale_y;
    const float* scale_z;
    const float* rot_w;
    const float* rot_x;
    const float* rot_y;
    const float* rot_z;
    const float* opacity;
    const float* sh;
} LumenExScene;

/* CameraConfig (src/core/config.mojo:17-42) + PinholeCamera pose
 * (src/scene/camera.mojo:37-40).  rotation is camera-to-world (w, x, y, z);
 * the camera looks along +Z with +Y down in view space
 * (src/scene/camera.mojo:69-79). */
typedef struct {
    uint32_t width;
    uint32_t height;
    float focal_x;
    float focal_y;
    float principal_x;
    float principal_y;
    float near_plane;
    float far_plane;
    float position[3];
    float rotation[4];
} LumenExCamera;

/* RenderConfig fields the reference reads (src/core/config.mojo:65-102). */
typedef struct {
    uint32_t tile_width;     /* 16 */
    uint32_t tile_height;    /* 16; tile_width*tile_height 64, 128, 192 or 256, tile_width dividing 64 */
    float lowpass_filter;    /* 0.3 */
    int32_t max_sh_degree;   /* 0..3 */
    int32_t white_background;
    int32_t compute_depth;
    uint32_t flags;          /* LUMEN_EX_* below */
    uint32_t batch_depth;    /* compositor LDS batch: 64, 128 or 256; 0 = 128 for tiles of a
                                multiple of 128 pixels (two pixels per lane), else 256.
                                The backward's batch is 128 for 128 or 256, else 64 */
} LumenExRenderConfig;

#define LUMEN_EX_KEEP_STAGES 1u   /* keep the unsorted keys/values for lumen_ex_copy_stage */
#define LUMEN_EX_SAVE_PIXELS 2u   /* keep final transmittance and evaluation counts per pixel */
#define LUMEN_EX_GRAD_ASSIGN 8u   /* backward: write the gradient planes (every Gaussian's entry,
                                     zero where it is not visible) instead of adding to them;
                                     viewspace_grad_2d / observation_count still accumulate */

typedef struct {
    uint64_t visible_splats;        /* RenderOutput.visible_splats (pipeline.mojo:159) */
    uint64_t total_splat_instances; /* RenderOutput.total_splat_instances (pipeline.mojo:160) */
    uint64_t tiles_x;
    uint64_t tiles_y;
    uint32_t sort_passes;           /* radix digit passes actually run */
    uint32_t key_bytes;             /* 4: 32-bit frame keys were sorted, 8: 64-bit keys */
    uint32_t emit_passes;           /* of those, passes run while the keys were emitted
                                       (0 or 1: the first digit, the depth key's low 9 bits,
                                       on single-view frames of at least 2048 tiles) */
    uint32_t reserved;
} LumenExRenderStats;

LumenExContext* lumen_ex_context_create(int device);
void lumen_ex_context_destroy(LumenExContext* ctx);
int lumen_ex_abi_version(void);

/* ---- workspace ----------------------------------------------------------
 * The reference allocates its buffers per frame (RenderOutput and the
 * pipeline's lists, src/raster/pipeline.mojo:21-45,141-157); a context keeps
 * them and grows them on demand instead.  These three calls let a caller size
 * that workspace up front so that no frame allocates.
 *
 * lumen_ex_workspace_size: device bytes a context holds after frames of
 * n_gaussians splats over these cameras and config with at most
 * max_instances tile instances per frame (RenderOutput.total_splat_instances,
 * summed over the views), with the backward and the photometric loss when
 * with_backward.  config->flags count: LUMEN_EX_SAVE_PIXELS and
 * LUMEN_EX_KEEP_STAGES add their buffers, LUMEN_EX_GRAPH reserves the graph's
 * key headroom.  Host-only: touches no device.
 *
 * lumen_ex_context_reserve: allocates that workspace in the context; frames
 * within the bounds then allocate nothing (device or pinned host).  Its
 * bounds (n_views * n_gaussians visible splats, max_instances keys) are also
 * the capacities of the LUMEN_EX_RESERVED frames that follow, even where
 * larger frames have grown the buffers beyond them.
 *
 * lumen_ex_context_workspace: device bytes the context holds now. */
LumenStatus lumen_ex_workspace_size(const LumenExCamera* cameras, uint32_t n_views,
                                    const LumenExRenderConfig* config, uint64_t n_gaussians,
                                    uint64_t max_instances, int with_backward,
                                    size_t* out_bytes);
LumenStatus lumen_ex_context_reserve(LumenExContext* ctx, const LumenExCamera* cameras,
                                     uint32_t n_views, const LumenExRenderConfig* config,
                                     uint64_t n_gaussians, uint64_t max_instances,
                                     int with_backward);
LumenStatus lumen_ex_context_workspace(LumenExContext* ctx, size_t* held_bytes);

/* rasterize_scene (src/raster/pipeline.mojo:48-231).  out_rgb: H*W*3 floats,
 * out_depth: H*W floats (device).  Every pixel is written: empty tiles get the
 * background (0, or 1 with white_background) and depth 0.  Blocks the host
 * once per call to learn the instance count. */
LumenStatus lumen_ex_render(LumenExContext* ctx, const LumenExScene* scene,
                            const LumenExCamera* camera, const LumenExRenderConfig* config,
                            float* out_rgb, float* out_depth, LumenExRenderStats* out_stats,
                            void* stream);

/* ---- many views of one scene in one frame ---------------------------------
 * Renders rasterize_scene for every camera with one launch per stage: view k
 * owns the output pixels [pix_base_k, pix_base_k + W_k*H_k) of out_rgb
 * (x3 floats) and out_depth, pix_base_k = sum of W*H over views < k.  Each
 * view's image, and its keys/values/ranges with the view's frame offsets
 * removed, equal a single-view lumen_ex_render of that camera.  All views
 * share the config (tile shape, SH degree, ...); resolutions may differ. */
#define LUMEN_EX_GRAPH 4u   /* capture the frame as a hipGraph on first use (needs a
                               non-default stream) and replay it while the scene,
                               cameras, config, outputs and stream are unchanged; no
                               host round trip per frame */
#define LUMEN_EX_RESERVED 16u /* enqueue the whole frame with the visible-splat and key
                                 capacities of the context's last
                                 lumen_ex_context_reserve (with none: all its buffers
                                 hold, as grown by earlier frames) and return at once:
                                 no host round trip, any cameras.  The sort's grids
                                 cover the key capacity, so reserve tightly.
                                 A frame that exceeds them renders background and
                                 lumen_ex_frame_status reports LUMEN_EX_ERROR_CAPACITY;
                                 on a context with no capacity yet the frame renders
                                 plainly (and sizes it).  The backward works after it. */

typedef struct {
    uint32_t vis_base;   /* first visible index of the view in the frame's arrays */
    uint32_t visible;    /* RenderOutput.visible_splats of the view */
    uint32_t key_base;   /* first key of the view in the frame's (unsorted) key array */
    uint32_t instances;  /* RenderOutput.total_splat_instances of the view */
} LumenExViewStats;

LumenStatus lumen_ex_render_batch(LumenExContext* ctx, const LumenExScene* scene,
                                  const LumenExCamera* cameras, uint32_t n_views,
                                  const LumenExRenderConfig* config, float* out_rgb,
                                  float* out_depth, LumenExViewStats* out_stats, void* stream);

/* Waits for the context's last frame on `stream` and copies its per-view
 * stats (out_stats may be NULL).  Returns LUMEN_EX_ERROR_CAPACITY when a graph
 * replay produced more tile instances than the capacity reserved when it was
 * captured (the frame's images are then background only). */
LumenStatus lumen_ex_frame_status(LumenExContext* ctx, LumenExViewStats* out_stats,
                                  uint32_t n_views, void* stream);

/* radix_sort_64bit (src/raster/radix_sort.mojo:39-87): stable ascending sort
 * of n 64-bit keys with 32-bit values, in place on device memory.  Digits
 * that are equal across all keys are skipped (a stable pass over them is the
 * identity).  Uses the context's frame buffers: the last frame's stages and
 * its backward are no longer available afterwards. */
LumenStatus lumen_ex_radix_sort(LumenExContext* ctx, uint64_t* keys, uint32_t* values,
                                uint64_t n, void* stream);

/* TileRangeIndex.build_from_sorted_keys (src/raster/binning.mojo:51-73):
 * ranges[2*t], ranges[2*t+1] = [start, end) of tile t in the sorted keys, or
 * (-1, -1) when the tile has no instance. */
LumenStatus lumen_ex_build_ranges(const uint64_t* sorted_keys, uint64_t n, uint32_t tiles,
                                  int32_t* ranges, void* stream);

/* Stage buffers of the last lumen_ex_render on this context, copied to host
 * memory (synchronous).  bytes must equal the stage's size. */
typedef enum {
    LUMEN_EX_STAGE_VIS_TO_ORIG = 0,   /* int32  [V]   visible -> original index */
    LUMEN_EX_STAGE_RECORDS = 1,       /* float  [V][12] cx cy i00 2*i01 | i11 op r g | b depth 0 0 */
    LUMEN_EX_STAGE_RADIUS = 2,        /* int32  [V]   3-sigma radius */
    LUMEN_EX_STAGE_KEYS_UNSORTED = 3, /* uint64 [M]   (needs LUMEN_EX_KEEP_STAGES) */
    LUMEN_EX_STAGE_VALS_UNSORTED = 4, /* uint32 [M]   (needs LUMEN_EX_KEEP_STAGES) */
    LUMEN_EX_STAGE_KEYS = 5,          /* uint64 [M]   sorted */
    LUMEN_EX_STAGE_VALS = 6,          /* uint32 [M]   sorted */
    LUMEN_EX_STAGE_RANGES = 7,        /* int32  [T][2] */
    LUMEN_EX_STAGE_T_FINAL = 8,       /* float  [H*W] (needs LUMEN_EX_SAVE_PIXELS) */
    LUMEN_EX_STAGE_N_EVAL = 9,        /* int32  [H*W] (needs LUMEN_EX_SAVE_PIXELS) */
    LUMEN_EX_STAGE_GRAD_VIS = 10      /* float  [V][10] of the last backward: dL/d mean x, mean y,
                                         inv00, inv01, inv11, opacity, r, g, b, depth */
} LumenExStage;

LumenStatus lumen_ex_copy_stage(LumenExContext* ctx, int stage, void* dst_host, size_t bytes);

/* Gradient planes in the layout of GaussianGradients
 * (src/optim/gradients.mojo:10-106), device pointers, ACCUMULATED into (the
 * caller zeroes them, as GaussianGradients.zero does), or written when the
 * backward's config has LUMEN_EX_GRAD_ASSIGN (then they need no zeroing).  viewspace_grad_2d and
 * observation_count follow accumulate_viewspace_grad (gradients.mojo:96-99):
 * every Gaussian visible in the view gets |dL/d mean2d| (pixels) added and
 * its count incremented.  Either of the two may be NULL. */
typedef struct {
    float* pos_x;
    float* pos_y;
    float* pos_z;
    float* scale_x;
    float* scale_y;
    float* scale_z;
    float* rot_w;
    float* rot_x;
    float* rot_y;
    float* rot_z;
    float* opacity;
    float* sh;                  /* count*48 */
    float* viewspace_grad_2d;   /* count */
    int32_t* observation_count; /* count */
} LumenExGradients;

/* Backward of the last lumen_ex_render on this context, which must have
 * been called with LUMEN_EX_SAVE_PIXELS and the same scene, camera and config.
 * grad_rgb: H*W*3 floats, grad_depth: H*W floats or NULL (device).  Gradients
 * are ADDED to the planes of grads (written with LUMEN_EX_GRAD_ASSIGN in
 * config->flags).  The compositor's reverse pass recomputes
 * every alpha exactly as the forward did, recovers transmittance from the
 * saved final value, and stores each (splat, tile) instance's partial
 * gradient in its own slot; a second kernel sums a splat's slots in a fixed
 * order and a third applies the EWA, SH and covariance adjoints per visible
 * Gaussian.
 * No atomics are involved for one view, so the result is the same on every
 * run and does not depend on what else runs on the device. */
LumenStatus lumen_ex_render_backward(LumenExContext* ctx, const LumenExScene* scene,
                                     const LumenExCamera* camera,
                                     const LumenExRenderConfig* config, const float* grad_rgb,
                                     const float* grad_depth, LumenExGradients* grads,
                                     void* stream);

/* Backward of the last lumen_ex_render_batch (same cameras and config, with
 * LUMEN_EX_SAVE_PIXELS); grad_rgb/grad_depth in the batch's output layout.
 * Works after a graph replay too.  One thread per Gaussian adds up its views'
 * contributions in view order, so this too is the same on every run. */
LumenStatus lumen_ex_render_batch_backward(LumenExContext* ctx, const LumenExScene* scene,
                                           const LumenExCamera* cameras, uint32_t n_views,
                                           const LumenExRenderConfig* config,
                                           const float* grad_rgb, const float* grad_depth,
                                           LumenExGradients* grads, void* stream);

/* The training step's photometric loss on device images (H*W*3 floats,
 * interleaved RGB): compute_l1_loss, compute_ssim_window (8x8
 * non-overlapping windows, channels pooled) and compute_combined_loss
 * (src/optim/loss.mojo:11-35, 65-133, 136-144).  out (device, 3 floats)
 * receives L1, SSIM and (1 - lambda) L1 + lambda (1 - SSIM); grad (device,
 * H*W*3, or NULL) receives d(loss)/d(pred) with the reference's L1
 * subgradient (+1/N where pred >= target, -1/N elsewhere).  Partial sums
 * are added in a fixed order (the same result on every run). */
LumenStatus lumen_ex_photometric_loss(LumenExContext* ctx, const float* pred, const float* target,
                                      uint32_t width, uint32_t height, float lambda_dssim,
                                      float* out, float* grad, void* stream);

/* lumen_ex_photometric_loss with no context and no allocation, ever: its
 * partial sums live in a fixed 2 MiB device array of the library (one per
 * GPU, held until the process ends), large enough for images up to
 * 8192 x 8192; larger images return LUMEN_ERROR_INVALID_ARGUMENT.  The
 * images, out and grad are on the current device; the result is bit for bit
 * lumen_ex_photometric_loss's.  It has no counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under hipGraph
 * capture, serving loops that forbid allocation).  Every call on one GPU
 * shares the array: calls that overlap on one GPU (different streams, or
 * different host threads) race, so serialize them. */
LumenStatus lumen_ex_photometric_loss_noalloc(const float* pred, const float* target,
                                              uint32_t width, uint32_t height,
                                              float lambda_dssim, float* out, float* grad,
                                              void* stream);

/* ---- the optimizer side of a training step --------------------------------
 * Mutable Gaussian planes in the GaussianSoA layout (the scene itself, or a
 * per-Gaussian state such as Adam's moments): 11 planes of count floats and
 * sh = count*48 floats. */
typedef struct {
    float* pos_x;
    float* pos_y;
    float* pos_z;
    float* scale_x;
    float* scale_y;
    float* scale_z;
    float* rot_w;
    float* rot_x;
    float* rot_y;
    float* rot_z;
    float* opacity;
    float* sh;
} LumenExPlanes;

/* TrainingConfig fields the optimizer and density control read
 * (src/core/config.mojo:113-166) and the Adam constants
 * (src/optim/optimizer.mojo:80-83).  Reference defaults in comments. */
typedef struct {
    int64_t iterations;              /* 30000 */
    float position_lr_init;          /* 0.00016 */
    float position_lr_final;         /* 0.0000016 */
    float feature_lr;                /* 0.0025 */
    float opacity_lr;                /* 0.05 */
    float scaling_lr;                /* 0.005 */
    float rotation_lr;               /* 0.001 */
    float beta1;                     /* 0.9 */
    float beta2;                     /* 0.999 */
    float epsilon;                   /* 1e-8 */
    float densify_grad_threshold;    /* 0.0002 */
    float prune_opacity_threshold;   /* 0.005 */
} LumenExTrainConfig;

#define LUMEN_EX_ADAM_ALL_PARAMS 1u  /* also scale, rotation and SH (see below) */

/* AdamOptimizer.step (src/optim/optimizer.mojo:112-170) for the step count
 * `step` (>= 1, the reference's step_count after its increment): pos x, y, z
 * with get_position_lr(step) (optimizer.mojo:103-110) and opacity with
 * opacity_lr, clamped to [0.01, 0.99], in place on device planes; m and v
 * are the first and second moments (zero before the first step).  Only the
 * pos and opacity planes of grads, m and v are read unless
 * LUMEN_EX_ADAM_ALL_PARAMS, which also updates scale (scaling_lr), rotation
 * (rotation_lr) and SH (feature_lr) the same way without a clamp (the
 * reference declares these rates but does not apply them).  The same
 * result, bit for bit, as the reference's scalar loop. */
LumenStatus lumen_ex_adam_step(const LumenExPlanes* params, const LumenExGradients* grads,
                               const LumenExPlanes* m, const LumenExPlanes* v, uint64_t count,
                               const LumenExTrainConfig* config, int64_t step, uint32_t flags,
                               void* stream);

/* reset_opacities (src/optim/density_control.mojo:101-109): every opacity
 * above max_value becomes max_value (device plane, in place). */
LumenStatus lumen_ex_reset_opacities(float* opacity, uint64_t count, float max_value,
                                     void* stream);

typedef struct {
    uint64_t cloned;  /* DensityControlStats (density_control.mojo:14-22) */
    uint64_t split;
    uint64_t pruned;
    uint64_t count;   /* Gaussians written to out */
} LumenExDensifyStats;

/* densify_and_prune (src/optim/density_control.mojo:25-98) from the scene
 * and its view-space gradient statistics (GaussianGradients'
 * viewspace_grad_2d and observation_count) into the caller's `out` planes,
 * which hold out_capacity Gaussians (2 * count always suffices).  Output in
 * the reference's order, rotations normalized as its Gaussian3D constructor
 * does.  moments_in / moments_out (both NULL, or each an array {m, v} of
 * full LumenExPlanes) carry Adam's state over: kept Gaussians and a clone's
 * original keep their moments, the new ones start at zero.  Blocks the
 * host once (the new count); returns LUMEN_EX_ERROR_CAPACITY with stats
 * filled and nothing written when the result exceeds out_capacity. */
LumenStatus lumen_ex_densify_and_prune(LumenExContext* ctx, const LumenExScene* scene,
                                       const float* viewspace_grad_2d,
                                       const int32_t* observation_count,
                                       const LumenExTrainConfig* config, float scene_extent,
                                       const LumenExPlanes* out, uint64_t out_capacity,
                                       const LumenExPlanes* moments_in,
                                       const LumenExPlanes* moments_out,
                                       LumenExDensifyStats* stats, void* stream);

/* Per-stage GPU timing with HIP events recorded on the render's stream around
 * each stage of lumen_ex_render (off by default).  lumen_ex_stage_times waits
 * for the last render's events and writes up to n durations in ms, in the
 * order of LumenExTimedStage. */
typedef enum {
    LUMEN_EX_TIME_COUNT = 0,     /* k_count + compaction scan */
    LUMEN_EX_TIME_EMIT = 1,      /* records + keys */
    LUMEN_EX_TIME_SORT = 2,      /* all radix passes */
    LUMEN_EX_TIME_RANGES = 3,    /* tile ranges */
    LUMEN_EX_TIME_COMPOSITE = 4, /* compositor */
    LUMEN_EX_TIME_BWD_COMPOSITE = 5,  /* compositor reverse pass (lumen_ex_render_backward) */
    LUMEN_EX_TIME_BWD_PREPROCESS = 6, /* EWA / SH / covariance adjoints */
    LUMEN_EX_TIME_STAGES = 7
} LumenExTimedStage;

LumenStatus lumen_ex_set_profiling(LumenExContext* ctx, int enable);
LumenStatus lumen_ex_stage_times(LumenExContext* ctx, float* ms_out, int n);

/* Message of the last failing lumen_ex_* call on this thread ("" if none). */
const char* lumen_ex_last_error(void);

/* ---- lumen_rasterize's renderer ------------------------------------------
 * lumen_rasterize (lumen.h) renders with one of two algorithms:
 *   LUMEN_EX_RASTER_PIPELINE (the default): the splat pipeline of
 *     lumen_ex_render (anisotropic EWA footprints, SH colour up to degree 3,
 *     16x16 front-to-back compositing), composited into the framebuffer's
 *     8-bit colour and float alpha/depth planes.  Its pixels differ from the
 *     reference C library's, whose renderer is a simpler model.
 *   LUMEN_EX_RASTER_LEGACY: the reference C library's own renderer
 *     (c_api/src/lumen_c.c:380-575) on the GPU, bit for bit: isotropic
 *     radius sigma * max(scale) * f / z, DC-only colour, splats in ascending
 *     depth (ties in cloud order), saturating 8-bit accumulation into the
 *     framebuffer.  Tile size and image sides are limited to 1..65535.
 * The mode belongs to the rasterizer handle and persists across calls.  The
 * reference has no such switch: it only has the legacy renderer. */
#define LUMEN_EX_RASTER_PIPELINE 0
#define LUMEN_EX_RASTER_LEGACY 1
LumenStatus lumen_ex_rasterizer_set_mode(LumenRasterizer* rasterizer, int mode);
int lumen_ex_rasterizer_get_mode(const LumenRasterizer* rasterizer);

#ifdef __cplusplus
}
#endif

#endif /* LUMEN_EX_H */
