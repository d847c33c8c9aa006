/*
 * lumen.h -- liblumen C ABI, MI355X build.
 *
 * Same 22 functions and 8 public types, with identical names, signatures,
 * field order and LP64 sizes, as the reference's c_api/include/lumen.h:29-159,
 * so that existing callers (c_api/test/test_c_api.c, the Python SDK's ctypes
 * loader python/lumen/client.py:47-65) bind to this library unchanged.
 * lumen_rasterize here runs on the GPU with one of two renderers, chosen per
 * rasterizer with lumen_ex_rasterizer_set_mode (lumen_ex.h):
 *   - by default the splat pipeline (anisotropic EWA footprints, SH colour
 *     to degree 3), whose pixels differ from the reference C library's: the
 *     C test's observable contract holds (c_api/test/test_c_api.c:127-172),
 *     the pixel values do not;
 *   - LUMEN_EX_RASTER_LEGACY: the reference C library's own isotropic,
 *     DC-only renderer (c_api/src/lumen_c.c:380-575), bit for bit.
 * See lumen_ex.h for the full-precision entry points.
 */
#ifndef LUMEN_H
#define LUMEN_H

#ifdef __cplusplus
extern "C" {
#endif

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#define LUMEN_VERSION_MAJOR 0
#define LUMEN_VERSION_MINOR 1
#define LUMEN_VERSION_PATCH 0
#define LUMEN_VERSION_STRING "0.1.0"

/* c_api/include/lumen.h:29-36 */
typedef enum {
    LUMEN_SUCCESS = 0,
    LUMEN_ERROR_INVALID_ARGUMENT = -1,
    LUMEN_ERROR_OUT_OF_MEMORY = -2,
    LUMEN_ERROR_UNINITIALIZED = -3,
    LUMEN_ERROR_IO_FAILURE = -4,
    LUMEN_ERROR_MATH_SINGULAR = -5
} LumenStatus;

/* c_api/include/lumen.h:39-47 (28 bytes) */
typedef struct {
    uint32_t image_width;
    uint32_t image_height;
    uint32_t tile_size;
    float near_plane;
    float far_plane;
    float early_termination_alpha;
    float gaussian_radius_sigma;
} LumenConfig;

/* c_api/include/lumen.h:50-54 */
typedef struct {
    float x;
    float y;
    float z;
} LumenVec3;

/* c_api/include/lumen.h:57-62 */
typedef struct {
    float w;
    float x;
    float y;
    float z;
} LumenQuat;

/* c_api/include/lumen.h:65-73 (52 bytes) */
typedef struct {
    LumenVec3 position;
    LumenVec3 target;
    LumenVec3 up;
    float fov_x_degrees;
    float fov_y_degrees;
    float near_plane;
    float far_plane;
} LumenCamera;

/* c_api/include/lumen.h:76-82 (236 bytes) */
typedef struct {
    LumenVec3 position;
    LumenVec3 scale;
    LumenQuat rotation;
    float opacity;
    float sh_coeffs[48];
} LumenGaussian;

/* c_api/include/lumen.h:85-91 (32 bytes) */
typedef struct {
    uint32_t width;
    uint32_t height;
    uint8_t* rgb_data;
    float* depth_data;
    float* alpha_data;
} LumenFrameBuffer;

/* c_api/include/lumen.h:94-102 (48 bytes) */
typedef struct {
    size_t total_splats;
    size_t visible_splats;
    size_t culled_splats;
    size_t total_tiles;
    size_t touched_tiles;
    float rasterize_time_ms;
    float frames_per_second;
} LumenStats;

typedef struct LumenCloud LumenCloud;
typedef struct LumenRasterizer LumenRasterizer;

/* c_api/include/lumen.h:112-118 */
const char* lumen_get_version(void);
int lumen_get_version_major(void);
int lumen_get_version_minor(void);
int lumen_get_version_patch(void);
LumenConfig lumen_default_config(void);

/* c_api/include/lumen.h:124-135 */
LumenCloud* lumen_cloud_create(size_t initial_capacity);
void lumen_cloud_destroy(LumenCloud* cloud);
LumenStatus lumen_cloud_reserve(LumenCloud* cloud, size_t new_capacity);
size_t lumen_cloud_get_count(const LumenCloud* cloud);
size_t lumen_cloud_get_capacity(const LumenCloud* cloud);
LumenStatus lumen_cloud_add_gaussian(LumenCloud* cloud, const LumenGaussian* gaussian);
LumenStatus lumen_cloud_get_gaussian(const LumenCloud* cloud, size_t index, LumenGaussian* out_gaussian);
void lumen_cloud_clear(LumenCloud* cloud);
LumenCloud* lumen_create_synthetic_cube(size_t points_per_axis, float cube_size);
LumenCloud* lumen_create_synthetic_sphere(size_t point_count, float radius);

/* c_api/include/lumen.h:141-143 */
LumenFrameBuffer* lumen_framebuffer_create(uint32_t width, uint32_t height);
void lumen_framebuffer_destroy(LumenFrameBuffer* fb);
void lumen_framebuffer_clear(LumenFrameBuffer* fb, uint8_t r, uint8_t g, uint8_t b);

/* c_api/include/lumen.h:149-159.  lumen_rasterize accumulates into the
 * framebuffer's existing contents; its renderer is the rasterizer's mode
 * (default: the splat pipeline, pixels differ from the reference C
 * library's; LUMEN_EX_RASTER_LEGACY: the reference's renderer bit for bit). */
LumenRasterizer* lumen_rasterizer_create(const LumenConfig* config);
void lumen_rasterizer_destroy(LumenRasterizer* rasterizer);
LumenStatus lumen_rasterizer_set_config(LumenRasterizer* rasterizer, const LumenConfig* config);
LumenStatus lumen_rasterize(LumenRasterizer* rasterizer, const LumenCloud* cloud,
                            const LumenCamera* camera, LumenFrameBuffer* out_framebuffer,
                            LumenStats* out_stats);

#ifdef __cplusplus
}
#endif

#endif /* LUMEN_H */
