// context.hpp -- the lumen_ex context (per-frame workspace and the state one
// render leaves for its backward pass) and the device view of a scene.
#pragma once

#include "common.hpp"
#include "lumen_ex.h"
#include "splat_math.hpp"

namespace lumen {

struct SceneDev {
  const float *px, *py, *pz, *sx, *sy, *sz, *rw, *rx, *ry, *rz, *op, *sh;
  uint32_t n;
};

__device__ __forceinline__ bool load_geometry(const SceneDev& s, const ViewParams& v, uint32_t i,
                                              SplatGeom& g) {
  return splat_geometry(v, s.px[i], s.py[i], s.pz[i], s.sx[i], s.sy[i], s.sz[i],
                        Quat{s.rw[i], s.rx[i], s.ry[i], s.rz[i]}, s.op[i], g);
}

// Per-view constants from the ABI camera and config (validated by the caller).
inline void fill_view(ViewParams& v, const LumenExCamera* cam, const LumenExRenderConfig* cfg) {
  memset(&v, 0, sizeof(v));
  v.width = (int32_t)cam->width;
  v.height = (int32_t)cam->height;
  v.tile_w = (int32_t)cfg->tile_width;
  v.tile_h = (int32_t)cfg->tile_height;
  v.tiles_x = (int32_t)((cam->width + cfg->tile_width - 1) / cfg->tile_width);
  v.tiles_y = (int32_t)((cam->height + cfg->tile_height - 1) / cfg->tile_height);
  v.fx = cam->focal_x;
  v.fy = cam->focal_y;
  v.cx = cam->principal_x;
  v.cy = cam->principal_y;
  v.near_plane = cam->near_plane;
  v.far_plane = cam->far_plane;
  v.lowpass = cfg->lowpass_filter;
  static const int kCoeffs[4] = {1, 4, 9, 16};
  v.sh_coeffs = kCoeffs[cfg->max_sh_degree];
  v.white_background = cfg->white_background ? 1 : 0;
  v.compute_depth = cfg->compute_depth ? 1 : 0;
  make_view(v, cam->position, cam->rotation);
}

inline SceneDev scene_dev(const LumenExScene* sc) {
  return SceneDev{sc->pos_x, sc->pos_y, sc->pos_z, sc->scale_x, sc->scale_y, sc->scale_z,
                  sc->rot_w, sc->rot_x, sc->rot_y, sc->rot_z, sc->opacity, sc->sh,
                  (uint32_t)sc->count};
}

}  // namespace lumen

struct LumenExContext {
  int device = 0;
  lumen::DevBuf packed, part, totals, bounds;
  lumen::DevBuf rec, vis2orig, radius;
  lumen::DevBuf keys[2], vals[2], keys_keep, vals_keep;
  lumen::DevBuf hist, hist_part;
  lumen::DevBuf ranges, t_final, n_eval, minmax;
  lumen::DevBuf grad_vis;  // backward: per visible splat dL/d(mean2d, inv cov, opacity, rgb, depth)
  void* pinned = nullptr;
  // metadata of the last render (for lumen_ex_copy_stage and the backward)
  uint64_t V = 0, M = 0, T = 0, P = 0, N = 0;
  uint32_t width = 0, height = 0, tile_w = 0, tile_h = 0;
  int final_buf = 0;
  bool kept = false, pixels = false, has_grad = false;
  // optional per-stage timing (lumen_ex_set_profiling)
  bool profile = false;
  bool events_ready = false;
  bool stage_ran[LUMEN_EX_TIME_STAGES] = {};
  hipEvent_t ev[2 * LUMEN_EX_TIME_STAGES] = {};
};


// Records the start (end = false) or end of a timed stage on the stream when
// profiling is enabled (lumen_ex_set_profiling).
inline LumenStatus mark(LumenExContext* c, int stage, bool end, hipStream_t st) {
  if (!c->profile) return LUMEN_SUCCESS;
  LUMEN_HIP(hipEventRecord(c->ev[2 * stage + (end ? 1 : 0)], st));
  if (end) c->stage_ran[stage] = true;
  return LUMEN_SUCCESS;
}
