// context.hpp -- the lumen_ex context (per-frame workspace and the state one
// frame leaves for its backward pass) and the device view of a scene.
#pragma once

#include <string.h>

#include <vector>

#include "common.hpp"
#include "lumen_ex.h"
#include "plan.hpp"
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

// Backward buffers: per tile instance and per visible splat, the gradients
// dL/d(mean x, mean y, inv00, inv01, inv11, opacity, r, g, b, depth).
constexpr int kGradVis = 10;
inline size_t partial_bytes(uint64_t m_cap) { return (m_cap ? m_cap : 1) * kGradVis * sizeof(float); }
inline size_t grad_vis_bytes(uint64_t v_cap) { return (v_cap ? v_cap : 1) * kGradVis * sizeof(float); }
// The per-splat sum cuts the instance slots into chunks of kSumChunk: per
// chunk, the splat holding its first slot and two partial sums.
constexpr uint32_t kSumChunk = 512;
inline size_t sum_seg_bytes(uint64_t m_cap) { return (size_t)div_up(m_cap ? m_cap : 1, kSumChunk) * 4; }
inline size_t sum_aux_bytes(uint64_t m_cap) {
  return (size_t)div_up(m_cap ? m_cap : 1, kSumChunk) * 2 * kGradVis * sizeof(float);
}

// lumen_ex_photometric_loss's partial sums for a width x height image: two
// floats per block of four 8x8 windows and per block of the pixels outside
// the windows (at most 2048 such blocks).
inline uint32_t loss_window_blocks(uint32_t width, uint32_t height) {
  const uint32_t nwin = (width / 8) * (height / 8);
  return div_up(nwin ? nwin : 1, 4);
}
inline uint32_t loss_rest_blocks(uint32_t width, uint32_t height) {
  const uint64_t nwin = (uint64_t)(width / 8) * (height / 8);
  const uint64_t rest = (uint64_t)width * height - nwin * 64u;
  const uint32_t br = div_up(rest ? rest : 1, 256);
  return br > 2048 ? 2048 : br;
}
inline size_t loss_part_bytes(uint32_t width, uint32_t height) {
  return (size_t)2 * (loss_window_blocks(width, height) + loss_rest_blocks(width, height)) *
         sizeof(float);
}

// Where each view's visible splats and keys sit in the frame's arrays.
struct ViewSpan {
  uint32_t vis_base, visible, key_base, instances;
};

}  // namespace lumen

struct GraphCache {
  bool valid = false;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t stream = nullptr;
  std::vector<uint8_t> key;  // bytes of everything the captured frame depends on
  uint64_t m_cap = 0;
};

struct LumenExContext {
  int device = 0;
  lumen::DevBuf packed, part, totals, plan, views, spans, segs, geo, block_bounds;
  lumen::DevBuf rec, vis2orig, radius;
  lumen::DevBuf vmeta;    // per visible splat: first key offset, tile rectangle (uint4)
  lumen::DevBuf keys[2], vals[2], keys_keep, vals_keep;
  lumen::DevBuf hist, hist_part;
  // single-view frames whose first sort pass runs in k_emit: per (digit,
  // count block) tile counts of k_emit's digit, their exclusive prefixes over
  // the blocks and the per-digit totals
  lumen::DevBuf etab, epre, etot;
  lumen::DevBuf ranges, t_final, n_eval, minmax;
  lumen::DevBuf tile_order;  // the compositors' tile order: heaviest key count class first
  lumen::DevBuf partial;   // backward: per tile instance (emission order) 10 partial gradients
  lumen::DevBuf grad_vis;  // backward: per visible splat dL/d(mean2d, inv cov, opacity, rgb, depth)
  lumen::DevBuf sum_seg, sum_aux;  // backward: per slot chunk, its first splat and partial sums
  lumen::DevBuf loss_part; // lumen_ex_photometric_loss: per-block partial sums
  // lumen_ex_densify_and_prune: per-Gaussian class, output count and offset;
  // stats, total and the scan's partial sums
  lumen::DevBuf dens_cls, dens_cnt, dens_off, dens_misc;
  void* pinned = nullptr;       // small host staging (plan, bounds)
  // mapped, coherent host words k_plan writes a plain frame's sizes to (the
  // host pointer and the device's address of them)
  uint64_t* ret_host = nullptr;
  uint64_t* ret_dev = nullptr;
  void* pinned_views = nullptr; // host staging of the per-view parameters
  size_t pinned_views_cap = 0;
  // the last frame (for lumen_ex_copy_stage and the backward)
  std::vector<lumen::ViewParams> host_views;
  uint64_t V = 0, M = 0, T = 0, P = 0, N = 0;
  uint32_t nviews = 0;
  uint32_t tile_w = 0, tile_h = 0;
  uint64_t v_cap = 0, m_cap = 0;  // capacities the frame's arrays were sized for
  // the bounds of the last lumen_ex_context_reserve: what a reserved frame
  // sizes its launches by (the buffers may hold more, from larger frames)
  bool reserved_set = false;
  uint64_t reserved_v = 0, reserved_m = 0;
  uint32_t key_bytes = 8;         // width of the last plain frame's sort keys
  uint32_t emit_passes = 0;       // of its passes, those k_emit ran (plan_frame)
  bool sizes_known = false;        // V and M were read back (not a graph replay)
  bool kept = false, pixels = false, has_grad = false;
  GraphCache graph;
  // optional per-stage timing (lumen_ex_set_profiling)
  bool profile = false;
  bool events_ready = false;
  bool stage_ran[LUMEN_EX_TIME_STAGES] = {};
  hipEvent_t ev[2 * LUMEN_EX_TIME_STAGES] = {};
  // the plain frame waits on this event (recorded after the totals' copy)
  hipEvent_t ev_totals = nullptr;
  // recorded after the per-view parameters' copy out of pinned_views; the
  // next upload waits on it before it overwrites the staging
  hipEvent_t ev_views = nullptr;
  bool ev_views_pending = false;
};

// Records the start (end = false) or end of a timed stage on the stream when
// profiling is enabled (lumen_ex_set_profiling).  Never inside a graph capture.
inline LumenStatus mark(LumenExContext* c, int stage, bool end, hipStream_t st) {
  if (!c->profile) return LUMEN_SUCCESS;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  LUMEN_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return LUMEN_SUCCESS;
  LUMEN_HIP(hipEventRecord(c->ev[2 * stage + (end ? 1 : 0)], st));
  if (end) c->stage_ran[stage] = true;
  return LUMEN_SUCCESS;
}
