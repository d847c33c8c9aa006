// legacy.hpp -- lumen_rasterize's legacy renderer mode on the GPU: the
// reference C library's own algorithm (c_api/src/lumen_c.c:380-575), bit
// for bit, selected with lumen_ex_rasterizer_set_mode.
#pragma once

#include "common.hpp"
#include "lumen.h"
#include "lumen_ex.h"

namespace lumen {

// Device buffers of the legacy mode (grow only; owned by a LumenRasterizer).
struct LegacyBuffers {
  DevBuf proj, flag, off, part, totals;   // per Gaussian: projection, visibility, scan
  DevBuf rec, rec_sorted, dkeys, dvals;   // per visible splat, and the depth sort
  DevBuf rect, cnt, koff, kpart;          // per visible splat in depth order: tiles
  DevBuf tkeys, tvals, ranges;            // (tile, depth rank) pairs and per-tile runs
  DevBuf rgb, depth, alpha, touched;      // the framebuffer's planes on the device
};

void legacy_release(LegacyBuffers& b);

// Renders `n` AoS Gaussians (device, 59 floats each) from `cam` with the
// legacy algorithm into fb (host planes, read and updated in place).
// stats: total, visible, culled, total tiles, touched tiles.
LumenStatus legacy_render(LumenExContext* ctx, LegacyBuffers& b, const float* aos, uint64_t n,
                          const LumenConfig& cfg, const LumenCamera& cam, LumenFrameBuffer* fb,
                          uint64_t stats[5]);

}  // namespace lumen
