// liblumen.hip -- forward splat rasterization on MI355X (gfx950) behind the
// lumen_ex.h C ABI.  A frame renders one scene from one or more views; each
// view is exactly rasterize_scene (src/raster/pipeline.mojo:48-231):
//
//   k_count      culling + covariance + EWA + 3-sigma radius + tile rectangle
//                per (view, Gaussian), one thread per Gaussian looping over
//                the views; writes (tile count << 32 | visible) for the
//                compaction scan (view-major)       pipeline.mojo:84-130
//   scan         visible index and key offset per (view, Gaussian); stable,
//                so each view's visible order is the reference's append
//                order                               :141-142
//   k_plan       sort digit plan, key count, per-view spans and sort
//                segments, on the device
//   k_emit       the same geometry (bit-identical), SH colour, inverse
//                covariance, the 48-byte splat record at its visible index,
//                and the keys: (tile within the view, depth key), 32-bit
//                frame keys when the frame's ranges fit, else 64-bit  :132-157
//   sort         stable LSD radix sort over the digits that vary, each
//                view's keys a segment of its own                   :166
//   k_ranges     per-tile [start, end)                               :169
//   k_composite  one workgroup per tile, one lane per pixel, records staged
//                through LDS in batches, per-pixel front-to-back blend with
//                T <= 1e-4 termination and a block-wide exit         :172-229
//
// A view's tiles are frame tiles [tile_base, tile_base + tiles); with one
// view the frame tile id is the reference's tile id, so keys (decoded),
// values and ranges are the reference's.  The plain path blocks the host once per frame
// (to size the key buffers); the graph path (LUMEN_EX_GRAPH) captures the
// whole frame once with reserved capacities and replays it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdio.h>
#include <string.h>

#include <string>

#include "common.hpp"
#include "context.hpp"
#include "lumen_ex.h"
#include "scan.hpp"
#include "sort.hpp"
#include "splat_math.hpp"

namespace lumen {

static thread_local std::string t_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
}
const char* last_error() { return t_last_error.c_str(); }
void clear_error() { t_last_error.clear(); }

__global__ void k_init_frame(uint32_t* __restrict__ bounds, SortPlan* __restrict__ plan) {
  bounds[0] = 0xFFFFFFFFu;
  bounds[1] = 0u;
  bounds[2] = 0xFFFFFFFFu;
  bounds[3] = 0u;
  *reinterpret_cast<uint64_t*>(bounds + 4) = 0ull;  // exact instance total
  plan->passes = 0;
  plan->first = 0;
  plan->m = 0;
  plan->v = 0;
  plan->overflow = 0;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, d, 64));
  return x;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, d, 64));
  return x;
}

// Geometry of a visible (view, Gaussian) kept by k_count for k_emit (32 B):
// screen centre, depth, EWA covariance, tile rectangle (16-bit fields).
struct GeoCache {
  float4 a;  // sx, sy, depth, c00
  float4 b;  // c01, c11, rect min (tx | ty << 16), rect max (tx | ty << 16) as bits
};

// Pass 1: visibility and tile count per (view, Gaussian), the geometry of
// the visible ones, and per-block bounds of depth keys and view-local tile
// ids (which decide the sort digits; k_plan folds them).  The packed word's
// count half is scanned modulo 2^32, so each block also sums its counts
// exactly in 64 bits (block_inst): the frame's exact instance total, against
// which k_plan and the host reject frames the 32-bit offsets cannot address.
// INL: the views come in the kernel's arguments (iv) instead of device
// memory, and block 0 writes them to views_out for the frame's later kernels
// (a frame of at most kInlineViews views needs no copy of its own).
// With bsum, each block also writes the sum of its packed words (single-view
// frames: the block's Gaussians are one run of the scanned array) to
// bsum[Gaussian block], the partial sums the scan's downsweep starts from.
// With etab (single-view frames, frame_fusable), each block also writes the
// histogram of k_emit's digit (depth key & 511) over its Gaussians, weighted
// by their tile counts: etab[digit * blocks + block], the rows k_emit's first
// sort pass needs.  Such frames count in blocks of kFusedBlock Gaussians (a
// half of the table rows 256 would give: fewer partial-line writes here and
// shorter column scans in k_plan).  Grid block i takes Gaussian
// block xcd_block(i) (k_emit uses the same numbering), so the neighbouring
// blocks that share a line of the table run on one XCD.
constexpr int kInlineViews = 4;
constexpr int kFusedBlock = 512;
struct InlineViews {
  ViewParams v[kInlineViews];
};
static_assert(sizeof(ViewParams) % 4 == 0 && kInlineViews * sizeof(ViewParams) / 4 <= 256,
              "block 0 copies the views a word per thread");
template <bool INL, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_count(SceneDev s, const ViewParams* __restrict__ views,
                                               int nviews, uint64_t* __restrict__ packed,
                                               GeoCache* __restrict__ geo,
                                               uint4* __restrict__ block_bounds,
                                               uint64_t* __restrict__ block_inst, InlineViews iv,
                                               ViewParams* __restrict__ views_out,
                                               uint32_t* __restrict__ etab,
                                               uint64_t* __restrict__ bsum) {
  constexpr int NW = BLOCK / 64;
  __shared__ uint32_t sb[NW][4];
  __shared__ uint64_t sinst[NW], spk[NW];
  __shared__ uint32_t eh[kDigitBins];
  const uint32_t gb = xcd_block(blockIdx.x, gridDim.x);  // the block's Gaussians
  const uint32_t i = gb * (uint32_t)BLOCK + threadIdx.x;
  if (INL && blockIdx.x == 0 && threadIdx.x < nviews * (int)(sizeof(ViewParams) / 4))
    reinterpret_cast<uint32_t*>(views_out)[threadIdx.x] =
        reinterpret_cast<const uint32_t*>(iv.v)[threadIdx.x];
  if (etab) {  // uniform: a kernel argument
    for (int e = threadIdx.x; e < kDigitBins; e += BLOCK) eh[e] = 0u;
    __syncthreads();
  }
  uint32_t dmin = 0xFFFFFFFFu, dmax = 0u, tmin = 0xFFFFFFFFu, tmax = 0u;
  uint64_t inst = 0ull, pk = 0ull;
  if (i < s.n) {
    const float px = s.px[i], py = s.py[i], pz = s.pz[i];
    const float s0 = s.sx[i], s1 = s.sy[i], s2 = s.sz[i];
    const Quat q{s.rw[i], s.rx[i], s.ry[i], s.rz[i]};
    const float op = s.op[i];
    for (int vw = 0; vw < nviews; ++vw) {
      const ViewParams& v = INL ? iv.v[vw] : views[vw];
      SplatGeom g;
      uint64_t val = 0;
      const size_t idx = (size_t)vw * s.n + i;
      if (splat_geometry(v, px, py, pz, s0, s1, s2, q, op, g)) {
        const uint32_t cnt =
            (uint32_t)(g.max_tx - g.min_tx + 1) * (uint32_t)(g.max_ty - g.min_ty + 1);
        val = ((uint64_t)cnt << 32) | 1ull;
        inst += cnt;
        const uint32_t dk = depth_key(g.depth);
        if (etab && vw == 0) atomicAdd(&eh[dk & kEmitDigitMask], cnt);
        dmin = min(dmin, dk);
        dmax = max(dmax, dk);
        tmin = min(tmin, (uint32_t)(g.min_ty * v.tiles_x + g.min_tx));  // view-local tiles
        tmax = max(tmax, (uint32_t)(g.max_ty * v.tiles_x + g.max_tx));
        GeoCache gc;
        gc.a = make_float4(g.sx, g.sy, g.depth, g.c00);
        gc.b = make_float4(g.c01, g.c11,
                           __uint_as_float((uint32_t)g.min_tx | ((uint32_t)g.min_ty << 16)),
                           __uint_as_float((uint32_t)g.max_tx | ((uint32_t)g.max_ty << 16)));
        geo[idx] = gc;
      }
      packed[idx] = val;
      pk += val;
    }
  }
  dmin = wave_min_u32_all(dmin);
  dmax = wave_max_u32_all(dmax);
  tmin = wave_min_u32_all(tmin);
  tmax = wave_max_u32_all(tmax);
  inst = wave_sum(inst);
  if (bsum) pk = wave_sum(pk);  // uniform: a kernel argument
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sb[w][0] = dmin;
    sb[w][1] = dmax;
    sb[w][2] = tmin;
    sb[w][3] = tmax;
    sinst[w] = inst;
    spk[w] = pk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t bi = sinst[0], bp = spk[0];
    for (int k = 1; k < NW; ++k) {
      dmin = min(dmin, sb[k][0]);
      dmax = max(dmax, sb[k][1]);
      tmin = min(tmin, sb[k][2]);
      tmax = max(tmax, sb[k][3]);
      bi += sinst[k];
      bp += spk[k];
    }
    block_bounds[blockIdx.x] = make_uint4(dmin, dmax, tmin, tmax);
    block_inst[blockIdx.x] = bi;
    if (bsum) bsum[gb] = bp;
  }
  if (etab) {  // the __syncthreads above ordered the LDS atomics
    for (int e = threadIdx.x; e < kDigitBins; e += BLOCK)
      etab[(size_t)e * gridDim.x + gb] = eh[e];
  }
}

// Sort plan, clamped sizes and per-view spans from the scan and the bounds.
// Block 0 first folds k_count's per-block key bounds into bounds[0..3] and
// their exact instance sums into the 64-bit total at bounds + 4.
// A frame whose exact instance total exceeds the key capacity, or the
// 2^31 - 1 that int32 tile ranges address (which also covers a count half
// that wrapped in the scan), is flagged as overflowed: k_emit then writes
// nothing and the frame renders background.  With host_ret (mapped, coherent
// host memory) block 0 also writes the scan total, the four bounds and the
// exact instance total there, the 32 bytes a plain frame sizes itself by.
// With etab (k_count wrote k_emit's digit histograms, one row per count
// block), plan_frame may let k_emit run the first pass, and blocks 1 to 512
// each scan one digit's column of the table over the count blocks into epre
// (exclusive prefixes; out of place, so a frame that enqueues k_plan again
// after growing its buffers scans the same counts) and its total into etot.
constexpr int kPlanScanItems = 4;
__global__ __launch_bounds__(1024) void k_plan(
    const uint64_t* __restrict__ offsets, const uint64_t* __restrict__ totals,
    const uint4* __restrict__ block_bounds, const uint64_t* __restrict__ block_inst,
    uint32_t nblocks, uint32_t* __restrict__ bounds, uint32_t n, int nviews, uint64_t m_cap,
    uint64_t v_cap, SortPlan* __restrict__ plan, ViewSpan* __restrict__ spans,
    SortSeg* __restrict__ segs, int2* __restrict__ ranges, uint32_t tiles, int force_items,
    const uint32_t* __restrict__ etab, uint32_t* __restrict__ epre, uint32_t* __restrict__ etot,
    uint64_t* __restrict__ host_ret) {
  // every block: empty tile ranges (-1, -1) before k_ranges fills the used ones
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tiles; t += gridDim.x * blockDim.x)
    ranges[t] = make_int2(-1, -1);
  const bool fuse = etab != nullptr;
  if (fuse && blockIdx.x >= 1 && blockIdx.x <= (uint32_t)kDigitBins) {
    __shared__ uint32_t s_w[16];
    const uint32_t d = blockIdx.x - 1u;
    const uint32_t* __restrict__ src = etab + (size_t)d * nblocks;
    uint32_t* __restrict__ dst = epre + (size_t)d * nblocks;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += 1024u * kPlanScanItems) {  // uniform
      uint32_t v[kPlanScanItems], sum = 0;
      const uint32_t j0 = b0 + threadIdx.x * kPlanScanItems;
#pragma unroll
      for (int j = 0; j < kPlanScanItems; ++j) {
        v[j] = j0 + j < nblocks ? src[j0 + j] : 0u;
        sum += v[j];
      }
      uint32_t tot;
      uint32_t ex = block_exclusive_scan<uint32_t, 1024>(sum, s_w, tot) + carry;
#pragma unroll
      for (int j = 0; j < kPlanScanItems; ++j) {
        if (j0 + j < nblocks) dst[j0 + j] = ex;
        ex += v[j];
      }
      carry += tot;
    }
    if (threadIdx.x == 0) etot[d] = carry;
    return;
  }
  if (blockIdx.x != 0) return;
  __shared__ uint32_t sb[16][4];
  __shared__ uint64_t sinst[16];
  __shared__ uint32_t s_bounds[4];
  __shared__ uint64_t s_exact;
  // the scan's total and this thread's first view offsets, loaded with the
  // fold's loads
  const uint64_t tot = *totals;
  const int vw0 = (int)threadIdx.x;
  const uint64_t off_a0 = vw0 < nviews ? offsets[(size_t)vw0 * n] : 0ull;
  const uint64_t off_b0 = vw0 + 1 < nviews ? offsets[(size_t)(vw0 + 1) * n] : 0ull;
  {
    uint32_t dmin = 0xFFFFFFFFu, dmax = 0u, tmin = 0xFFFFFFFFu, tmax = 0u;
    uint64_t inst = 0ull;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += 4096u) {
      uint4 x[4];
      uint64_t y[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // four loads of each kind in flight
        const uint32_t b = b0 + (uint32_t)j * 1024u + threadIdx.x;
        x[j] = b < nblocks ? block_bounds[b] : make_uint4(0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u);
        y[j] = b < nblocks ? block_inst[b] : 0ull;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        dmin = min(dmin, x[j].x);
        dmax = max(dmax, x[j].y);
        tmin = min(tmin, x[j].z);
        tmax = max(tmax, x[j].w);
        inst += y[j];
      }
    }
    inst = wave_sum(inst);
    dmin = wave_min_u32_all(dmin);
    dmax = wave_max_u32_all(dmax);
    tmin = wave_min_u32_all(tmin);
    tmax = wave_max_u32_all(tmax);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      sb[w][0] = dmin;
      sb[w][1] = dmax;
      sb[w][2] = tmin;
      sb[w][3] = tmax;
      sinst[w] = inst;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = (int)(blockDim.x >> 6);
      for (int k = 1; k < nw; ++k) {
        dmin = min(dmin, sb[k][0]);
        dmax = max(dmax, sb[k][1]);
        tmin = min(tmin, sb[k][2]);
        tmax = max(tmax, sb[k][3]);
        inst += sinst[k];
      }
      s_bounds[0] = bounds[0] = dmin;
      s_bounds[1] = bounds[1] = dmax;
      s_bounds[2] = bounds[2] = tmin;
      s_bounds[3] = bounds[3] = tmax;
      s_exact = inst;
      *reinterpret_cast<uint64_t*>(bounds + 4) = inst;
    }
    __syncthreads();
  }
  const uint32_t V = (uint32_t)tot, M = (uint32_t)(tot >> 32);
  const uint64_t exact_m = s_exact;
  const bool over = exact_m > m_cap || exact_m >= 0x7FFFFFFFull || (uint64_t)V > v_cap;
  // keys per thread of the sort blocks: a graph fixes it, a plain frame takes
  // it from its key count (the host decides the same from the read-back)
  const int items = force_items ? force_items : sort_items_for(M);
  const uint32_t tile = 256u * (uint32_t)items;
  if (threadIdx.x == 0) {
    SortPlan p;
    plan_frame(V > 0, s_bounds[0], s_bounds[1], s_bounds[2], s_bounds[3], fuse, p);
    plan->passes = p.passes;
    plan->first = p.first;
    for (int k = 0; k < 8; ++k) {
      plan->shift[k] = p.shift[k];
      plan->width[k] = p.width[k];
    }
    plan->mode = p.mode;
    plan->tmin = p.tmin;
    plan->dmin = p.dmin;
    plan->dbits = p.dbits;
    plan->m = (uint64_t)M <= m_cap ? M : (uint32_t)m_cap;
    plan->v = (uint64_t)V <= v_cap ? V : (uint32_t)v_cap;
    plan->overflow = over ? 1u : 0u;
    plan->items = items;
    if (host_ret) {
      host_ret[0] = tot;
      host_ret[1] = (uint64_t)s_bounds[0] | ((uint64_t)s_bounds[1] << 32);
      host_ret[2] = (uint64_t)s_bounds[2] | ((uint64_t)s_bounds[3] << 32);
      host_ret[3] = exact_m;
    }
  }
  const uint32_t mc = (uint64_t)M <= m_cap ? M : (uint32_t)m_cap;
  for (int vw = threadIdx.x; vw < nviews; vw += blockDim.x) {
    const uint64_t a = vw == vw0 ? off_a0 : offsets[(size_t)vw * n];
    const uint64_t b = vw + 1 < nviews ? (vw == vw0 ? off_b0 : offsets[(size_t)(vw + 1) * n]) : tot;
    ViewSpan sp;
    sp.vis_base = (uint32_t)a;
    sp.visible = (uint32_t)b - (uint32_t)a;
    sp.key_base = (uint32_t)(a >> 32);
    sp.instances = (uint32_t)(b >> 32) - (uint32_t)(a >> 32);
    spans[vw] = sp;
    // the view's keys as a sort segment, clipped to the key capacity
    SortSeg sg;
    sg.key_base = sp.key_base < mc ? sp.key_base : mc;
    const uint32_t kend = sp.key_base + sp.instances < mc ? sp.key_base + sp.instances : mc;
    sg.count = kend - sg.key_base;
    sg.nblocks = (sg.count + tile - 1) / tile;
    sg.block_base = 0;
    segs[vw] = sg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // number the sort blocks through the segments
    uint32_t b = 0;
    for (int vw = 0; vw < nviews; ++vw) {
      segs[vw].block_base = b;
      b += segs[vw].nblocks;
    }
    plan->blocks = b;
  }
}

// Radius exactly as splat_geometry computes it (solve_eigenvalues_2x2,
// src/math/cholesky.mojo:28-51), from the cached screen covariance.
__device__ __forceinline__ int32_t radius_of(float c00, float c01, float c11) {
  const float mid = (c00 + c11) * 0.5f;
  const float diff = (c00 - c11) * 0.5f;
  const float root = sqrtf(diff * diff + c01 * c01);
  const int64_t rad = trunc_to_i64(ceilf(3.0f * sqrtf(mid + root)));
  return rad > INT32_MAX ? INT32_MAX : (int32_t)rad;
}

// Half-widths of a box around the splat centre outside of which the
// compositor's computed power exceeds 4.5 at every pixel centre (so alpha is
// 0 there and the splat changes nothing).  In exact arithmetic the power is
// <= 4.5 only inside the ellipse q <= 9 of the inverse covariance (i00, i01,
// i11), whose extents are 3 sqrt(i11/det) and 3 sqrt(i00/det); the box is
// 2% + 0.02 px larger, which covers the rounding of the power for inverse
// covariances with condition number up to ~1e4.  Worse-conditioned, singular
// or indefinite inverses get an infinite box (never skipped).
__device__ __forceinline__ void splat_box(float i00, float i01, float i11, float& hx, float& hy) {
  const float det = i00 * i11 - i01 * i01;
  const float tr = i00 + i11;
  if (!(i00 > 0.0f) || !(i11 > 0.0f) || !(det > 0.0f) || !(tr * tr <= 1.0e4f * det)) {
    hx = __builtin_inff();
    hy = __builtin_inff();
    return;
  }
  hx = 3.0f * sqrtf(i11 / det) * 1.02f + 0.02f;
  hy = 3.0f * sqrtf(i00 / det) * 1.02f + 0.02f;
}

// Pass 2: records and keys of every visible (view, Gaussian) at its scanned
// offsets, from the geometry k_count cached.  Keys are written by the whole
// block together: the block's keys are one contiguous run (the scan is in
// Gaussian order), each thread takes every 256th key of the run, finds the
// splat that owns it in the block's table of start offsets in LDS, and
// derives the tile from its rectangle -- coalesced stores, and a splat with
// thousands of tiles is shared by all 256 threads instead of one wave.  A
// thread's keys only move forward, so its search continues from the last
// splat it found (one LDS read while it stays inside a large splat).
// Emission order within a splat is ty-major then tx (pipeline.mojo:152-157).
//
// When the plan's first pass is k_emit's (plan->first == 1, plan_frame), the
// keys go to buffer 1 in the order that LSD pass 0 over the depth key's low 9
// bits would give, instead of emission order.  A splat's keys share that
// digit and stay together, so each splat's run moves as one: its output start
// is the view's first key + the digit's start in the view (exclusive scan of
// the digit's totals over the view's blocks) + the block's prefix in the
// digit's column (both from k_count's table, scanned in k_plan) +
// the tile counts of the block's earlier splats with the same digit.  The
// last term is a wave's earlier lanes with the same digit (found with one
// ballot per digit bit) plus the earlier waves' digit sums.  Emission order
// breaks the digit's ties, as the stable pass does.
// MULTI: more than one view, the Gaussian's SH coefficients loaded once and
// kept in registers across the views.  FUSABLE: a single-view frame whose
// k_count wrote the digit table in blocks of BLOCK Gaussians (the plan decides
// whether the pass runs here); without it the kernel carries none of that code.
template <bool MULTI, bool FUSABLE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_emit(SceneDev s, const ViewParams* __restrict__ views,
                                              int nviews, const uint64_t* __restrict__ offsets,
                                              const uint64_t* __restrict__ totals,
                                              const GeoCache* __restrict__ geo, uint64_t m_cap,
                                              uint64_t v_cap, float4* __restrict__ rec,
                                              int32_t* __restrict__ vis2orig,
                                              int32_t* __restrict__ radius,
                                              uint4* __restrict__ vmeta,
                                              uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ vals,
                                              SortPlan* __restrict__ plan,
                                              const uint32_t* __restrict__ etab,
                                              const uint32_t* __restrict__ etot,
                                              uint64_t* __restrict__ keys1,
                                              uint32_t* __restrict__ vals1) {
  static_assert(!(MULTI && FUSABLE), "the first pass runs in k_emit for single views only");
  constexpr int NW = BLOCK / 64;
  // digits per thread for the table row (512 digits over the block's threads)
  constexpr int DPT = kDigitBins / BLOCK > 0 ? kDigitBins / BLOCK : 1;
  // LDS, in one block: first (single view) each wave's SH staging, 16
  // Gaussians of 12 quads in rows padded to kShPad quads; then per thread
  // of the block (key start - block base, rect min, rect max, depth key) and
  // the visible index, and for the first sort pass each thread's output
  // start, per-wave digit sums (then their exclusive prefixes over the
  // waves) and the block's start of each digit in the view's keys
  constexpr int kShPad = 13;
  constexpr size_t kShBytes = MULTI ? 0 : (size_t)NW * 16 * kShPad * sizeof(float4);
  constexpr size_t kMetaBytes = (size_t)BLOCK * sizeof(uint4) + (size_t)BLOCK * 4 +
                                (FUSABLE ? (size_t)BLOCK * 4 + (size_t)NW * kDigitBins * 4 : 0) +
                                (size_t)kDigitBins * 4;
  __shared__ __align__(16) uint8_t s_raw[kShBytes > kMetaBytes ? kShBytes : kMetaBytes];
  __shared__ uint32_t s_w[NW];
  float4* const s_sh = reinterpret_cast<float4*>(s_raw);
  uint4* const s_meta = reinterpret_cast<uint4*>(s_raw);
  uint32_t* const s_vi = reinterpret_cast<uint32_t*>(s_meta + BLOCK);
  uint32_t* const s_delta = s_vi + BLOCK;
  uint32_t* const s_dst = s_delta + kDigitBins;
  uint32_t (*const s_wh)[kDigitBins] = reinterpret_cast<uint32_t (*)[kDigitBins]>(s_dst + BLOCK);
  if (plan->overflow) return;  // k_plan: the frame exceeds the capacities
  const bool fused = FUSABLE && plan->first == 1;  // only ever with etab (k_plan's fuse)
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const uint32_t nblk = gridDim.x;
  const uint32_t gb = xcd_block(blockIdx.x, nblk);  // as k_count numbers them
  const uint32_t first = gb * (uint32_t)BLOCK;
  const uint32_t i = first + (uint32_t)t;
  const bool valid = i < s.n;
  const uint32_t last = min(first + (uint32_t)BLOCK, s.n) - 1u;  // the block's last Gaussian
  const uint64_t nv = (uint64_t)nviews * s.n;
  float px = 0.0f, py = 0.0f, pz = 0.0f, op = 0.0f;
  if (valid) {
    px = s.px[i];
    py = s.py[i];
    pz = s.pz[i];
    op = s.op[i];
  }
  // keys hold the view-local tile: mode 1 plans sort 32-bit frame keys
  // (tile - tmin) << dbits | (depth key - dmin), mode 0 tile << 32 | depth key
  const bool frame_keys = plan->mode == 1;
  const uint32_t kt0 = plan->tmin, kd0 = plan->dmin;
  const int kdb = plan->dbits;
  uint64_t* __restrict__ kout = fused ? keys1 : keys;
  uint32_t* __restrict__ vout = fused ? vals1 : vals;
  uint32_t* __restrict__ keys32 = reinterpret_cast<uint32_t*>(kout);
  bool have_sh = false;
  float c[48];
  if (!MULTI) nviews = 1;
  for (int vw = 0; vw < nviews; ++vw) {
    const ViewParams& v = views[vw];
    // the block's runs: keys [base, end), visible splats [vbase, vend)
    const uint64_t row = (uint64_t)vw * s.n;
    const uint64_t o_first = offsets[row + first];
    const uint64_t o_end = row + last + 1 < nv ? offsets[row + last + 1] : *totals;
    const uint32_t base = (uint32_t)(o_first >> 32), end = (uint32_t)(o_end >> 32);
    const uint32_t vend = (uint32_t)o_end;
    if (base > end || (uint64_t)end > m_cap || (uint64_t)vend > v_cap) {  // uniform across the block
      if (t == 0) atomicOr(&plan->overflow, 1u);
      continue;
    }
    // first pass here: this block's row of the digit table (DPT digits per
    // thread), loaded ahead of the splat work
    uint32_t gl[DPT], ct[DPT], vkb = 0u;
#pragma unroll
    for (int e = 0; e < DPT; ++e) gl[e] = ct[e] = 0u;
    if (fused) {
      if (t * DPT < kDigitBins) {
#pragma unroll
        for (int e = 0; e < DPT; ++e) {
          const uint32_t col = (uint32_t)(t * DPT + e);
          gl[e] = etab[(size_t)col * nblk + gb];
          ct[e] = etot[col];
        }
      }
      vkb = (uint32_t)(offsets[row] >> 32);  // the view's first key
    }
    uint32_t vis = 0, ko = base, vi = 0, cnt = 0;
    if (valid) {
      const uint64_t off0 = offsets[row + i];
      const uint64_t off1 = row + i + 1 < nv ? offsets[row + i + 1] : *totals;
      vis = (uint32_t)off1 - (uint32_t)off0;
      vi = (uint32_t)off0;
      ko = (uint32_t)(off0 >> 32);
      cnt = (uint32_t)(off1 >> 32) - ko;
    }
    uint32_t rmin = 0, rmax = 0, dk = 0;
    // evaluate_sh_color(sh, i*48, normalize(pos - eye), degree)
    // src/math/spherical_harmonics.mojo:133-172, pipeline.mojo:133-139:
    // 0.5 + the coefficients times the basis, each channel summed in
    // coefficient order
    float cr = 0.5f, cg = 0.5f, cb = 0.5f;
    if (!MULTI) {
      // The wave's SH coefficients, loaded together 16 Gaussians at a time:
      // their 3 KB are contiguous, so each load instruction reads 1 KB of
      // consecutive addresses (a lane loading its own 192 B at a 192 B
      // stride meets a different cache line in every lane), and pass through
      // LDS (rows padded to 13 quads, so the 16 readers' quads fall in
      // distinct banks) to the lanes that own them, which add them into the
      // colour as they read them (the basis is ready before the loads, so
      // no lane holds its 48 coefficients).  The next group's loads are
      // issued before this group's exchange.  Only visible splats' quads
      // are loaded.
      float Y[16];
      {
        float dx = px - v.eye[0], dy = py - v.eye[1], dz = pz - v.eye[2];
        normalize3(dx, dy, dz);
        sh_basis(dx, dy, dz, Y);
      }
      const uint64_t visb = __ballot(vis != 0u);
      const int nq = (3 * v.sh_coeffs + 3) / 4;  // quads of coefficients the degree uses
      const uint32_t wfirst = first + (uint32_t)(w * 64);
      const float4* __restrict__ sh4 = reinterpret_cast<const float4*>(s.sh);
      float4* __restrict__ shw = s_sh + w * 16 * kShPad;
      float4 r[2][3];
      auto fetch = [&](int ph, float4 (&dst)[3]) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int idx = k * 64 + lane, gl = idx / 12, q = idx - gl * 12;
          dst[k] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (q < nq && ((visb >> (16 * ph + gl)) & 1ull))
            dst[k] = sh4[(size_t)(wfirst + 16u * ph + gl) * 12u + q];
        }
      };
      fetch(0, r[0]);
#pragma unroll
      for (int ph = 0; ph < 4; ++ph) {
        if (ph < 3) fetch(ph + 1, r[(ph + 1) & 1]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int idx = k * 64 + lane, gl = idx / 12, q = idx - gl * 12;
          shw[gl * kShPad + q] = r[ph & 1][k];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if ((lane >> 4) == ph && vis) {
#pragma unroll
          for (int q = 0; q < 12; ++q) {
            if (q < nq) {
              const float4 x = shw[(lane & 15) * kShPad + q];
              const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const int k = (4 * q + e) / 3, ch = (4 * q + e) % 3;  // constants once unrolled
                if (k < v.sh_coeffs) {
                  if (ch == 0) cr += xs[e] * Y[k];
                  else if (ch == 1) cg += xs[e] * Y[k];
                  else cb += xs[e] * Y[k];
                }
              }
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      __syncthreads();  // the staging's LDS is the tables' below
    }
    if (fused)
      for (int e = t; e < NW * kDigitBins; e += BLOCK) (&s_wh[0][0])[e] = 0u;
    if (vis) {
      const GeoCache gc = geo[row + i];
      const float sx = gc.a.x, sy = gc.a.y, depth = gc.a.z;
      const float c00 = gc.a.w, c01 = gc.b.x, c11 = gc.b.y;
      rmin = __float_as_uint(gc.b.z);
      rmax = __float_as_uint(gc.b.w);
      if (MULTI && !have_sh) {
        const float4* sh4 = reinterpret_cast<const float4*>(s.sh) + (size_t)i * 12u;
        const int nfloat = 3 * v.sh_coeffs;
#pragma unroll
        for (int qd = 0; qd < 12; ++qd) {
          const float4 x = (4 * qd < nfloat) ? sh4[qd] : make_float4(0.f, 0.f, 0.f, 0.f);
          c[4 * qd + 0] = x.x;
          c[4 * qd + 1] = x.y;
          c[4 * qd + 2] = x.z;
          c[4 * qd + 3] = x.w;
        }
        have_sh = true;
      }
      if (MULTI) {  // (the single view's colour was summed as its coefficients arrived)
        float dx = px - v.eye[0], dy = py - v.eye[1], dz = pz - v.eye[2];
        normalize3(dx, dy, dz);
        float Y[16];
        sh_basis(dx, dy, dz, Y);
        cr = cg = cb = 0.5f;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (k < v.sh_coeffs) {
            cr += c[3 * k + 0] * Y[k];
            cg += c[3 * k + 1] * Y[k];
            cb += c[3 * k + 2] * Y[k];
          }
        }
      }
      if (cr < 0.0f) cr = 0.0f; else if (cr > 1.0f) cr = 1.0f;
      if (cg < 0.0f) cg = 0.0f; else if (cg > 1.0f) cg = 1.0f;
      if (cb < 0.0f) cb = 0.0f; else if (cb > 1.0f) cb = 1.0f;
      float i00, i01, i11, hx, hy;
      inverse2d(c00, c01, c11, i00, i01, i11);
      splat_box(i00, i01, i11, hx, hy);
      const SplatRecord sr = make_record(sx, sy, i00, i01, i11, op, cr, cg, cb, depth, hx, hy);
      rec[4u * vi + 0u] = sr.q0;
      rec[4u * vi + 1u] = sr.q1;
      rec[4u * vi + 2u] = sr.q2;
      rec[4u * vi + 3u] = sr.q3;
      vis2orig[vi] = (int32_t)i;
      radius[vi] = radius_of(c00, c01, c11);
      vmeta[vi] = make_uint4(ko, rmin, rmax, 0u);
      dk = depth_key(depth);
    }
    // threads past the scene never own a key; one with no key shares its
    // start with the next splat that has keys, and the search takes the last
    // of equal starts
    s_meta[t] = make_uint4(valid ? ko - base : 0xFFFFFFFFu, rmin, rmax, dk);
    s_vi[t] = vi;
    if (fused) {
      // the digit's start in the view: an exclusive scan of its totals
      uint32_t all, csum = 0;
#pragma unroll
      for (int e = 0; e < DPT; ++e) csum += ct[e];
      uint32_t ex = block_exclusive_scan<uint32_t, BLOCK>(csum, s_w, all);
      if (t * DPT < kDigitBins) {
#pragma unroll
        for (int e = 0; e < DPT; ++e) {
          s_delta[t * DPT + e] = vkb + ex + gl[e];
          ex += ct[e];
        }
      }
      // tile counts of this wave's earlier lanes with the same digit (such
      // lanes are rare: the loop runs as often as the most of them below
      // any one lane)
      const bool has = cnt > 0;
      const uint32_t d = dk & kEmitDigitMask;
      uint64_t peers = __ballot(has);
#pragma unroll
      for (int bit = 0; bit < kEmitDigitBits; ++bit) {
        const bool on = (d >> bit) & 1u;
        const uint64_t bb = __ballot(on);
        peers &= on ? bb : ~bb;
      }
      uint64_t below = has ? (peers & lanemask_lt()) : 0ull;
      uint32_t r = 0;
      while (__ballot(below != 0ull)) {  // wave-uniform
        const int src = below ? __ffsll((unsigned long long)below) - 1 : lane;
        const uint32_t cs = (uint32_t)__shfl((int)cnt, src, 64);
        if (below) {
          r += cs;
          below &= below - 1ull;
        }
      }
      if (has) atomicAdd(&s_wh[w][d], cnt);
      __syncthreads();
      // each digit's sums become the earlier waves' totals
      for (int dd = t; dd < kDigitBins; dd += BLOCK) {
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          const uint32_t x = s_wh[k][dd];
          s_wh[k][dd] = run;
          run += x;
        }
      }
      __syncthreads();
      if (has) s_dst[t] = s_delta[d] + s_wh[w][d] + r;
    }
    __syncthreads();
    const uint32_t total = end - base;
    int g = 0;  // the splat owning key p (thread 0 starts at 0)
    for (uint32_t p = (uint32_t)t; p < total; p += (uint32_t)BLOCK) {
      if (g < BLOCK - 1 && s_meta[g + 1].x <= p) {
        int lo = g + 1, hi = BLOCK - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_meta[mid].x <= p) lo = mid;
          else hi = mid - 1;
        }
        g = lo;
      }
      const uint4 m = s_meta[g];
      const uint32_t local = p - m.x;
      const uint32_t min_tx = m.y & 0xFFFFu, min_ty = m.y >> 16;
      const uint32_t wtiles = (m.z & 0xFFFFu) - min_tx + 1u;
      const uint32_t r = local / wtiles;
      const uint32_t ty = min_ty + r;
      const uint32_t tx = min_tx + (local - r * wtiles);
      const uint32_t tile = ty * (uint32_t)v.tiles_x + tx;  // within the view
      const uint32_t dst = fused ? s_dst[g] + local : base + p;
      if (fused && dst >= m_cap) continue;  // never with a consistent digit table
      if (frame_keys)
        keys32[dst] = ((tile - kt0) << kdb) | (m.w - kd0);
      else
        kout[dst] = ((uint64_t)tile << 32) | (uint64_t)m.w;
      vout[dst] = s_vi[g];
    }
    __syncthreads();
  }
}

// The segment (view) holding key position i.
__device__ __forceinline__ int seg_of(const SortSeg* __restrict__ segs, int nseg, uint32_t i) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].key_base <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint64_t sorted_key(const void* keys, uint32_t i, bool frame_keys) {
  if (frame_keys) return static_cast<const uint32_t*>(keys)[i];
  return static_cast<const uint64_t*>(keys)[i];
}

// Tile ranges from the sorted keys.  A run of one tile starts where the
// view-local tile changes or a view's segment begins; the thread at a run's
// start also closes the previous run, and the last key closes the final
// one.  A block takes 2048 consecutive keys, 8 per thread (all loads issued
// before any use), and finds each key's predecessor in an LDS copy of the
// block's keys; thread 0 loads the key before the block.
constexpr int kRangeItems = 8;
constexpr uint32_t kRangeKeys = 256u * kRangeItems;

__global__ __launch_bounds__(256) void k_ranges(const uint64_t* __restrict__ k0,
                                                const uint64_t* __restrict__ k1,
                                                const SortPlan* __restrict__ plan,
                                                const SortSeg* __restrict__ segs, int nseg,
                                                const ViewParams* __restrict__ views,
                                                int2* __restrict__ ranges, uint32_t tiles) {
  __shared__ uint64_t s_key[kRangeKeys + 1];  // s_key[1 + j]: the block's key j
  const uint32_t b0 = blockIdx.x * kRangeKeys;
  // the segment of the block's first key (a search on uniform values that
  // reads only segs, issued with the plan's loads), then forward with the
  // thread's keys
  int sv = seg_of(segs, nseg, b0);
  const uint32_t m = plan->m;
  if (b0 >= m || plan->overflow) return;  // uniform across the block
  const void* keys = (plan->passes & 1) ? (const void*)k1 : (const void*)k0;
  const SortPlan p = *plan;
  const bool fk = p.mode == 1;
  const int t = threadIdx.x;
  uint64_t k[kRangeItems];
#pragma unroll
  for (int j = 0; j < kRangeItems; ++j) {
    const uint32_t i = b0 + (uint32_t)j * 256u + (uint32_t)t;
    k[j] = i < m ? sorted_key(keys, i, fk) : 0ull;
  }
  if (t == 0) s_key[0] = b0 > 0 ? sorted_key(keys, b0 - 1, fk) : 0ull;
#pragma unroll
  for (int j = 0; j < kRangeItems; ++j) s_key[1 + j * 256 + t] = k[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRangeItems; ++j) {
    const uint32_t i = b0 + (uint32_t)j * 256u + (uint32_t)t;
    if (i >= m) break;
    const uint64_t kp = s_key[j * 256 + t];  // key i - 1
    while (sv + 1 < nseg && segs[sv + 1].key_base <= i) ++sv;
    const bool seg_start = i == segs[sv].key_base;
    const uint32_t tile = frame_key_tile(k[j], p);
    const uint32_t gt = views[sv].tile_base + tile;
    if (i == 0 || seg_start || frame_key_tile(kp, p) != tile) {
      if (gt < tiles) ranges[gt].x = (int32_t)i;
      if (i > 0) {  // close the previous run, in its own view
        int sp = sv;
        if (seg_start) {
          sp = sv - 1;
          while (sp > 0 && segs[sp].count == 0) --sp;
        }
        const uint32_t gp = views[sp].tile_base + frame_key_tile(kp, p);
        if (gp < tiles) ranges[gp].y = (int32_t)i;
      }
    }
    if (i + 1 == m && gt < tiles) ranges[gt].y = (int32_t)m;
  }
}

// Frame keys (either width) back to the reference's 64-bit keys, for
// lumen_ex_copy_stage.
__global__ void k_decode_keys(const void* __restrict__ keys, uint32_t n,
                              const SortPlan* __restrict__ plan, const SortSeg* __restrict__ segs,
                              int nseg, const ViewParams* __restrict__ views,
                              uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const SortPlan p = *plan;
  const uint64_t k = sorted_key(keys, i, p.mode == 1);
  out[i] = frame_key_to_key(k, p, views[seg_of(segs, nseg, i)].tile_base);
}

// Ranges of an arbitrary sorted key array (lumen_ex_build_ranges).
__global__ __launch_bounds__(256) void k_ranges_plain(const uint64_t* __restrict__ keys,
                                                      uint32_t m, int2* __restrict__ ranges,
                                                      uint32_t tiles) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= m) return;
  const uint32_t tile = (uint32_t)(keys[i] >> 32);
  if (tile >= tiles) return;
  if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != tile) ranges[tile].x = (int32_t)i;
  if (i == m - 1 || (uint32_t)(keys[i + 1] >> 32) != tile) ranges[tile].y = (int32_t)(i + 1);
}

// The view owning frame tile gt (views are in tile_base order).
__device__ __forceinline__ int view_of_tile(const ViewParams* __restrict__ views, int nviews,
                                            uint32_t gt) {
  int lo = 0, hi = nviews - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (views[mid].tile_base <= gt) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One workgroup per tile; lane t owns PIX vertically adjacent pixels
// (column t % tile_w, rows (t / tile_w) PIX + p), so a wave covers 64 PIX /
// tile_w full rows.  Splat records of the tile's sorted range are staged
// through LDS BATCH at a time with the strip mask of each (and the power
// limit past which alpha < 1/255); every wave then compacts the batch to the
// splats whose box meets its strip (ballot + prefix count, order kept), and
// each pixel walks that list front to back exactly as PixelAccumulator does
// (src/raster/compositor.mojo:25-89).  A lane's pixels share the splat's
// loads and the x terms of the power, i00 dx dx and 2 i01 dx (each pixel's
// power keeps the reference's expression tree).  A pixel stops at
// T <= 1e-4, a lane when both of its pixels have, the block when all lanes have.
// Frame tiles ordered by key count, heaviest first (classes floor(log2 n),
// any order inside a class): the compositor's blocks take tiles in this
// order, so the longest tiles start first instead of landing in the launch's
// tail (a longest-processing-time-first schedule).  Each tile's pixels do not
// depend on the order.  kOrderBlocks blocks each count every tile's class
// (per class: all tiles, and the tiles before the block's own range), then
// place their own range's tiles; a wave adds to a class counter once per
// distinct class among its lanes.
constexpr int kOrderClasses = 32;
__device__ __forceinline__ uint32_t tile_class(const int2* __restrict__ ranges, bool over,
                                               uint32_t i) {
  const int2 r = ranges[i];  // loaded alongside the flag, then dropped on overflow
  const uint32_t n = (!over && r.x >= 0 && r.y > r.x) ? (uint32_t)(r.y - r.x) : 0u;
  return n ? (uint32_t)__clz(n) : 31u;  // 31 - floor(log2 n): the heaviest class is 0
}
// Lanes of the wave whose class equals this lane's (inactive lanes excluded).
__device__ __forceinline__ uint64_t class_peers(bool valid, uint32_t c) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 5; ++b) {
    const bool on = (c >> b) & 1u;
    const uint64_t bb = __ballot(on);
    peers &= on ? bb : ~bb;
  }
  return peers;
}
// Frames of more tiles than this keep their tiles in order (written by a
// grid of blocks): ordering them would take longer than the launch's tail
// costs, which is short against so many tiles.
constexpr uint32_t kOrderMaxTiles = 65536;
constexpr uint32_t kOrderBlocks = 8;
// Tiles per ordering block: a multiple of 64, so each wave's tiles lie on one
// side of every block's range start.
__host__ __device__ inline uint32_t order_chunk(uint32_t tiles) {
  return (tiles + kOrderBlocks * 64u - 1u) / (kOrderBlocks * 64u) * 64u;
}
__global__ __launch_bounds__(1024) void k_tile_order(const int2* __restrict__ ranges,
                                                     uint32_t tiles,
                                                     const SortPlan* __restrict__ plan,
                                                     uint32_t* __restrict__ order) {
  constexpr int kLoads = 8;  // range loads a thread has in flight
  __shared__ uint32_t cnt[kOrderClasses], pre[kOrderClasses];
  __shared__ uint8_t s_cls[kOrderMaxTiles / kOrderBlocks];  // the block's own tiles' classes
  const int t = threadIdx.x;
  if (tiles > kOrderMaxTiles) {
    const uint32_t i = blockIdx.x * 1024u + (uint32_t)t;
    if (i < tiles) order[i] = i;
    return;
  }
  const uint32_t chunk = order_chunk(tiles);
  const uint32_t lo = blockIdx.x * chunk;
  const uint32_t hi = min(tiles, lo + chunk);
  const bool over = plan->overflow != 0;
  const uint64_t lt = lanemask_lt();
  if (t < kOrderClasses) cnt[t] = pre[t] = 0;
  __syncthreads();
  for (uint32_t base = 0; base < tiles; base += 1024u * kLoads) {
    uint32_t c[kLoads];
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
      const uint32_t i = base + (uint32_t)j * 1024u + (uint32_t)t;
      c[j] = i < tiles ? tile_class(ranges, over, i) : 0u;
    }
#pragma unroll
    for (int j = 0; j < kLoads; ++j) {
      const uint32_t i = base + (uint32_t)j * 1024u + (uint32_t)t;
      const bool valid = i < tiles;
      if (i >= lo && i < hi) s_cls[i - lo] = (uint8_t)c[j];
      const uint64_t peers = class_peers(valid, c[j]);
      if (valid && (peers & lt) == 0) {
        const uint32_t k = (uint32_t)__popcll(peers);
        atomicAdd(&cnt[c[j]], k);
        if (i < lo) atomicAdd(&pre[c[j]], k);  // the whole wave lies before lo
      }
    }
  }
  __syncthreads();
  if (t == 0) {  // the block's first position in each class, heaviest class first
    uint32_t s = 0;
    for (int c = 0; c < kOrderClasses; ++c) {
      const uint32_t n = cnt[c];
      cnt[c] = s + pre[c];
      s += n;
    }
  }
  __syncthreads();
  for (uint32_t base = lo; base < hi; base += 1024) {
    const uint32_t i = base + t;
    const bool valid = i < hi;
    const uint32_t c = valid ? (uint32_t)s_cls[i - lo] : 0u;
    const uint64_t peers = class_peers(valid, c);
    const int leader = peers ? __ffsll((unsigned long long)peers) - 1 : 0;
    uint32_t first = 0;
    if (valid && (peers & lt) == 0) first = atomicAdd(&cnt[c], (uint32_t)__popcll(peers));
    first = __shfl(first, leader, 64);
    if (valid) order[first + (uint32_t)__popcll(peers & lt)] = i;
  }
}

// SAVE: the frame keeps each pixel's final transmittance and evaluation
// count (for the backward); without it the blend does not track where each
// pixel terminated.
template <int BATCH, int PIX, bool SAVE>
__global__ __launch_bounds__(256) void k_composite(
    const int2* __restrict__ ranges, const uint32_t* __restrict__ v0,
    const uint32_t* __restrict__ v1, const SortPlan* __restrict__ plan,
    const float4* __restrict__ rec, const ViewParams* __restrict__ views, int nviews,
    const ViewSpan* __restrict__ spans, float* __restrict__ rgb, float* __restrict__ depth,
    float* __restrict__ t_final, int32_t* __restrict__ n_eval, const uint32_t* __restrict__ order) {
  static_assert(BATCH * sizeof(float4) <= 65536, "list entries are 16-bit byte offsets");
  __shared__ float4 s0[BATCH], s1[BATCH], s2[BATCH];
  __shared__ uint8_t s_mask[BATCH];
  // one list per wave, of the staged splats' byte offsets in s0 / s1 / s2
  // (the loop's LDS reads then need no index arithmetic)
  __shared__ uint16_t s_list[4 / PIX][BATCH];
  const uint32_t gt = order[blockIdx.x];
  const int vw = view_of_tile(views, nviews, gt);
  const ViewParams& v = views[vw];
  const int nthreads = v.tile_w * v.tile_h / PIX;
  const int nwaves = nthreads >> 6;
  const int rows = 64 * PIX / v.tile_w;
  const int tile = (int)(gt - v.tile_base);
  const int tx = tile % v.tiles_x, ty = tile / v.tiles_x;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int px = tx * v.tile_w + t % v.tile_w;
  const int py0 = ty * v.tile_h + (t / v.tile_w) * PIX;
  const uint32_t* __restrict__ vals = (plan->passes & 1) ? v1 : v0;
  const int2 rg0 = ranges[gt];  // loaded alongside the flag
  const int2 rg = plan->overflow ? make_int2(-1, -1) : rg0;
  // A pixel is live while T > kTransmittanceEps: it starts at 1 inside the
  // image (0 outside, where nothing is written) and the blend that takes it
  // to kTransmittanceEps or below is the pixel's last.
  float C0[PIX], C1[PIX], C2[PIX], T[PIX], D[PIX];
  // the range position of the pixel's last blended splat (SAVE): a pixel
  // that terminated blends nothing after the splat that ended it, so its
  // evaluation count is that position + 1; one that did not gets the list
  // length
  int32_t lastb[PIX];
#pragma unroll
  for (int p = 0; p < PIX; ++p) {
    C0[p] = C1[p] = C2[p] = D[p] = 0.0f;
    T[p] = (px < v.width && py0 + p < v.height) ? 1.0f : 0.0f;
    lastb[p] = 0;
  }
  if (rg.x != -1 && rg.x < rg.y) {
    const float pxc = (float)px + 0.5f;
    float pyc[PIX];
#pragma unroll
    for (int p = 0; p < PIX; ++p) pyc[p] = (float)(py0 + p) + 0.5f;
    bool any = false;
#pragma unroll
    for (int p = 0; p < PIX; ++p) any |= T[p] > kTransmittanceEps;
    for (int base = rg.x; base < rg.y; base += BATCH) {
      if (__syncthreads_count(any) == 0) break;
      const int n = min(BATCH, rg.y - base);
      for (int j = t; j < n; j += nthreads) {
        const uint32_t vi = vals[base + j];
        SplatRecord r;
        r.q0 = rec[4u * vi + 0u];
        r.q1 = rec[4u * vi + 1u];
        r.q2 = rec[4u * vi + 2u];
        r.q3 = rec[4u * vi + 3u];
        // the conic comes halved (the power's factor 0.5; a power small
        // enough for the halving to lose bits, < 2^-125, gives exp(-power) =
        // 1 either way)
        s0[j] = r.q0;  // x, y, i00/2, i01
        s1[j] = r.q1;  // i11/2, opacity, power limit (the cut is not used here)
        s2[j] = r.q2;  // r, g, b, depth
        s_mask[j] = (uint8_t)record_mask(r, tx * v.tile_w, ty * v.tile_h, v.tile_w, rows, 1,
                                         nwaves);
      }
      __syncthreads();
      // this wave's list of the batch's splats, in order
      uint32_t cnt = 0;
      for (int c0 = 0; c0 < n; c0 += 64) {
        const int j = c0 + lane;
        const bool hit = j < n && ((s_mask[j] >> w) & 1u);
        const uint64_t bal = __ballot(hit);
        if (hit) s_list[w][cnt + __popcll(bal & lanemask_lt())] = (uint16_t)(j * sizeof(float4));
        cnt += __popcll(bal);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // every lane walks the list until none of the wave's pixels is live
      // (a lane whose pixels are done fails their T tests), so the loop's
      // exit is wave-uniform
      if (__ballot(any)) {
        for (uint32_t q = 0; q < cnt; ++q) {
          const uint32_t jo = s_list[w][q];
          auto staged = [jo](const float4* arr) {
            return *reinterpret_cast<const float4*>(reinterpret_cast<const uint8_t*>(arr) + jo);
          };
          const float4 a = staged(s0);  // x, y, i00/2, i01
          const float4 b = staged(s1);  // i11/2, opacity, power limit
          const float4 c = staged(s2);  // r, g, b, depth
          // compute_gaussian_pixel_alpha, compositor.mojo:62-89
          const float dx = pxc - a.x;
          const float txx = a.z * dx * dx;
          const float txy = a.w * dx;
#pragma unroll
          for (int p = 0; p < PIX; ++p) {
            const float dy = pyc[p] - a.y;
            float power = txx + txy * dy + b.x * dy * dy;
            if (power < 0.0f) power = 0.0f;
            // a finished pixel, or power > 4.5 or alpha < 1/255 (power_limit)
            if (!(T[p] > kTransmittanceEps) || power > b.z) continue;
            float alpha = b.y * splat_exp_neg(-power);
            if (alpha > 0.99f) alpha = 0.99f;
            // PixelAccumulator.blend, compositor.mojo:25-53
            if (alpha < kMinAlpha) continue;
            const float wgt = alpha * T[p];
            C0[p] += c.x * wgt;
            C1[p] += c.y * wgt;
            C2[p] += c.z * wgt;
            D[p] += c.w * wgt;
            T[p] *= (1.0f - alpha);
            if (SAVE) lastb[p] = base + (int)(jo / sizeof(float4));
          }
          // T is never negative, so its bits order as the values do: one
          // unsigned maximum and one compare (a NaN T, from NaN inputs, counts
          // as live here and keeps the wave walking; its pixel's own test
          // above skips every splat, as the reference's does)
          uint32_t tbits = 0u;
#pragma unroll
          for (int p = 0; p < PIX; ++p) tbits = max(tbits, __float_as_uint(T[p]));
          if (__ballot(tbits > __float_as_uint(kTransmittanceEps)) == 0) break;
        }
        any = false;
#pragma unroll
        for (int p = 0; p < PIX; ++p) any |= T[p] > kTransmittanceEps;
      }
    }
  }
  // PixelAccumulator.finalize + RenderOutput writes, pipeline.mojo:182-229.
  // rasterize_scene returns its zero-filled output before any background fill
  // when the view has no tile instance (pipeline.mojo:63-65, 162-163).
  const bool white = v.white_background && spans[vw].instances > 0;
#pragma unroll
  for (int p = 0; p < PIX; ++p) {
    if (px >= v.width || py0 + p >= v.height) continue;
    const size_t pix = v.pix_base + (size_t)(py0 + p) * v.width + px;
    const float bg = white ? T[p] : 0.0f;
    rgb[3 * pix + 0] = C0[p] + bg;
    rgb[3 * pix + 1] = C1[p] + bg;
    rgb[3 * pix + 2] = C2[p] + bg;
    if (depth) depth[pix] = v.compute_depth ? D[p] : 0.0f;
    if (SAVE) {
      t_final[pix] = T[p];
      n_eval[pix] = T[p] <= kTransmittanceEps ? lastb[p] - rg.x + 1 : rg.y - rg.x;
    }
  }
}

// The records in the 48-byte layout lumen_ex_copy_stage documents (cx cy i00
// 2 i01 | i11 op r g | b depth 0 0): the reference's projected splat, from
// the 64-byte compositor records (doubling the halved conic is exact).
__global__ void k_records_out(const float4* __restrict__ rec, uint32_t v, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v) return;
  const float4 q0 = rec[4u * i + 0u], q1 = rec[4u * i + 1u], q2 = rec[4u * i + 2u];
  out[3u * i + 0u] = make_float4(q0.x, q0.y, 2.0f * q0.z, 2.0f * q0.w);
  out[3u * i + 1u] = make_float4(2.0f * q1.x, q1.y, q2.x, q2.y);
  out[3u * i + 2u] = make_float4(q2.z, q2.w, 0.0f, 0.0f);
}

__global__ void k_minmax_u64(const uint64_t* __restrict__ keys, uint64_t n,
                             unsigned long long* __restrict__ mm) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long lo = ~0ull, hi = 0ull;
  for (uint64_t j = i; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = keys[j];
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long a = __shfl_xor(lo, d, 64);
    const unsigned long long b = __shfl_xor(hi, d, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&mm[0], lo);
    atomicMax(&mm[1], hi);
  }
}

__global__ void k_set_plan(SortPlan* plan, SortSeg* seg, int passes, const int32_t* digits,
                           uint32_t m, int items) {
  seg->key_base = 0;  // one segment: all the keys
  seg->count = m;
  seg->block_base = 0;
  seg->nblocks = (m + 256u * items - 1) / (256u * items);
  plan->blocks = seg->nblocks;
  plan->items = items;
  plan->passes = passes;
  plan->first = 0;
  for (int k = 0; k < 8; ++k) {
    plan->shift[k] = k < passes ? 8 * digits[k] : 0;
    plan->width[k] = k < passes ? 8 : 0;
  }
  plan->mode = 0;
  plan->m = m;
  plan->v = 0;
  plan->overflow = 0;
}

}  // namespace lumen

using namespace lumen;

// ---- validation -------------------------------------------------------------

static bool plane_ok(const float* p) { return p != nullptr; }

static LumenStatus check_config(const LumenExRenderConfig* cfg) {
  const uint32_t tw = cfg->tile_width, th = cfg->tile_height;
  const uint32_t threads = tw * th;
  if (tw == 0 || th == 0 || threads < 64 || threads > 256 || threads % 64 != 0 || 64 % tw != 0)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                "tile %ux%u unsupported (64, 128, 192 or 256 pixels per tile, width dividing 64)",
                tw, th);
  const uint32_t batch = cfg->batch_depth ? cfg->batch_depth : 256u;
  if (batch != 64 && batch != 128 && batch != 256)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "batch_depth %u not in {64,128,256}", batch);
  if (cfg->max_sh_degree < 0 || cfg->max_sh_degree > 3)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "max_sh_degree %d not in 0..3", cfg->max_sh_degree);
  return LUMEN_SUCCESS;
}

static LumenStatus check_scene(const LumenExScene* sc) {
  if (sc->count >= 0x7FFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "%llu Gaussians exceed 2^31-1",
                (unsigned long long)sc->count);
  if (sc->count > 0 &&
      !(plane_ok(sc->pos_x) && plane_ok(sc->pos_y) && plane_ok(sc->pos_z) &&
        plane_ok(sc->scale_x) && plane_ok(sc->scale_y) && plane_ok(sc->scale_z) &&
        plane_ok(sc->rot_w) && plane_ok(sc->rot_x) && plane_ok(sc->rot_y) && plane_ok(sc->rot_z) &&
        plane_ok(sc->opacity) && plane_ok(sc->sh)))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "null scene plane");
  return LUMEN_SUCCESS;
}

static LumenStatus check_cameras(const LumenExCamera* cams, uint32_t n) {
  if (n == 0 || n > 4096) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "%u views (1..4096)", n);
  for (uint32_t k = 0; k < n; ++k)
    if (cams[k].width == 0 || cams[k].height == 0 || cams[k].width > 16384 ||
        cams[k].height > 16384)
      return fail(LUMEN_ERROR_INVALID_ARGUMENT, "view %u: bad image size %ux%u", k, cams[k].width,
                  cams[k].height);
  return LUMEN_SUCCESS;
}

// ---- frame construction ---------------------------------------------------

struct FrameShape {
  uint32_t nviews;
  uint64_t tiles;   // frame tiles (sum over views)
  uint64_t pixels;  // frame pixels (sum over views)
  bool inline_views = false;  // k_count takes the views as arguments and writes the device copy
  // k_count writes per-block histograms of k_emit's digit, so the sort's
  // first pass can run inside k_emit (frame_fusable)
  bool fusable = false;
};

// Frames whose first sort pass may run inside k_emit: one view of at least
// 2048 tiles.  The per-block histograms cost about 8 bytes per Gaussian
// (written, scanned, read), against 20 bytes per key for a sort pass; frames
// of smaller images hold few keys per Gaussian, and their passes are short.
// A frame that keeps its unsorted keys (LUMEN_EX_KEEP_STAGES) emits them in
// the reference's order instead.
static bool frame_fusable(uint32_t nviews, uint64_t tiles, uint64_t n, bool keep) {
  return !keep && n > 0 && nviews == 1 && tiles >= 2048ull;
}

// Gaussians per block of k_count and k_emit (the digit table's rows).
static uint32_t count_block(const FrameShape& fs) { return fs.fusable ? kFusedBlock : 256u; }

// Whether k_count's block sums replace the scan's reduce pass: one view (its
// blocks are consecutive pieces of the scanned array, and count_block
// divides the scan's chunk) of at most kScanFuse chunks.
static bool scan_pieces(const FrameShape& fs, uint64_t n) {
  static_assert(kScanChunk % kFusedBlock == 0 && kScanChunk % 256 == 0, "pieces of a chunk");
  return fs.nviews == 1 && n > 0 && div_up(n, (uint64_t)kScanChunk) <= kScanFuse;
}

// Per-view parameters on the host and their device copy: through pinned
// staging, ordered on the stream, or (allow_inline, at most kInlineViews
// views, a frame that runs k_count) written by k_count from its arguments.
static LumenStatus upload_views(LumenExContext* c, const LumenExCamera* cams, uint32_t n,
                                const LumenExRenderConfig* cfg, FrameShape& fs, hipStream_t st,
                                bool allow_inline = false) {
  c->host_views.resize(n);
  uint64_t tb = 0, pb = 0;
  for (uint32_t k = 0; k < n; ++k) {
    ViewParams& v = c->host_views[k];
    fill_view(v, &cams[k], cfg);
    v.tile_base = (uint32_t)tb;
    v.view = k;
    v.pix_base = pb;
    tb += (uint64_t)v.tiles_x * v.tiles_y;
    pb += (uint64_t)v.width * v.height;
  }
  if (tb >= 0x7FFFFFFFull) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "too many tiles in the frame");
  fs.nviews = n;
  fs.tiles = tb;
  fs.pixels = pb;
  const size_t bytes = sizeof(ViewParams) * n;
  if (allow_inline && n <= (uint32_t)kInlineViews) {
    LUMEN_TRY(ensure(c->views, bytes));
    LUMEN_TRY(ensure(c->spans, sizeof(ViewSpan) * n));
    LUMEN_TRY(ensure(c->segs, sizeof(SortSeg) * n));
    fs.inline_views = true;
    return LUMEN_SUCCESS;
  }
  // the staging buffer is reused: the previous frame's copy out of it (which
  // may still be queued behind that frame's work) must have run
  if (c->ev_views_pending) {
    LUMEN_HIP(hipEventSynchronize(c->ev_views));
    c->ev_views_pending = false;
  }
  if (c->pinned_views_cap < bytes) {
    if (c->pinned_views) LUMEN_HIP(hipHostFree(c->pinned_views));
    c->pinned_views = nullptr;
    LUMEN_HIP(hipHostMalloc(&c->pinned_views, bytes, hipHostMallocDefault));
    c->pinned_views_cap = bytes;
  }
  memcpy(c->pinned_views, c->host_views.data(), bytes);
  LUMEN_TRY(ensure(c->views, bytes));
  LUMEN_TRY(ensure(c->spans, sizeof(ViewSpan) * n));
  LUMEN_TRY(ensure(c->segs, sizeof(SortSeg) * n));
  LUMEN_HIP(hipMemcpyAsync(c->views.ptr, c->pinned_views, bytes, hipMemcpyHostToDevice, st));
  if (!c->ev_views) LUMEN_HIP(hipEventCreateWithFlags(&c->ev_views, hipEventDisableTiming));
  LUMEN_HIP(hipEventRecord(c->ev_views, st));
  c->ev_views_pending = true;
  return LUMEN_SUCCESS;
}

static LumenStatus composite_launch(uint32_t batch, uint64_t tiles, uint32_t tile_w,
                                    uint32_t tile_h, hipStream_t st, LumenExContext* c, uint32_t nviews,
                                    float* rgb, float* depth, float* t_final, int32_t* n_eval) {
  const int2* ranges = c->ranges.as<int2>();
  const uint32_t* v0 = c->vals[0].as<uint32_t>();
  const uint32_t* v1 = c->vals[1].as<uint32_t>();
  const SortPlan* plan = c->plan.as<SortPlan>();
  const float4* rec = c->rec.as<float4>();
  const ViewParams* views = c->views.as<ViewParams>();
  const ViewSpan* spans = c->spans.as<ViewSpan>();
  const uint32_t threads = tile_w * tile_h;
  uint32_t* order = c->tile_order.as<uint32_t>();
  const uint32_t ob = tiles > kOrderMaxTiles ? div_up((uint32_t)tiles, 1024u) : kOrderBlocks;
  hipLaunchKernelGGL(k_tile_order, dim3(ob), dim3(1024), 0, st, ranges, (uint32_t)tiles, plan,
                     order);
  LUMEN_LAUNCHED();
  // two pixels per lane where the tile has at least 128 pixels
  const bool two = threads % 128 == 0;
  const dim3 grid((uint32_t)tiles), block(two ? threads / 2 : threads);
  const bool save = t_final != nullptr && n_eval != nullptr;
#define LUMEN_COMPOSITE(B, P)                                                                   \
  hipLaunchKernelGGL((save ? k_composite<B, P, true> : k_composite<B, P, false>), grid, block, 0, \
                     st, ranges, v0, v1, plan, rec, views, (int)nviews, spans, rgb, depth,       \
                     t_final, n_eval, order)
  if (two) {
    if (batch == 64) LUMEN_COMPOSITE(64, 2);
    else if (batch == 128) LUMEN_COMPOSITE(128, 2);
    else LUMEN_COMPOSITE(256, 2);
  } else {
    if (batch == 64) LUMEN_COMPOSITE(64, 1);
    else if (batch == 128) LUMEN_COMPOSITE(128, 1);
    else LUMEN_COMPOSITE(256, 1);
  }
#undef LUMEN_COMPOSITE
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

// The buffers below are named to a `need(DevBuf&, bytes)` callable: ensure()
// when rendering, a byte count for lumen_ex_workspace_size.
static LumenStatus need_ensure(DevBuf& b, size_t bytes) { return ensure(b, bytes); }

// Buffers that depend on N, the views and the pixel count.
template <class Need>
static LumenStatus frame_buffers(LumenExContext* c, Need&& need, uint64_t N, const FrameShape& fs,
                                 bool save_pixels) {
  const uint64_t nv = (uint64_t)fs.nviews * N;
  LUMEN_TRY(need(c->packed, (nv ? nv : 1) * 8));
  LUMEN_TRY(need(c->geo, (nv ? nv : 1) * sizeof(GeoCache)));
  // per count block: bounds (16 B), then after all of them the exact instance sums (8 B)
  LUMEN_TRY(need(c->block_bounds, (uint64_t)div_up(N ? N : 1, 256) * 24));
  // the scan's chunk sums, or k_count's block sums (scan_pieces)
  LUMEN_TRY(need(c->part, std::max<uint64_t>(div_up(nv ? nv : 1, kScanChunk),
                                             div_up(N ? N : 1, 256)) * 8));
  // 8 B totals, the 16 B bounds and the 8 B exact instance total (one read-back)
  LUMEN_TRY(need(c->totals, 32));
  LUMEN_TRY(need(c->plan, sizeof(SortPlan)));
  LUMEN_TRY(need(c->ranges, fs.tiles * 8));
  LUMEN_TRY(need(c->tile_order, fs.tiles * 4));
  if (fs.fusable) {
    LUMEN_TRY(need(c->etab, (uint64_t)kDigitBins * div_up(N, kFusedBlock) * 4));
    LUMEN_TRY(need(c->epre, (uint64_t)kDigitBins * div_up(N, kFusedBlock) * 4));
    LUMEN_TRY(need(c->etot, (uint64_t)kDigitBins * 4));
  }
  if (save_pixels) {
    LUMEN_TRY(need(c->t_final, fs.pixels * 4));
    LUMEN_TRY(need(c->n_eval, fs.pixels * 4));
  }
  return LUMEN_SUCCESS;
}

static LumenStatus ensure_frame_buffers(LumenExContext* c, uint64_t N, const FrameShape& fs,
                                        bool save_pixels) {
  return frame_buffers(c, need_ensure, N, fs, save_pixels);
}

// Buffers sized by the frame's visible splats and tile instances.
template <class Need>
static LumenStatus capacity_buffers(LumenExContext* c, Need&& need, uint64_t v_cap,
                                    uint64_t m_cap) {
  LUMEN_TRY(need(c->rec, (v_cap ? v_cap : 1) * sizeof(SplatRecord)));
  LUMEN_TRY(need(c->vis2orig, (v_cap ? v_cap : 1) * 4));
  LUMEN_TRY(need(c->radius, (v_cap ? v_cap : 1) * 4));
  LUMEN_TRY(need(c->vmeta, (v_cap ? v_cap : 1) * 16));
  for (int b = 0; b < 2; ++b) {
    LUMEN_TRY(need(c->keys[b], (m_cap ? m_cap : 1) * 8));
    LUMEN_TRY(need(c->vals[b], (m_cap ? m_cap : 1) * 4));
  }
  return LUMEN_SUCCESS;
}

static LumenStatus ensure_capacity(LumenExContext* c, uint64_t v_cap, uint64_t m_cap) {
  return capacity_buffers(c, need_ensure, v_cap, m_cap);
}

// Key capacity a graph frame reserves over the instances of its first frame.
static uint64_t graph_m_cap(uint64_t m) { return m + m / 4 + 65536; }

// The frame's key bounds (depth key min, max, tile id min, max) sit right
// after the 8-byte totals, followed by the exact 64-bit instance total, so
// one 32-byte copy reads all three back.
static uint32_t* frame_bounds(LumenExContext* c) {
  return reinterpret_cast<uint32_t*>(c->totals.as<uint8_t>() + 8);
}

// Stage 1 of a frame: per-(view, Gaussian) counts and their scan.
static LumenStatus enqueue_count(LumenExContext* c, const SceneDev& sd, const FrameShape& fs,
                                 hipStream_t st) {
  if (sd.n == 0) {  // no k_plan this frame: an empty plan
    hipLaunchKernelGGL(k_init_frame, dim3(1), dim3(1), 0, st, frame_bounds(c),
                       c->plan.as<SortPlan>());
    LUMEN_LAUNCHED();
    return LUMEN_SUCCESS;
  }
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_COUNT, false, st));
  const uint32_t cb = count_block(fs);
  const uint32_t nblocks = div_up(sd.n, cb);
  uint4* bb = c->block_bounds.as<uint4>();
  uint64_t* binst = reinterpret_cast<uint64_t*>(bb + nblocks);
  InlineViews iv;
  memset(&iv, 0, sizeof(iv));
  uint32_t* etab = fs.fusable ? c->etab.as<uint32_t>() : nullptr;
  if (fs.inline_views) memcpy(iv.v, c->host_views.data(), sizeof(ViewParams) * fs.nviews);
  ViewParams* vout = fs.inline_views ? c->views.as<ViewParams>() : nullptr;
  auto count = fs.fusable ? (fs.inline_views ? k_count<true, kFusedBlock> : k_count<false, kFusedBlock>)
                          : (fs.inline_views ? k_count<true, 256> : k_count<false, 256>);
  const uint64_t nv = (uint64_t)fs.nviews * sd.n;
  // a single view's blocks are consecutive pieces of the scanned array: their
  // sums stand in for the scan's reduce pass
  const bool pieces = scan_pieces(fs, sd.n);
  hipLaunchKernelGGL(count, dim3(nblocks), dim3(cb), 0, st, sd, c->views.as<ViewParams>(),
                     (int)fs.nviews, c->packed.as<uint64_t>(), c->geo.as<GeoCache>(), bb, binst,
                     iv, vout, etab, pieces ? c->part.as<uint64_t>() : nullptr);
  LUMEN_LAUNCHED();
  if (pieces)
    LUMEN_TRY(exclusive_scan_pieces<uint64_t>(c->packed.as<uint64_t>(), c->packed.as<uint64_t>(),
                                              nv, c->part.as<uint64_t>(), cb,
                                              c->totals.as<uint64_t>(), st));
  else
    LUMEN_TRY(exclusive_scan<uint64_t>(c->packed.as<uint64_t>(), c->packed.as<uint64_t>(), nv,
                                       c->part.as<uint64_t>(), c->totals.as<uint64_t>(), st));
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_COUNT, true, st));
  return LUMEN_SUCCESS;
}

// Stage 2 of a frame: plan and emission, with the given capacities (exact
// after a host read-back, reserved in a graph, or the buffers' present size
// ahead of the read-back; k_plan and k_emit flag a frame that exceeds them).
// With read_back, k_plan also writes the frame's sizes to the context's
// mapped host words and c->ev_totals is recorded right after it.
static LumenStatus enqueue_emit(LumenExContext* c, const SceneDev& sd, const FrameShape& fs,
                                uint64_t v_cap, uint64_t m_cap, int force_items,
                                hipStream_t st, bool read_back = false) {
  SortPlan* plan = c->plan.as<SortPlan>();
  if (sd.n > 0) {
    const uint32_t nblocks = div_up(sd.n, count_block(fs));  // k_count's blocks
    const uint4* bb = c->block_bounds.as<uint4>();
    const uint64_t* binst = reinterpret_cast<const uint64_t*>(bb + nblocks);
    uint32_t pb = div_up(fs.tiles ? fs.tiles : 1, 1024 * 16);
    if (fs.fusable && pb < 1u + kDigitBins) pb = 1u + kDigitBins;  // the column scans
    hipLaunchKernelGGL(k_plan, dim3(pb), dim3(1024), 0, st, c->packed.as<uint64_t>(),
                       c->totals.as<uint64_t>(), bb, binst, nblocks, frame_bounds(c), sd.n,
                       (int)fs.nviews, m_cap, v_cap, plan, c->spans.as<ViewSpan>(),
                       c->segs.as<SortSeg>(), c->ranges.as<int2>(), (uint32_t)fs.tiles,
                       force_items, fs.fusable ? c->etab.as<uint32_t>() : nullptr,
                       fs.fusable ? c->epre.as<uint32_t>() : nullptr,
                       fs.fusable ? c->etot.as<uint32_t>() : nullptr, read_back ? c->ret_dev : nullptr);
    LUMEN_LAUNCHED();
    if (read_back) {
      if (!c->ev_totals) LUMEN_HIP(hipEventCreateWithFlags(&c->ev_totals, hipEventDisableTiming));
      LUMEN_HIP(hipEventRecord(c->ev_totals, st));
    }
  } else {
    LUMEN_TRY(fill<uint32_t>(c->spans.as<uint32_t>(), 0u, 4ull * fs.nviews, st));
    LUMEN_TRY(fill<int32_t>(c->ranges.as<int32_t>(), -1, 2ull * fs.tiles, st));
  }
  if (sd.n > 0 && m_cap > 0) {
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_EMIT, false, st));
    // (a frame that cannot run the first pass here keeps the SH coefficients
    // across views whatever their number: measured faster at one view too)
    const uint32_t eb = count_block(fs);
    auto emit = fs.fusable ? k_emit<false, true, kFusedBlock> : k_emit<true, false, 256>;
    hipLaunchKernelGGL(emit, dim3(div_up(sd.n, eb)), dim3(eb), 0, st, sd,
                       c->views.as<ViewParams>(), (int)fs.nviews, c->packed.as<uint64_t>(),
                       c->totals.as<uint64_t>(), c->geo.as<GeoCache>(), m_cap,
                       v_cap, c->rec.as<float4>(), c->vis2orig.as<int32_t>(),
                       c->radius.as<int32_t>(), c->vmeta.as<uint4>(), c->keys[0].as<uint64_t>(),
                       c->vals[0].as<uint32_t>(), plan,
                       fs.fusable ? c->epre.as<uint32_t>() : nullptr,
                       fs.fusable ? c->etot.as<uint32_t>() : nullptr, c->keys[1].as<uint64_t>(),
                       c->vals[1].as<uint32_t>());
    LUMEN_LAUNCHED();
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_EMIT, true, st));
  }
  return LUMEN_SUCCESS;
}

// Stage 3 of a frame: sort, ranges and compositing of at most m_cap keys.
static LumenStatus enqueue_finish(LumenExContext* c, const SceneDev& sd, const FrameShape& fs,
                                  const LumenExRenderConfig* cfg, uint64_t m_cap, int sort_first,
                                  int sort_slots, int sort_modes, int items, bool keep,
                                  float* out_rgb, float* out_depth, hipStream_t st) {
  SortPlan* plan = c->plan.as<SortPlan>();
  if (sd.n > 0 && m_cap > 0) {
    if (keep) {
      LUMEN_TRY(ensure(c->keys_keep, m_cap * 8));
      LUMEN_TRY(ensure(c->vals_keep, m_cap * 4));
      LUMEN_HIP(hipMemcpyAsync(c->keys_keep.ptr, c->keys[0].ptr, m_cap * 8,
                               hipMemcpyDeviceToDevice, st));
      LUMEN_HIP(hipMemcpyAsync(c->vals_keep.ptr, c->vals[0].ptr, m_cap * 4,
                               hipMemcpyDeviceToDevice, st));
    }
    uint64_t* kb[2] = {c->keys[0].as<uint64_t>(), c->keys[1].as<uint64_t>()};
    uint32_t* vb[2] = {c->vals[0].as<uint32_t>(), c->vals[1].as<uint32_t>()};
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_SORT, false, st));
    LUMEN_TRY(radix_sort_planned(kb, vb, plan, c->segs.as<SortSeg>(), (int)fs.nviews, m_cap,
                                 sort_first, sort_slots, sort_modes, items,
                                 SortScratch{&c->hist, &c->hist_part}, st));
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_SORT, true, st));
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_RANGES, false, st));
    hipLaunchKernelGGL(k_ranges, dim3(div_up(m_cap, kRangeKeys)), dim3(256), 0, st, kb[0], kb[1], plan,
                       c->segs.as<SortSeg>(), (int)fs.nviews, c->views.as<ViewParams>(),
                       c->ranges.as<int2>(), (uint32_t)fs.tiles);
    LUMEN_LAUNCHED();
    LUMEN_TRY(mark(c, LUMEN_EX_TIME_RANGES, true, st));
  }
  const bool save = (cfg->flags & LUMEN_EX_SAVE_PIXELS) != 0;
  float* tf = save ? c->t_final.as<float>() : nullptr;
  int32_t* ne = save ? c->n_eval.as<int32_t>() : nullptr;
  // default LDS batch: 128 records (two pixels per lane keep 2 waves per
  // block resident; 128 leaves room for 8 waves per SIMD), 256 for one pixel
  const uint32_t tpix = cfg->tile_width * cfg->tile_height;
  const uint32_t batch = cfg->batch_depth ? cfg->batch_depth : (tpix % 128 == 0 ? 128u : 256u);
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_COMPOSITE, false, st));
  LUMEN_TRY(composite_launch(batch, fs.tiles, cfg->tile_width, cfg->tile_height, st, c,
                             fs.nviews, out_rgb, out_depth, tf, ne));
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_COMPOSITE, true, st));
  return LUMEN_SUCCESS;
}

// Stages 2 and 3 with one set of capacities (the graph frame).
static LumenStatus enqueue_rest(LumenExContext* c, const SceneDev& sd, const FrameShape& fs,
                                const LumenExRenderConfig* cfg, uint64_t v_cap, uint64_t m_cap,
                                int sort_slots, int sort_modes, bool keep, float* out_rgb,
                                float* out_depth, hipStream_t st) {
  const int items = sort_items_for(m_cap);
  LUMEN_TRY(enqueue_emit(c, sd, fs, v_cap, m_cap, items, st));
  return enqueue_finish(c, sd, fs, cfg, m_cap, 0, sort_slots, sort_modes, items, keep, out_rgb,
                        out_depth, st);
}

// The visible-splat and key capacities the buffers hold now (0 before a
// frame has sized them).
static void current_capacity(const LumenExContext* c, uint64_t& v_cap, uint64_t& m_cap) {
  auto cap = [](const DevBuf& b, uint64_t unit) -> uint64_t { return b.ptr ? b.cap / unit : 0; };
  v_cap = cap(c->rec, sizeof(SplatRecord));
  v_cap = std::min(v_cap, cap(c->vis2orig, 4));
  v_cap = std::min(v_cap, cap(c->radius, 4));
  v_cap = std::min(v_cap, cap(c->vmeta, 16));
  m_cap = cap(c->keys[0], 8);
  m_cap = std::min(m_cap, cap(c->keys[1], 8));
  m_cap = std::min(m_cap, cap(c->vals[0], 4));
  m_cap = std::min(m_cap, cap(c->vals[1], 4));
}

// Sort passes that can ever be needed for this frame shape: depth keys are
// at most trunc(far * 1000) (culling keeps depth <= far), frame tile ids at
// most tiles - 1.
static int max_sort_slots(const LumenExContext* c, const FrameShape& fs) {
  float far_max = 0.0f;
  for (const ViewParams& v : c->host_views) far_max = v.far_plane > far_max ? v.far_plane : far_max;
  const uint32_t dmax = depth_key(far_max);
  int32_t dg[8];
  int n = varying_digits(0, dmax, 0, 4, dg);
  uint32_t tiles = 0;  // keys hold view-local tiles
  for (const ViewParams& v : c->host_views) {
    const uint32_t t = (uint32_t)v.tiles_x * (uint32_t)v.tiles_y;
    tiles = t > tiles ? t : tiles;
  }
  (void)fs;
  n += varying_digits(0, tiles ? tiles - 1 : 0, 4, 4, dg + n);
  return n;
}

// The plain frame: count, read back the totals, size, render.
static LumenStatus render_plain(LumenExContext* c, const LumenExScene* sc,
                                const LumenExCamera* cams, uint32_t nviews,
                                const LumenExRenderConfig* cfg, float* out_rgb, float* out_depth,
                                int* passes_out, hipStream_t st) {
  // the scan's visible half counts (view, Gaussian) pairs in 32 bits
  if ((uint64_t)nviews * sc->count >= 0xFFFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "%u views x %llu Gaussians exceed 2^32 - 1", nviews,
                (unsigned long long)sc->count);
  // until this frame completes, the last frame's stages and sizes are gone
  c->sizes_known = false;
  c->kept = false;
  c->pixels = false;
  c->has_grad = false;
  FrameShape fs;
  LUMEN_TRY(upload_views(c, cams, nviews, cfg, fs, st, /*allow_inline=*/sc->count > 0));
  const bool save = (cfg->flags & LUMEN_EX_SAVE_PIXELS) != 0;
  const bool keep_flag = (cfg->flags & LUMEN_EX_KEEP_STAGES) != 0;
  fs.fusable = frame_fusable(nviews, fs.tiles, sc->count, keep_flag);
  LUMEN_TRY(ensure_frame_buffers(c, sc->count, fs, save));
  const SceneDev sd = scene_dev(sc);
  for (bool& r : c->stage_ran) r = false;
  LUMEN_TRY(enqueue_count(c, sd, fs, st));
  uint64_t V = 0, M = 0;
  uint32_t hb[4] = {0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u};
  uint64_t v_now = 0, m_now = 0;
  bool ahead = false;  // plan and emission enqueued before the read-back
  if (sd.n > 0) {
    // Plan and emission go ahead with the capacity the buffers already hold,
    // so the device works on them while the host waits for the sizes k_plan
    // writes to mapped host memory (an event after k_plan, not the stream);
    // a frame that outgrows them is flagged by k_plan / k_emit and both run
    // again below.  Without capacity only k_plan runs ahead.
    current_capacity(c, v_now, m_now);
    ahead = v_now > 0 && m_now > 0;
    if (!ahead) v_now = m_now = 0;
    LUMEN_TRY(enqueue_emit(c, sd, fs, v_now, m_now, 0, st, /*read_back=*/true));
    LUMEN_HIP(hipEventSynchronize(c->ev_totals));
    const volatile uint64_t* h = c->ret_host;
    const uint64_t tot = h[0];
    const uint64_t b01 = h[1], b23 = h[2];
    hb[0] = (uint32_t)b01;
    hb[1] = (uint32_t)(b01 >> 32);
    hb[2] = (uint32_t)b23;
    hb[3] = (uint32_t)(b23 >> 32);
    M = h[3];  // exact: the scan's count half is modulo 2^32
    V = tot & 0xFFFFFFFFull;
    if (M >= 0x7FFFFFFFull) {
      if (ahead) LUMEN_HIP(hipStreamSynchronize(st));
      return fail(LUMEN_EX_ERROR_CAPACITY, "%llu tile instances exceed 2^31-1",
                  (unsigned long long)M);
    }
  }
  if (!(ahead && V <= v_now && M <= m_now)) {
    // the buffers may be replaced: the emission enqueued above must be done
    if (ahead) LUMEN_HIP(hipStreamSynchronize(st));
    LUMEN_TRY(ensure_capacity(c, V, M));
    LUMEN_TRY(enqueue_emit(c, sd, fs, V, M, 0, st));
  }
  // the plan k_plan will write, known here from the bounds: only its passes
  // and key width are enqueued
  SortPlan hp;
  plan_frame(V > 0, hb[0], hb[1], hb[2], hb[3], fs.fusable, hp);
  const int np = hp.passes;
  const bool keep = keep_flag && M > 0;
  LUMEN_TRY(enqueue_finish(c, sd, fs, cfg, M, hp.first, np, hp.mode == 1 ? 2 : 1,
                           sort_items_for(M), keep, out_rgb, out_depth, st));
  c->V = V;
  c->M = M;
  c->T = fs.tiles;
  c->P = fs.pixels;
  c->N = sc->count;
  c->nviews = nviews;
  c->tile_w = cfg->tile_width;
  c->tile_h = cfg->tile_height;
  c->v_cap = V;
  c->m_cap = M;
  c->key_bytes = hp.mode == 1 ? 4u : 8u;
  c->emit_passes = (uint32_t)hp.first;
  c->sizes_known = true;
  c->kept = keep;
  c->pixels = save;
  c->has_grad = false;
  if (passes_out) *passes_out = np;
  return LUMEN_SUCCESS;
}

// The reserved frame: the whole frame enqueued with the capacities of the
// last reservation (all the buffers hold, with none), the sizes and the sort
// plan left on the device (as a graph replay does, without capture): the host
// returns at once.  The sort's and the ranges' grids cover the key capacity,
// so a tight reservation keeps a small frame's launches small in a context
// that larger frames have grown.
static LumenStatus render_reserved(LumenExContext* c, const LumenExScene* sc,
                                   const LumenExCamera* cams, uint32_t nviews,
                                   const LumenExRenderConfig* cfg, float* out_rgb,
                                   float* out_depth, hipStream_t st) {
  uint64_t v_now = 0, m_now = 0;
  current_capacity(c, v_now, m_now);
  if (c->reserved_set) {
    v_now = std::min(v_now, c->reserved_v);
    m_now = std::min(m_now, c->reserved_m);
  }
  if (v_now == 0 || m_now == 0)  // nothing sized yet: a plain frame sizes the buffers
    return render_plain(c, sc, cams, nviews, cfg, out_rgb, out_depth, nullptr, st);
  if ((uint64_t)nviews * sc->count >= 0xFFFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "%u views x %llu Gaussians exceed 2^32 - 1", nviews,
                (unsigned long long)sc->count);
  c->sizes_known = false;
  c->kept = false;
  c->pixels = false;
  c->has_grad = false;
  FrameShape fs;
  LUMEN_TRY(upload_views(c, cams, nviews, cfg, fs, st, /*allow_inline=*/sc->count > 0));
  const bool save = (cfg->flags & LUMEN_EX_SAVE_PIXELS) != 0;
  fs.fusable = frame_fusable(nviews, fs.tiles, sc->count, false);
  LUMEN_TRY(ensure_frame_buffers(c, sc->count, fs, save));
  const int items = sort_items_for(m_now);
  LUMEN_TRY(ensure(c->hist, sort_hist_len(m_now, (int)nviews, items) * sizeof(uint32_t)));
  LUMEN_TRY(ensure(c->hist_part, sort_part_bytes(m_now, (int)nviews, items)));
  const SceneDev sd = scene_dev(sc);
  for (bool& r : c->stage_ran) r = false;
  LumenExRenderConfig pc = *cfg;
  pc.flags &= ~LUMEN_EX_KEEP_STAGES;
  LUMEN_TRY(enqueue_count(c, sd, fs, st));
  LUMEN_TRY(enqueue_rest(c, sd, fs, &pc, v_now, m_now, max_sort_slots(c, fs), 3, false, out_rgb,
                         out_depth, st));
  c->T = fs.tiles;
  c->P = fs.pixels;
  c->N = sc->count;
  c->nviews = nviews;
  c->tile_w = cfg->tile_width;
  c->tile_h = cfg->tile_height;
  c->v_cap = v_now;
  c->m_cap = m_now;
  c->pixels = save;
  return LUMEN_SUCCESS;
}

static void graph_release(GraphCache& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
  g.valid = false;
  g.key.clear();
}

static std::vector<uint8_t> graph_key(const LumenExScene* sc, const LumenExCamera* cams,
                                      uint32_t n, const LumenExRenderConfig* cfg, float* rgb,
                                      float* depth, hipStream_t st) {
  std::vector<uint8_t> k;
  auto put = [&k](const void* p, size_t b) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    k.insert(k.end(), q, q + b);
  };
  put(sc, sizeof(*sc));
  put(&n, sizeof(n));
  put(cams, sizeof(LumenExCamera) * n);
  put(cfg, sizeof(*cfg));
  put(&rgb, sizeof(rgb));
  put(&depth, sizeof(depth));
  put(&st, sizeof(st));
  return k;
}

// Graph frame: on a new key, render once plainly (which also sizes the
// buffers), reserve capacity, capture the whole frame; replay afterwards.
static LumenStatus render_graph(LumenExContext* c, const LumenExScene* sc,
                                const LumenExCamera* cams, uint32_t nviews,
                                const LumenExRenderConfig* cfg, float* out_rgb, float* out_depth,
                                hipStream_t st) {
  if (st == nullptr)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "LUMEN_EX_GRAPH needs a non-default stream");
  std::vector<uint8_t> key = graph_key(sc, cams, nviews, cfg, out_rgb, out_depth, st);
  if (c->graph.valid && key == c->graph.key) {
    for (bool& r : c->stage_ran) r = false;
    LUMEN_HIP(hipGraphLaunch(c->graph.exec, st));
    c->sizes_known = false;
    c->kept = false;
    c->pixels = (cfg->flags & LUMEN_EX_SAVE_PIXELS) != 0;
    c->has_grad = false;
    return LUMEN_SUCCESS;
  }
  graph_release(c->graph);
  LumenExRenderConfig pc = *cfg;
  pc.flags &= ~LUMEN_EX_KEEP_STAGES;
  LUMEN_TRY(render_plain(c, sc, cams, nviews, &pc, out_rgb, out_depth, nullptr, st));
  const uint64_t m_cap = graph_m_cap(c->M);
  const uint64_t v_cap = (uint64_t)nviews * sc->count;
  LUMEN_TRY(ensure_capacity(c, v_cap, m_cap));
  FrameShape fs;
  LUMEN_TRY(upload_views(c, cams, nviews, cfg, fs, st));
  fs.fusable = frame_fusable(nviews, fs.tiles, sc->count, false);
  LUMEN_TRY(ensure_frame_buffers(c, sc->count, fs, (cfg->flags & LUMEN_EX_SAVE_PIXELS) != 0));
  // sort scratch for the capacity, allocated before capture
  const int items = sort_items_for(m_cap);
  LUMEN_TRY(ensure(c->hist, sort_hist_len(m_cap, (int)nviews, items) * sizeof(uint32_t)));
  LUMEN_TRY(ensure(c->hist_part, sort_part_bytes(m_cap, (int)nviews, items)));
  LUMEN_HIP(hipStreamSynchronize(st));
  const int slots = max_sort_slots(c, fs);
  const SceneDev sd = scene_dev(sc);
  const bool prof = c->profile;
  c->profile = false;
  LUMEN_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  LumenStatus s1 = enqueue_count(c, sd, fs, st);
  LumenStatus s2 = s1 == LUMEN_SUCCESS
                       ? enqueue_rest(c, sd, fs, &pc, v_cap, m_cap, slots, 3, false, out_rgb,
                                      out_depth, st)
                       : s1;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(st, &g);
  c->profile = prof;
  if (s2 != LUMEN_SUCCESS) {
    if (g) (void)hipGraphDestroy(g);
    return s2;
  }
  if (ec != hipSuccess) return fail(LUMEN_EX_ERROR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(LUMEN_EX_ERROR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  }
  c->graph.graph = g;
  c->graph.exec = ex;
  c->graph.key = key;
  c->graph.valid = true;
  c->graph.stream = st;
  c->graph.m_cap = m_cap;
  c->v_cap = v_cap;
  c->m_cap = m_cap;
  return LUMEN_SUCCESS;
}

// ---- workspace ------------------------------------------------------------

// Every device buffer of a context.
static std::vector<DevBuf*> context_buffers(LumenExContext* c) {
  return {&c->packed, &c->part,     &c->totals,    &c->plan,      &c->views,    &c->spans,
          &c->geo,    &c->block_bounds, &c->rec,   &c->vis2orig,  &c->radius,   &c->vmeta,
          &c->keys[0], &c->keys[1], &c->vals[0],   &c->vals[1],   &c->keys_keep, &c->vals_keep,
          &c->hist,   &c->hist_part, &c->etab, &c->epre, &c->etot, &c->ranges, &c->tile_order,   &c->t_final,   &c->n_eval,   &c->minmax,
          &c->partial, &c->grad_vis, &c->sum_seg, &c->sum_aux, &c->loss_part, &c->dens_cls,
          &c->dens_cnt, &c->dens_off, &c->dens_misc, &c->segs};
}

struct WorkspaceShape {
  FrameShape fs;
  uint32_t max_w = 0, max_h = 0;  // the largest view (the loss's partial sums)
};

static LumenStatus workspace_shape(const LumenExCamera* cams, uint32_t n,
                                   const LumenExRenderConfig* cfg, WorkspaceShape& ws) {
  uint64_t tb = 0, pb = 0;
  for (uint32_t k = 0; k < n; ++k) {
    const uint64_t tx = (cams[k].width + cfg->tile_width - 1) / cfg->tile_width;
    const uint64_t ty = (cams[k].height + cfg->tile_height - 1) / cfg->tile_height;
    tb += tx * ty;
    pb += (uint64_t)cams[k].width * cams[k].height;
    if ((uint64_t)cams[k].width * cams[k].height > (uint64_t)ws.max_w * ws.max_h) {
      ws.max_w = cams[k].width;
      ws.max_h = cams[k].height;
    }
  }
  if (tb >= 0x7FFFFFFFull) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "too many tiles in the frame");
  ws.fs.nviews = n;
  ws.fs.tiles = tb;
  ws.fs.pixels = pb;
  return LUMEN_SUCCESS;
}

// Every buffer that frames of this shape use with at most `max_instances`
// tile instances (and the backward and its loss when `backward`), named to
// `need`: the sizes render_plain, render_graph, the sort and the backward
// ask for, at their largest.
template <class Need>
static LumenStatus workspace(LumenExContext* c, Need&& need, uint64_t N, const WorkspaceShape& ws,
                             const LumenExRenderConfig* cfg, uint64_t max_instances,
                             bool backward) {
  const uint32_t flags = cfg->flags;
  const uint64_t v_cap = (uint64_t)ws.fs.nviews * N;
  const uint64_t m_cap = (flags & LUMEN_EX_GRAPH) ? graph_m_cap(max_instances) : max_instances;
  const bool save = backward || (flags & LUMEN_EX_SAVE_PIXELS);
  LUMEN_TRY(need(c->views, sizeof(ViewParams) * ws.fs.nviews));
  LUMEN_TRY(need(c->spans, sizeof(ViewSpan) * ws.fs.nviews));
  LUMEN_TRY(need(c->segs, sizeof(SortSeg) * ws.fs.nviews));
  FrameShape fs = ws.fs;
  fs.fusable = frame_fusable(fs.nviews, fs.tiles, N, (flags & LUMEN_EX_KEEP_STAGES) != 0);
  LUMEN_TRY(frame_buffers(c, need, N, fs, save));
  LUMEN_TRY(capacity_buffers(c, need, v_cap, m_cap));
  if (m_cap > 0) {
    // the larger table of the two tile sizes a frame of up to m_cap keys may use
    LUMEN_TRY(need(c->hist, sort_hist_len(m_cap, (int)ws.fs.nviews, 12) * sizeof(uint32_t)));
    LUMEN_TRY(need(c->hist_part, sort_part_bytes(m_cap, (int)ws.fs.nviews, 12)));
  }
  if ((flags & LUMEN_EX_KEEP_STAGES) && m_cap > 0) {
    LUMEN_TRY(need(c->keys_keep, m_cap * 8));
    LUMEN_TRY(need(c->vals_keep, m_cap * 4));
  }
  if (backward) {
    LUMEN_TRY(need(c->partial, partial_bytes(m_cap)));
    LUMEN_TRY(need(c->grad_vis, grad_vis_bytes(v_cap)));
    LUMEN_TRY(need(c->sum_seg, sum_seg_bytes(m_cap)));
    LUMEN_TRY(need(c->sum_aux, sum_aux_bytes(m_cap)));
    LUMEN_TRY(need(c->loss_part, loss_part_bytes(ws.max_w, ws.max_h)));
  }
  return LUMEN_SUCCESS;
}

static LumenStatus check_workspace_args(const LumenExCamera* cams, uint32_t n,
                                        const LumenExRenderConfig* cfg, const char* who) {
  if (!cams || !cfg || n == 0) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "%s: null argument", who);
  LUMEN_TRY(check_config(cfg));
  LUMEN_TRY(check_cameras(cams, n));
  return LUMEN_SUCCESS;
}

extern "C" {

int lumen_ex_abi_version(void) { return LUMEN_EX_ABI_VERSION; }

const char* lumen_ex_last_error(void) { return lumen::last_error(); }

LumenExContext* lumen_ex_context_create(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    set_error("lumen_ex_context_create: device %d not available (%d devices)", device, count);
    return nullptr;
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_error("lumen_ex_context_create: hipSetDevice(%d) failed", device);
    return nullptr;
  }
  LumenExContext* c = new LumenExContext();
  c->device = device;
  if (hipHostMalloc(&c->pinned, 256, hipHostMallocDefault) != hipSuccess) {
    set_error("lumen_ex_context_create: hipHostMalloc failed");
    delete c;
    return nullptr;
  }
  void* ret = nullptr;
  if (hipHostMalloc(&ret, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&c->ret_dev), ret, 0) != hipSuccess) {
    if (ret) (void)hipHostFree(ret);
    (void)hipHostFree(c->pinned);
    set_error("lumen_ex_context_create: mapped host memory unavailable");
    delete c;
    return nullptr;
  }
  c->ret_host = static_cast<uint64_t*>(ret);
  return c;
}

void lumen_ex_context_destroy(LumenExContext* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  graph_release(c->graph);
  if (c->events_ready)
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->ev_totals) (void)hipEventDestroy(c->ev_totals);
  if (c->ev_views) (void)hipEventDestroy(c->ev_views);
  for (DevBuf* b : context_buffers(c)) release(*b);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->ret_host) (void)hipHostFree(c->ret_host);
  if (c->pinned_views) (void)hipHostFree(c->pinned_views);
  delete c;
}

LumenStatus lumen_ex_workspace_size(const LumenExCamera* cameras, uint32_t n_views,
                                    const LumenExRenderConfig* config, uint64_t n_gaussians,
                                    uint64_t max_instances, int with_backward,
                                    size_t* out_bytes) {
  clear_error();
  LUMEN_TRY(check_workspace_args(cameras, n_views, config, "lumen_ex_workspace_size"));
  if (!out_bytes) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_workspace_size: null argument");
  WorkspaceShape ws;
  LUMEN_TRY(workspace_shape(cameras, n_views, config, ws));
  // a context that is never given to the device: only its buffers' names
  LumenExContext sizing;
  std::vector<std::pair<DevBuf*, size_t>> asked;
  auto count = [&asked](DevBuf& b, size_t bytes) -> LumenStatus {
    for (auto& a : asked)
      if (a.first == &b) {
        a.second = bytes > a.second ? bytes : a.second;
        return LUMEN_SUCCESS;
      }
    asked.emplace_back(&b, bytes);
    return LUMEN_SUCCESS;
  };
  LUMEN_TRY(workspace(&sizing, count, n_gaussians, ws, config, max_instances, with_backward != 0));
  size_t total = 0;
  for (auto& a : asked) total += alloc_size(a.second);
  *out_bytes = total;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_context_reserve(LumenExContext* c, const LumenExCamera* cameras,
                                     uint32_t n_views, const LumenExRenderConfig* config,
                                     uint64_t n_gaussians, uint64_t max_instances,
                                     int with_backward) {
  clear_error();
  if (!c) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_context_reserve: null argument");
  LUMEN_TRY(check_workspace_args(cameras, n_views, config, "lumen_ex_context_reserve"));
  WorkspaceShape ws;
  LUMEN_TRY(workspace_shape(cameras, n_views, config, ws));
  LUMEN_HIP(hipSetDevice(c->device));
  if (c->graph.valid) graph_release(c->graph);  // a growing buffer would move under it
  LUMEN_TRY(workspace(c, need_ensure, n_gaussians, ws, config, max_instances, with_backward != 0));
  // the capacities reserved frames use from now on (workspace() sizes the
  // buffers for exactly these)
  c->reserved_set = true;
  c->reserved_v = (uint64_t)ws.fs.nviews * n_gaussians;
  c->reserved_m = (config->flags & LUMEN_EX_GRAPH) ? graph_m_cap(max_instances) : max_instances;
  // the host staging of the per-view parameters
  const size_t vbytes = sizeof(ViewParams) * n_views;
  if (c->pinned_views_cap < vbytes) {
    if (c->ev_views_pending) {
      LUMEN_HIP(hipEventSynchronize(c->ev_views));
      c->ev_views_pending = false;
    }
    if (c->pinned_views) LUMEN_HIP(hipHostFree(c->pinned_views));
    c->pinned_views = nullptr;
    c->pinned_views_cap = 0;
    LUMEN_HIP(hipHostMalloc(&c->pinned_views, vbytes, hipHostMallocDefault));
    c->pinned_views_cap = vbytes;
  }
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_context_workspace(LumenExContext* c, size_t* held_bytes) {
  clear_error();
  if (!c || !held_bytes)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_context_workspace: null argument");
  size_t total = 0;
  for (DevBuf* b : context_buffers(c)) total += b->ptr ? b->cap : 0;
  *held_bytes = total;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_render(LumenExContext* c, const LumenExScene* sc, const LumenExCamera* cam,
                            const LumenExRenderConfig* cfg, float* out_rgb, float* out_depth,
                            LumenExRenderStats* stats, void* stream) {
  clear_error();
  if (!c || !sc || !cam || !cfg || !out_rgb)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render: null argument");
  LUMEN_TRY(check_config(cfg));
  LUMEN_TRY(check_scene(sc));
  LUMEN_TRY(check_cameras(cam, 1));
  LUMEN_HIP(hipSetDevice(c->device));
  if (c->graph.valid) graph_release(c->graph);  // its buffers may be resized below
  int passes = 0;
  LUMEN_TRY(render_plain(c, sc, cam, 1, cfg, out_rgb, out_depth, &passes, (hipStream_t)stream));
  if (stats) {
    stats->visible_splats = c->V;
    stats->total_splat_instances = c->M;
    stats->tiles_x = (uint64_t)c->host_views[0].tiles_x;
    stats->tiles_y = (uint64_t)c->host_views[0].tiles_y;
    stats->sort_passes = (uint32_t)passes;
    stats->key_bytes = c->key_bytes;
    stats->emit_passes = c->emit_passes;
    stats->reserved = 0;
  }
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_render_batch(LumenExContext* c, const LumenExScene* sc,
                                  const LumenExCamera* cams, uint32_t n_views,
                                  const LumenExRenderConfig* cfg, float* out_rgb,
                                  float* out_depth, LumenExViewStats* out_stats, void* stream) {
  clear_error();
  if (!c || !sc || !cams || !cfg || !out_rgb)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render_batch: null argument");
  LUMEN_TRY(check_config(cfg));
  LUMEN_TRY(check_scene(sc));
  LUMEN_TRY(check_cameras(cams, n_views));
  LUMEN_HIP(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  if (cfg->flags & LUMEN_EX_GRAPH) {
    LUMEN_TRY(render_graph(c, sc, cams, n_views, cfg, out_rgb, out_depth, st));
  } else {
    if (c->graph.valid) graph_release(c->graph);
    if (cfg->flags & LUMEN_EX_RESERVED)
      LUMEN_TRY(render_reserved(c, sc, cams, n_views, cfg, out_rgb, out_depth, st));
    else
      LUMEN_TRY(render_plain(c, sc, cams, n_views, cfg, out_rgb, out_depth, nullptr, st));
  }
  if (out_stats) return lumen_ex_frame_status(c, out_stats, n_views, stream);
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_frame_status(LumenExContext* c, LumenExViewStats* out_stats, uint32_t n,
                                  void* stream) {
  clear_error();
  if (!c) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_frame_status: null context");
  hipStream_t st = (hipStream_t)stream;
  LUMEN_HIP(hipSetDevice(c->device));
  if (!c->plan.ptr) return fail(LUMEN_ERROR_UNINITIALIZED, "no frame rendered yet");
  SortPlan p;
  LUMEN_HIP(hipMemcpyAsync(c->pinned, c->plan.ptr, sizeof(SortPlan), hipMemcpyDeviceToHost, st));
  LUMEN_HIP(hipStreamSynchronize(st));
  memcpy(&p, c->pinned, sizeof(SortPlan));
  if (out_stats && n > 0) {
    if (n != c->nviews)
      return fail(LUMEN_ERROR_INVALID_ARGUMENT, "frame has %u views, asked for %u", c->nviews, n);
    LUMEN_HIP(hipMemcpy(out_stats, c->spans.ptr, sizeof(ViewSpan) * n, hipMemcpyDeviceToHost));
  }
  if (!c->sizes_known) {
    c->V = p.v;
    c->M = p.m;
    c->sizes_known = !p.overflow;
  }
  if (p.overflow)
    return fail(LUMEN_EX_ERROR_CAPACITY,
                "the frame exceeded its reserved key capacity (%llu); render it again without "
                "LUMEN_EX_GRAPH or with new cameras to re-plan",
                (unsigned long long)c->m_cap);
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_radix_sort(LumenExContext* c, uint64_t* keys, uint32_t* values, uint64_t n,
                                void* stream) {
  clear_error();
  if (!c || (n > 0 && (!keys || !values)))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_radix_sort: null argument");
  if (n >= 0xFFFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_radix_sort: n too large");
  if (n <= 1) return LUMEN_SUCCESS;
  hipStream_t st = (hipStream_t)stream;
  LUMEN_HIP(hipSetDevice(c->device));
  if (c->graph.valid) graph_release(c->graph);
  LUMEN_TRY(ensure(c->minmax, 16));
  LUMEN_TRY(ensure(c->plan, sizeof(SortPlan)));
  unsigned long long* mm = c->minmax.as<unsigned long long>();
  LUMEN_TRY(fill<unsigned long long>(mm, ~0ull, 1, st));
  LUMEN_TRY(fill<unsigned long long>(mm + 1, 0ull, 1, st));
  const uint32_t blocks = div_up(n, 256) < 2048 ? div_up(n, 256) : 2048;
  hipLaunchKernelGGL(k_minmax_u64, dim3(blocks), dim3(256), 0, st, keys, n, mm);
  LUMEN_LAUNCHED();
  uint64_t hmm[2];
  LUMEN_HIP(hipMemcpyAsync(c->pinned, mm, 16, hipMemcpyDeviceToHost, st));
  LUMEN_HIP(hipStreamSynchronize(st));
  memcpy(hmm, c->pinned, 16);
  int32_t digits[8];
  const int nd = varying_digits(hmm[0], hmm[1], 0, 8, digits);
  LUMEN_TRY(ensure(c->keys[1], n * 8));
  LUMEN_TRY(ensure(c->vals[1], n * 4));
  LUMEN_TRY(ensure(c->minmax, 16 + 32));
  LUMEN_TRY(ensure(c->segs, sizeof(SortSeg)));
  int32_t* ddig = reinterpret_cast<int32_t*>(static_cast<uint8_t*>(c->minmax.ptr) + 16);
  memcpy(static_cast<uint8_t*>(c->pinned) + 64, digits, sizeof(digits));
  LUMEN_HIP(hipMemcpyAsync(ddig, static_cast<uint8_t*>(c->pinned) + 64, sizeof(digits),
                           hipMemcpyHostToDevice, st));
  const int items = sort_items_for(n);
  hipLaunchKernelGGL(k_set_plan, dim3(1), dim3(1), 0, st, c->plan.as<SortPlan>(),
                     c->segs.as<SortSeg>(), nd, ddig, (uint32_t)n, items);
  LUMEN_LAUNCHED();
  uint64_t* kb[2] = {keys, c->keys[1].as<uint64_t>()};
  uint32_t* vb[2] = {values, c->vals[1].as<uint32_t>()};
  LUMEN_TRY(radix_sort_planned(kb, vb, c->plan.as<SortPlan>(), c->segs.as<SortSeg>(), 1, n, 0,
                               nd, 1, items, SortScratch{&c->hist, &c->hist_part}, st));
  if (nd & 1) {
    LUMEN_HIP(hipMemcpyAsync(keys, kb[1], n * 8, hipMemcpyDeviceToDevice, st));
    LUMEN_HIP(hipMemcpyAsync(values, vb[1], n * 4, hipMemcpyDeviceToDevice, st));
  }
  LUMEN_HIP(hipStreamSynchronize(st));  // the staging above is reused by the next call
  // the sort reused the frame's plan and ping-pong buffers: the last frame's
  // stages and its backward are gone
  c->sizes_known = false;
  c->M = 0;
  c->V = 0;
  c->kept = false;
  c->pixels = false;
  c->has_grad = false;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_build_ranges(const uint64_t* sorted_keys, uint64_t n, uint32_t tiles,
                                  int32_t* ranges, void* stream) {
  clear_error();
  if (!ranges || (n > 0 && !sorted_keys))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_build_ranges: null argument");
  if (n >= 0x7FFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_build_ranges: n too large");
  hipStream_t st = (hipStream_t)stream;
  LUMEN_TRY(fill<int32_t>(ranges, -1, 2ull * tiles, st));
  if (n > 0) {
    hipLaunchKernelGGL(k_ranges_plain, dim3(div_up(n, 256)), dim3(256), 0, st, sorted_keys,
                       (uint32_t)n, reinterpret_cast<int2*>(ranges), tiles);
    LUMEN_LAUNCHED();
  }
  return LUMEN_SUCCESS;
}

// The frame's c->M keys at `keys` (either width, view-local tiles) as the
// reference's 64-bit keys tile << 32 | depth key with frame tile ids, written
// to ping-pong buffer keys[into] (for the sorted keys the other buffer; for
// the kept unsorted keys, into = -1 takes whichever buffer does not hold the
// sorted result); *src points there.
static LumenStatus decode_frame_keys(LumenExContext* c, const void* keys, int into,
                                     const void** src) {
  SortPlan p;
  LUMEN_HIP(hipSetDevice(c->device));
  LUMEN_HIP(hipDeviceSynchronize());
  LUMEN_HIP(hipMemcpy(&p, c->plan.ptr, sizeof(p), hipMemcpyDeviceToHost));
  if (into < 0) into = 1 - (p.passes & 1);
  uint64_t* out = c->keys[into].as<uint64_t>();
  if (keys == out) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "decode_frame_keys: in place");
  hipLaunchKernelGGL(k_decode_keys, dim3(div_up(c->M, 256)), dim3(256), 0, nullptr, keys,
                     (uint32_t)c->M, c->plan.as<SortPlan>(), c->segs.as<SortSeg>(),
                     (int)c->nviews, c->views.as<ViewParams>(), out);
  LUMEN_LAUNCHED();
  LUMEN_HIP(hipDeviceSynchronize());
  *src = out;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_copy_stage(LumenExContext* c, int stage, void* dst, size_t bytes) {
  clear_error();
  if (!c || !dst) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_copy_stage: null argument");
  if (!c->sizes_known)
    return fail(LUMEN_ERROR_UNINITIALIZED,
                "frame sizes unknown (graph replay): call lumen_ex_frame_status first");
  const int fb = 0;
  const void* src = nullptr;
  size_t want = 0;
  switch (stage) {
    case LUMEN_EX_STAGE_VIS_TO_ORIG: src = c->vis2orig.ptr; want = c->V * 4; break;
    case LUMEN_EX_STAGE_RECORDS: {  // converted to the documented 48-byte layout
      want = c->V * 48;
      if (bytes != want) break;
      if (want == 0) return LUMEN_SUCCESS;
      DevBuf tmp;
      LUMEN_HIP(hipSetDevice(c->device));
      LUMEN_TRY(ensure(tmp, want));
      hipLaunchKernelGGL(k_records_out, dim3(div_up(c->V, 256)), dim3(256), 0, nullptr,
                         c->rec.as<float4>(), (uint32_t)c->V, tmp.as<float4>());
      const hipError_t e1 = hipGetLastError();
      const hipError_t e2 = e1 == hipSuccess ? hipDeviceSynchronize() : e1;
      const hipError_t e3 = e2 == hipSuccess ? hipMemcpy(dst, tmp.ptr, want, hipMemcpyDeviceToHost) : e2;
      release(tmp);
      if (e3 != hipSuccess)
        return fail(LUMEN_EX_ERROR_DEVICE, "lumen_ex_copy_stage: %s", hipGetErrorString(e3));
      return LUMEN_SUCCESS;
    }
    case LUMEN_EX_STAGE_RADIUS: src = c->radius.ptr; want = c->V * 4; break;
    case LUMEN_EX_STAGE_KEYS_UNSORTED:
      if (!c->kept && c->M) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_KEEP_STAGES)");
      src = c->keys_keep.ptr; want = c->M * 8;
      if (c->M) LUMEN_TRY(decode_frame_keys(c, c->keys_keep.ptr, -1, &src));
      break;
    case LUMEN_EX_STAGE_VALS_UNSORTED:
      if (!c->kept && c->M) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_KEEP_STAGES)");
      src = c->vals_keep.ptr; want = c->M * 4; break;
    case LUMEN_EX_STAGE_KEYS:
    case LUMEN_EX_STAGE_VALS: {
      // the sorted pairs sit in buffer (passes & 1)
      SortPlan p;
      if (c->M) {
        LUMEN_HIP(hipSetDevice(c->device));
        LUMEN_HIP(hipDeviceSynchronize());
        LUMEN_HIP(hipMemcpy(&p, c->plan.ptr, sizeof(p), hipMemcpyDeviceToHost));
      } else {
        p.passes = fb;
      }
      const int b = p.passes & 1;
      if (stage == LUMEN_EX_STAGE_KEYS) {
        src = c->keys[b].ptr;
        want = c->M * 8;
        if (c->M) LUMEN_TRY(decode_frame_keys(c, c->keys[b].ptr, 1 - b, &src));
      } else {
        src = c->vals[b].ptr;
        want = c->M * 4;
      }
      break;
    }
    case LUMEN_EX_STAGE_RANGES: src = c->ranges.ptr; want = c->T * 8; break;
    case LUMEN_EX_STAGE_T_FINAL:
      if (!c->pixels) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_SAVE_PIXELS)");
      src = c->t_final.ptr; want = c->P * 4; break;
    case LUMEN_EX_STAGE_N_EVAL:
      if (!c->pixels) return fail(LUMEN_ERROR_UNINITIALIZED, "stage not kept (LUMEN_EX_SAVE_PIXELS)");
      src = c->n_eval.ptr; want = c->P * 4; break;
    case LUMEN_EX_STAGE_GRAD_VIS:
      if (!c->has_grad) return fail(LUMEN_ERROR_UNINITIALIZED, "no backward since the last render");
      src = c->grad_vis.ptr; want = c->V * 40; break;
    default: return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_copy_stage: unknown stage %d", stage);
  }
  if (bytes != want)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_copy_stage: stage %d holds %zu bytes, got %zu",
                stage, want, bytes);
  if (want == 0) return LUMEN_SUCCESS;
  LUMEN_HIP(hipSetDevice(c->device));
  LUMEN_HIP(hipDeviceSynchronize());
  LUMEN_HIP(hipMemcpy(dst, src, want, hipMemcpyDeviceToHost));
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_set_profiling(LumenExContext* c, int enable) {
  clear_error();
  if (!c) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_set_profiling: null context");
  LUMEN_HIP(hipSetDevice(c->device));
  if (enable && !c->events_ready) {
    for (hipEvent_t& e : c->ev) LUMEN_HIP(hipEventCreate(&e));
    c->events_ready = true;
  }
  c->profile = enable != 0;
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_stage_times(LumenExContext* c, float* ms, int n) {
  clear_error();
  if (!c || !ms) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_stage_times: null argument");
  if (!c->events_ready) return fail(LUMEN_ERROR_UNINITIALIZED, "profiling was never enabled");
  for (int s = 0; s < n && s < LUMEN_EX_TIME_STAGES; ++s) {
    ms[s] = 0.0f;
    if (!c->stage_ran[s]) continue;
    LUMEN_HIP(hipEventSynchronize(c->ev[2 * s + 1]));
    LUMEN_HIP(hipEventElapsedTime(&ms[s], c->ev[2 * s], c->ev[2 * s + 1]));
  }
  return LUMEN_SUCCESS;
}

}  // extern "C"
