// backward.hip -- gradient of the splat rasterization path (lumen_ex_render_backward).
//
// The reference has no backward (src/optim/trainer.mojo:59-65 injects dummy
// gradients); the output layout is its GaussianGradients
// (src/optim/gradients.mojo:10-106).  Three kernels:
//
//   k_composite_bwd   one workgroup per tile, one lane per pixel, walking the
//                     tile's sorted list BACK to front from each pixel's last
//                     evaluated splat.  Alphas are recomputed exactly as the
//                     forward computed them (same expressions, same exp), so
//                     the skip / clamp / termination decisions are the
//                     forward's; transmittance before each blend is recovered
//                     from the saved final value (times 1/(1 - alpha), the
//                     hardware reciprocal).  A wave skips the splats past all
//                     of its pixels' last evaluated one and those whose 3-sigma
//                     box misses its pixels.  Per-splat gradients (mean2d,
//                     inverse covariance, opacity, rgb, depth) are summed over
//                     the wave (gfx950 lane swaps + DPP) into a per-wave LDS
//                     slot and over the waves at the end of each batch, then
//                     stored (no atomics) in the slot of that (splat, tile) instance in emission
//                     order: the splat's first key offset plus the tile's
//                     position in its rectangle, row-major.
//   k_sum_chunks      sums each visible splat's instance slots in a fixed
//   + k_sum_fixup     order (so the result does not depend on scheduling and
//                     is the same run to run), chunk by chunk through LDS.
//   k_preprocess_bwd  one thread per Gaussian, over every view it is visible
//                     in: adjoints of the SH colour (and view direction), the
//                     projection to screen, the 2D inverse and the EWA
//                     projection per view, then of the world covariance
//                     R S S^T R^T with its quaternion normalization once;
//                     writes the Gaussian's gradient planes (no atomics) and
//                     the view-space positional gradient norm for densification.
#include <hip/hip_runtime.h>
#include <string.h>

#include "common.hpp"
#include "context.hpp"
#include "lumen_ex.h"
#include "splat_math.hpp"

namespace lumen {


__device__ __forceinline__ int view_of_tile_b(const ViewParams* __restrict__ views, int nviews,
                                              uint32_t gt) {
  int lo = 0, hi = nviews - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (views[mid].tile_base <= gt) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Records, for every kSumChunk-slot chunk of the instance slots, the visible
// splat whose slots contain the chunk's first slot.
__global__ void k_chunk_seg(const SortPlan* __restrict__ plan, const uint4* __restrict__ vmeta,
                            uint32_t* __restrict__ chunk_seg) {
  if (plan->overflow) return;  // the frame was not rendered: its metadata is stale
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t V = plan->v, M = plan->m;
  for (uint64_t k = i0; k < V; k += stride) {
    const uint32_t q0 = vmeta[k].x;
    const uint32_t q1 = k + 1u < V ? vmeta[k + 1u].x : M;
    for (uint32_t c = (q0 + kSumChunk - 1u) / kSumChunk; c * kSumChunk < q1; ++c)
      chunk_seg[c] = (uint32_t)k;
  }
}

// DEPTH: a depth gradient is given (without it the depth channel's terms,
// one of the ten per-splat sums included, are left out: they are zero).
template <int BATCH, int PIX, bool DEPTH>
__global__ __launch_bounds__(256) void k_composite_bwd(
    const int2* __restrict__ ranges, const uint32_t* __restrict__ v0,
    const uint32_t* __restrict__ v1, const SortPlan* __restrict__ plan,
    const float4* __restrict__ rec, const ViewParams* __restrict__ views, int nviews,
    const ViewSpan* __restrict__ spans, const uint4* __restrict__ vmeta,
    const float* __restrict__ t_final, const int32_t* __restrict__ n_eval,
    const float* __restrict__ g_rgb, const float* __restrict__ g_depth,
    float* __restrict__ partial, const uint32_t* __restrict__ order) {
  static_assert(BATCH <= 256, "list entries are bytes");
  __shared__ float4 s0[BATCH], s1[BATCH], s2[BATCH];
  __shared__ uint32_t s_dst[BATCH];  // emission-order slot of each staged instance
  __shared__ uint8_t s_mask[BATCH];
  __shared__ uint8_t s_list[4 / PIX][BATCH];  // one list per wave
  __shared__ float s_part[4 / PIX][BATCH * kGradVis];  // per-wave sums, one slot per splat
  __shared__ int s_last;
  if (plan->overflow) return;
  // tiles in the forward's order (heaviest key-count class first, see
  // k_tile_order), each XCD (grid block i runs on XCD i % 8) taking runs of 8
  // consecutive positions: every XCD starts on the heavy tiles, and a run's
  // neighbouring tiles (a wave's tiles of one class keep their order) put
  // the slots of a splat covering them through one L2
  uint32_t gt;
  {
    const uint32_t n = gridDim.x, x = blockIdx.x % 8u, k = blockIdx.x / 8u;
    uint32_t pos = (k / 8u) * 64u + x * 8u + k % 8u;
    if (blockIdx.x >= (n / 64u) * 64u) pos = blockIdx.x;  // the last partial group: as is
    gt = order[pos];
  }
  const int vw = view_of_tile_b(views, nviews, gt);
  const ViewParams& v = views[vw];
  const uint32_t* __restrict__ vals = (plan->passes & 1) ? v1 : v0;
  const int nthreads = v.tile_w * v.tile_h / PIX;
  const int nwaves = nthreads >> 6;
  static_assert(PIX == 1, "one pixel per lane");
  // each wave covers a qw x qh block of the tile (8 x 8 in 16 x 16 tiles):
  // squarer than full-width strips, so fewer (splat, wave) pairs meet
  int qh = v.tile_h < 8 ? v.tile_h : 8;
  int qw = 64 / qh;
  if (qw > v.tile_w) {
    qw = v.tile_w;
    qh = 64 / qw;
  }
  const int nwx = v.tile_w / qw;
  const int tile = (int)(gt - v.tile_base);
  const int2 rg = ranges[gt];
  if (rg.x == -1 || rg.x >= rg.y) return;  // uniform across the block
  const bool white = v.white_background && spans[vw].instances > 0;
  const int tx = tile % v.tiles_x, ty = tile / v.tiles_x;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  const int px = tx * v.tile_w + (wv % nwx) * qw + lane % qw;
  const int py0 = ty * v.tile_h + (wv / nwx) * qh + lane / qw;
  for (int e = t; e < (4 / PIX) * BATCH * kGradVis; e += nthreads) (&s_part[0][0])[e] = 0.0f;
  if (t == 0) s_last = rg.x - 1;
  __syncthreads();
  int last[PIX];
  float T[PIX], gC0[PIX], gC1[PIX], gC2[PIX], gD[PIX];
  // Q = sum over channels of dL/dC_k times the colour (and depth) blended
  // behind the current splat, which is all the opacity gradient needs of it;
  // a white background is blended behind every splat (T_final per channel),
  // so Q starts at sum_k dL/dC_k T_final there
  float Q[PIX], pyc[PIX];
  int lane_last = rg.x - 1;
#pragma unroll
  for (int p = 0; p < PIX; ++p) {
    last[p] = rg.x - 1;
    T[p] = 1.0f;
    gC0[p] = gC1[p] = gC2[p] = gD[p] = 0.0f;
    Q[p] = 0.0f;
    pyc[p] = (float)(py0 + p) + 0.5f;
    if (px < v.width && py0 + p < v.height) {
      const size_t pix = v.pix_base + (size_t)(py0 + p) * v.width + px;
      last[p] = rg.x + n_eval[pix] - 1;
      T[p] = t_final[pix];
      gC0[p] = g_rgb[3 * pix + 0];
      gC1[p] = g_rgb[3 * pix + 1];
      gC2[p] = g_rgb[3 * pix + 2];
      gD[p] = (DEPTH && v.compute_depth) ? g_depth[pix] : 0.0f;
    }
    if (white) Q[p] = (gC0[p] + gC1[p] + gC2[p]) * T[p];
    lane_last = max(lane_last, last[p]);
  }
  // splats this wave's pixels evaluated: positions [0, wave_cnt) of the list
  const uint32_t wave_cnt = wave_max_u32_all((uint32_t)(lane_last - rg.x + 1));
  if (lane == 0) atomicMax(&s_last, rg.x + (int)wave_cnt - 1);
  __syncthreads();
  const int block_last = s_last;
  const float pxc = (float)px + 0.5f;
  // the slots of the instances past every pixel's last evaluated splat are
  // zero (the reverse pass never reaches them; it writes every other slot)
  for (int q = block_last + 1 + t; q < rg.y; q += nthreads) {
    const uint4 mt = vmeta[vals[q]];
    const uint32_t min_tx = mt.y & 0xFFFFu, min_ty = mt.y >> 16;
    const uint32_t wtiles = (mt.z & 0xFFFFu) - min_tx + 1u;
    const uint32_t dst = mt.x + ((uint32_t)ty - min_ty) * wtiles + ((uint32_t)tx - min_tx);
    float2* d2 = reinterpret_cast<float2*>(partial + (size_t)dst * kGradVis);
#pragma unroll
    for (int c = 0; c < kGradVis / 2; ++c) d2[c] = make_float2(0.0f, 0.0f);
  }
  for (int end = block_last + 1; end > rg.x; end -= BATCH) {
    const int start = max(rg.x, end - BATCH);
    const int n = end - start;
    for (int j = t; j < n; j += nthreads) {
      const uint32_t vi = vals[start + j];
      const uint4 mt = vmeta[vi];  // first key, rect min, rect max (tx | ty << 16)
      const uint32_t min_tx = mt.y & 0xFFFFu, min_ty = mt.y >> 16;
      const uint32_t wtiles = (mt.z & 0xFFFFu) - min_tx + 1u;
      s_dst[j] = mt.x + ((uint32_t)ty - min_ty) * wtiles + ((uint32_t)tx - min_tx);
      SplatRecord r;
      r.q0 = rec[4u * vi + 0u];
      r.q1 = rec[4u * vi + 1u];
      r.q2 = rec[4u * vi + 2u];
      r.q3 = rec[4u * vi + 3u];
      // the conic halved as the forward stages it, so the power below is the
      // forward's expression itself
      s0[j] = r.q0;  // x, y, i00/2, i01
      s1[j] = r.q1;  // i11/2, opacity, power limit
      s2[j] = r.q2;  // r, g, b, depth
      s_mask[j] = (uint8_t)record_mask(r, tx * v.tile_w, ty * v.tile_h, qw, qh, nwx, nwaves);
    }
    __syncthreads();
    // this wave's list: splats up to its pixels' last evaluated one whose box
    // meets its strip (the others have no gradient at its pixels), in order
    uint32_t cnt = 0;
    for (int c0 = 0; c0 < n; c0 += 64) {
      const int j = c0 + lane;
      const bool hit = j < n && (uint32_t)(start + j - rg.x) < wave_cnt &&
                       ((s_mask[j] >> wv) & 1u);
      const uint64_t bal = __ballot(hit);
      if (hit) s_list[wv][cnt + __popcll(bal & lanemask_lt())] = (uint8_t)j;
      cnt += __popcll(bal);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // each pixel's last evaluated splat as a position in this batch
    int lrel[PIX];
#pragma unroll
    for (int p = 0; p < PIX; ++p) lrel[p] = last[p] - start;
    for (int q = (int)cnt - 1; q >= 0; --q) {  // back to front; every lane runs every q
      const int j = s_list[wv][q];
      const float4 a = s0[j];  // x, y, i00/2, i01
      const float4 b = s1[j];  // i11/2, opacity, power limit
      const float4 c = s2[j];  // r, g, b, depth
      const float dx = pxc - a.x;
      const float txx = a.z * dx * dx;
      const float txy = a.w * dx;
      float gv[kGradVis];  // set by the first pixel, added to by the others
      // a blend weight of a pixel that blended this splat: never zero (alpha
      // >= 1/255, and T >= T_final, which the forward's termination keeps
      // >= 1e-4 x 0.01), so its compare is the wave's "any contribution"
      // test, taken straight from the weight's register
      float wany = 0.0f;
#pragma unroll
      for (int p = 0; p < PIX; ++p) {
        const float dy = pyc[p] - a.y;
        // the pixel's blend weight, opacity term and power term, zero where
        // the splat did not blend (the branches below only set these three,
        // so the ten gradient components are formed once, after them)
        float wm = 0.0f, g5 = 0.0f, dpm = 0.0f;
        float power = txx + txy * dy + b.x * dy * dy;
        const bool pclamp = power < 0.0f;
        if (pclamp) power = 0.0f;
        // splats up to the pixel's last evaluated one, within the power
        // limit (the forward skipped the others too); one branch, the
        // clamps' cases below are selects
        if (j <= lrel[p] && !(power > b.z)) {
          const float G = splat_exp_neg(-power);
          float alpha = b.y * G;
          const bool aclamp = alpha > 0.99f;
          if (aclamp) alpha = 0.99f;
          if (!(alpha < kMinAlpha)) {
            const float om = 1.0f - alpha;
            const float rom = __builtin_amdgcn_rcpf(om);
            const float Tj = T[p] * rom;
            const float w = alpha * Tj;
            // dL/dalpha = sum_k g_k (c_k Tj - S_k / (1 - alpha)), S_k the
            // channel blended behind; sum_k g_k S_k is carried as Q
            const float gc = DEPTH ? gC0[p] * c.x + gC1[p] * c.y + gC2[p] * c.z + gD[p] * c.w
                                   : gC0[p] * c.x + gC1[p] * c.y + gC2[p] * c.z;
            const float dal = gc * Tj - Q[p] * rom;
            Q[p] += gc * w;
            T[p] = Tj;
            wm = w;
            wany = w;
            const float g5v = dal * G;
            const float dpmv = -dal * b.y * G;
            g5 = aclamp ? 0.0f : g5v;
            dpm = (aclamp || pclamp) ? 0.0f : dpmv;
          }
        }
        // the power's gradient as moments of the pixel offset: sum dpm dx,
        // dpm dy, dpm dx^2, dpm dx dy, dpm dy^2 over the wave's pixels; the
        // mean and inverse-covariance gradients are formed from the sums once
        // per (splat, tile) (slot_gradients)
        const float m0 = dpm * dx, m1 = dpm * dy;
        const float t[kGradVis] = {m0,          m1,          m0 * dx,     m0 * dy,
                                   m1 * dy,     g5,          gC0[p] * wm, gC1[p] * wm,
                                   gC2[p] * wm, DEPTH ? gD[p] * wm : 0.0f};
#pragma unroll
        for (int k = 0; k < kGradVis; ++k) gv[k] = p == 0 ? t[k] : gv[k] + t[k];
      }
      if (__ballot(wany != 0.0f)) {  // wave-uniform; the whole wave is active here
        float* dst = &s_part[wv][j * kGradVis];
        if (DEPTH) {
          float sa, sb, sc;
          wave_sum10(gv, sa, sb, sc);
          if ((lane & 15) == 15) {
            const int r = lane >> 4;
            dst[sum10_slot(r)] = sa;
            dst[4 + sum10_slot(r)] = sb;
            if ((r & 1) == 0) dst[8 + (r >> 1)] = sc;
          }
        } else {  // nine sums (the depth slot stays zero)
          float sa, sb;
          wave_sum8(gv, sa, sb);
          const float sc = wave_sum_lane63(gv[8]);
          if ((lane & 15) == 15) {
            const int r = lane >> 4;
            dst[sum10_slot(r)] = sa;
            dst[4 + sum10_slot(r)] = sb;
            if (r == 3) dst[8] = sc;
          }
        }
      }
    }
    __syncthreads();
    for (int e = t; e < n * kGradVis; e += nthreads) {  // the waves' sums, into wave 0's slot
      float x = s_part[0][e];
      for (int w = 1; w < nwaves; ++w) {
        x += s_part[w][e];
        s_part[w][e] = 0.0f;
      }
      s_part[0][e] = x;
    }
    __syncthreads();
    for (int j = t; j < n; j += nthreads) {
      float* m = &s_part[0][j * kGradVis];
      const float i00 = 2.0f * s0[j].z, i01 = s0[j].w, i11 = 2.0f * s1[j].x;
      // dL/d(mean) = -Sigma^-1 sum dpm d, dL/d(i00, i01, i11) = (1/2, 1, 1/2) second moments
      const float g[kGradVis] = {-(i00 * m[0] + i01 * m[1]), -(i01 * m[0] + i11 * m[1]),
                                 0.5f * m[2], m[3], 0.5f * m[4], m[5], m[6], m[7], m[8], m[9]};
      float2* d2 = reinterpret_cast<float2*>(partial + (size_t)s_dst[j] * kGradVis);
#pragma unroll
      for (int c = 0; c < kGradVis / 2; ++c) d2[c] = make_float2(g[2 * c], g[2 * c + 1]);
#pragma unroll
      for (int c = 0; c < kGradVis; ++c) m[c] = 0.0f;
    }
    __syncthreads();
  }
}

// The per-splat sums of the instance slots, in a fixed order (so the result
// does not depend on scheduling).  A splat's slots are contiguous and the
// splats' runs follow each other in visible order, so the slots are cut into
// chunks of kSumChunk: one block per chunk loads it into LDS with coalesced
// reads, then one thread per splat meeting the chunk sums its slots there in
// order.  A splat inside the chunk is written out; the piece of a splat that
// continues past the chunk ("open") and the piece of one that began before it
// ("cont", also a whole chunk inside one splat) go to aux, and k_sum_fixup
// adds a crossing splat's pieces in chunk order.
// Portions of at most kSumLane slots are summed by one thread; longer ones
// (at most kSumChunk / (kSumLane + 1) per chunk) by a wave, lane l taking
// slots l, l + 64, ... and the wave folding its 64 sums with DPP.
constexpr uint32_t kSumLane = 32;

__device__ __forceinline__ void sum_store(const float (&g)[kGradVis], float* dst) {
  float2* d2 = reinterpret_cast<float2*>(dst);
#pragma unroll
  for (int c = 0; c < kGradVis / 2; ++c) d2[c] = make_float2(g[2 * c], g[2 * c + 1]);
}

__global__ __launch_bounds__(256) void k_sum_chunks(const uint4* __restrict__ vmeta,
                                                    const float* __restrict__ partial,
                                                    const SortPlan* __restrict__ plan,
                                                    const uint32_t* __restrict__ chunk_seg,
                                                    float* __restrict__ grad_vis,
                                                    float* __restrict__ aux) {
  __shared__ float2 s_rec[kSumChunk * kGradVis / 2];
  __shared__ uint32_t s_long[kSumChunk / (kSumLane + 1) + 1];
  __shared__ uint32_t s_nlong;
  if (plan->overflow) return;
  const uint32_t V = plan->v, M = plan->m;
  const uint32_t c0 = blockIdx.x * kSumChunk;
  if (c0 >= M || V == 0) return;  // uniform across the block
  const uint32_t c1 = min(c0 + kSumChunk, M);
  if (threadIdx.x == 0) s_nlong = 0;
  const float2* src = reinterpret_cast<const float2*>(partial + (size_t)c0 * kGradVis);
  const uint32_t n2 = (c1 - c0) * (kGradVis / 2);
  for (uint32_t e = threadIdx.x; e < n2; e += 256u) s_rec[e] = src[e];
  __syncthreads();
  // where a splat's portion of this chunk goes: the splat's sum when the
  // splat lies inside the chunk, else aux ("cont" if it began before, "open")
  auto dst_of = [&](uint32_t k, uint32_t q0, uint32_t q1) -> float* {
    return q0 >= c0 && q1 <= c1 ? grad_vis + (size_t)k * kGradVis
           : q0 < c0            ? aux + ((size_t)blockIdx.x * 2u + 1u) * kGradVis
                                : aux + (size_t)blockIdx.x * 2u * kGradVis;
  };
  const uint32_t seg0 = chunk_seg[blockIdx.x];
  for (uint32_t k = seg0 + threadIdx.x; k < V; k += 256u) {
    const uint32_t q0 = vmeta[k].x;
    if (q0 >= c1) break;
    const uint32_t q1 = k + 1u < V ? vmeta[k + 1u].x : M;
    const uint32_t a = max(q0, c0), b = min(q1, c1);
    if (b > a + kSumLane) {  // a wave sums it below
      s_long[atomicAdd(&s_nlong, 1u)] = k;
      continue;
    }
    float g[kGradVis];
#pragma unroll
    for (int c = 0; c < kGradVis; ++c) g[c] = 0.0f;
    for (uint32_t q = a; q < b; ++q) {
      const float2* r = &s_rec[(q - c0) * (kGradVis / 2)];
#pragma unroll
      for (int c = 0; c < kGradVis / 2; ++c) {
        const float2 x = r[c];
        g[2 * c] += x.x;
        g[2 * c + 1] += x.y;
      }
    }
    sum_store(g, dst_of(k, q0, q1));
  }
  __syncthreads();
  const uint32_t nlong = s_nlong;
  const int lane = threadIdx.x & 63;
  for (uint32_t i = threadIdx.x >> 6; i < nlong; i += 4u) {  // wave-uniform
    const uint32_t k = s_long[i];
    const uint32_t q0 = vmeta[k].x;
    const uint32_t q1 = k + 1u < V ? vmeta[k + 1u].x : M;
    const uint32_t a = max(q0, c0), b = min(q1, c1);
    float g[kGradVis];
#pragma unroll
    for (int c = 0; c < kGradVis; ++c) g[c] = 0.0f;
    for (uint32_t q = a + (uint32_t)lane; q < b; q += 64u) {
      const float2* r = &s_rec[(q - c0) * (kGradVis / 2)];
#pragma unroll
      for (int c = 0; c < kGradVis / 2; ++c) {
        const float2 x = r[c];
        g[2 * c] += x.x;
        g[2 * c + 1] += x.y;
      }
    }
#pragma unroll
    for (int c = 0; c < kGradVis; ++c)
      g[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_sum_lane63(g[c])), 63));
    if (lane == 0) sum_store(g, dst_of(k, q0, q1));
  }
}

// Splats whose slots cross a chunk boundary: open piece of the first chunk,
// then the cont pieces of the following ones, in order.
__global__ __launch_bounds__(256) void k_sum_fixup(const uint4* __restrict__ vmeta,
                                                   const SortPlan* __restrict__ plan,
                                                   const float* __restrict__ aux,
                                                   float* __restrict__ grad_vis) {
  if (plan->overflow) return;
  const uint32_t V = plan->v, M = plan->m;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= V) return;
  const uint32_t q0 = vmeta[k].x;
  const uint32_t q1 = k + 1u < V ? vmeta[k + 1u].x : M;
  if (q1 <= q0) {  // no instance slot (a chunk may not reach such a splat)
#pragma unroll
    for (int c = 0; c < kGradVis; ++c) grad_vis[(size_t)k * kGradVis + c] = 0.0f;
    return;
  }
  const uint32_t ca = q0 / kSumChunk, cb = (q1 - 1u) / kSumChunk;
  if (ca == cb) return;
  float g[kGradVis];
  const float* o = aux + (size_t)ca * 2u * kGradVis;
#pragma unroll
  for (int c = 0; c < kGradVis; ++c) g[c] = o[c];
  for (uint32_t ch = ca + 1u; ch <= cb; ++ch) {
    const float* t = aux + ((size_t)ch * 2u + 1u) * kGradVis;
#pragma unroll
    for (int c = 0; c < kGradVis; ++c) g[c] += t[c];
  }
#pragma unroll
  for (int c = 0; c < kGradVis; ++c) grad_vis[(size_t)k * kGradVis + c] = g[c];
}

// dY_k/d(x, y, z) of the real SH polynomials of sh_basis, contracted with
// dL/dY_k: returns dL/d(x, y, z) for a unit direction.
__device__ __forceinline__ void sh_dir_grad(float x, float y, float z, const float dY[16], int ncoef,
                                            float g[3]) {
  const float xx = x * x, yy = y * y, zz = z * z;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  if (ncoef > 1) {
    gy += -kShC1 * dY[1];
    gz += kShC1 * dY[2];
    gx += -kShC1 * dY[3];
  }
  if (ncoef > 4) {
    gx += kShC2_0 * y * dY[4];
    gy += kShC2_0 * x * dY[4];
    gy += -kShC2_0 * z * dY[5];
    gz += -kShC2_0 * y * dY[5];
    gx += -2.0f * kShC2_1 * x * dY[6];
    gy += -2.0f * kShC2_1 * y * dY[6];
    gz += 4.0f * kShC2_1 * z * dY[6];
    gx += -kShC2_0 * z * dY[7];
    gz += -kShC2_0 * x * dY[7];
    gx += 2.0f * kShC2_2 * x * dY[8];
    gy += -2.0f * kShC2_2 * y * dY[8];
  }
  if (ncoef > 9) {
    gx += -6.0f * kShC3_0 * x * y * dY[9];
    gy += -kShC3_0 * (3.0f * xx - 3.0f * yy) * dY[9];
    gx += kShC3_1 * y * z * dY[10];
    gy += kShC3_1 * x * z * dY[10];
    gz += kShC3_1 * x * y * dY[10];
    gx += 2.0f * kShC3_2 * x * y * dY[11];
    gy += -kShC3_2 * (4.0f * zz - xx - 3.0f * yy) * dY[11];
    gz += -8.0f * kShC3_2 * y * z * dY[11];
    gx += -6.0f * kShC3_3 * x * z * dY[12];
    gy += -6.0f * kShC3_3 * y * z * dY[12];
    gz += kShC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy) * dY[12];
    gx += -kShC3_2 * (4.0f * zz - 3.0f * xx - yy) * dY[13];
    gy += 2.0f * kShC3_2 * x * y * dY[13];
    gz += -8.0f * kShC3_2 * x * z * dY[13];
    gx += 2.0f * kShC3_4 * x * z * dY[14];
    gy += -2.0f * kShC3_4 * y * z * dY[14];
    gz += kShC3_4 * (xx - yy) * dY[14];
    gx += -kShC3_0 * (3.0f * xx - 3.0f * yy) * dY[15];
    gy += 6.0f * kShC3_0 * x * y * dY[15];
  }
  g[0] = gx;
  g[1] = gy;
  g[2] = gz;
}

struct GradPlanes {
  float* p[11];
  float* sh;
  float* viewspace;
  int32_t* count;
};

// One thread per Gaussian, looping over the frame's views: the per-view
// parts (SH colour and view direction, projection, 2D inverse and EWA
// adjoints) use that view's visible-splat gradient; their dL/dSigma (world
// covariance) contributions are summed and the view-independent part
// (Sigma = R S S^T R^T with its scale clamp and quaternion normalization)
// runs once.  Each gradient plane entry is read-modified-written by exactly
// one thread: no atomics, and the same result on every run for any number
// of views.  The view's visible index comes from the scanned per-(view,
// Gaussian) offsets (visible iff the next offset's visible count is larger).
//
// SINGLE (one view): a wave's SH coefficients are read, and their gradients
// written, 16 Gaussians at a time with coalesced accesses (their 3 KB are
// contiguous), passing through LDS to and from the lanes that own them; a
// lane reading or writing its own 192 B at a 192 B stride meets a different
// cache line in every lane.
constexpr int kBwdShPad = 13;  // LDS rows of 13 quads: the 16 owners hit distinct banks
template <bool SINGLE>
__global__ __launch_bounds__(256) void k_preprocess_bwd(SceneDev s,
                                                        const ViewParams* __restrict__ views,
                                                        int nviews,
                                                        const uint64_t* __restrict__ offsets,
                                                        const uint64_t* __restrict__ totals,
                                                        const SortPlan* __restrict__ plan,
                                                        const float* __restrict__ grad_vis,
                                                        GradPlanes out, bool assign) {
  __shared__ float4 s_sh[SINGLE ? 4 : 1][SINGLE ? 16 * kBwdShPad : 1];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const bool valid = i < s.n;
  if (!SINGLE && !valid) return;
  const int lane = threadIdx.x & 63;
  const uint32_t wfirst = blockIdx.x * 256u + (threadIdx.x & ~63u);
  float4* __restrict__ shw = &s_sh[SINGLE ? threadIdx.x >> 6 : 0][0];
  auto wave_sync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  if (plan->overflow) {  // the frame was not rendered: no gradient (uniform)
    if (valid && assign) {
      for (int k = 0; k < 11; ++k) out.p[k][i] = 0.0f;
      float4* z4 = reinterpret_cast<float4*>(out.sh) + (size_t)i * 12u;
#pragma unroll
      for (int qd = 0; qd < 12; ++qd) z4[qd] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return;
  }
  // the gradient planes are added to, or written (assign)
  auto put = [&](int plane, float g) {
    if (assign) out.p[plane][i] = g;
    else out.p[plane][i] += g;
  };
  const uint64_t nv = (uint64_t)nviews * s.n;
  const uint32_t ic = valid ? i : 0u;  // (lanes past the scene read Gaussian 0, unused)
  const float px = s.px[ic], py = s.py[ic], pz = s.pz[ic];
  const int ncoef = views[0].sh_coeffs;  // the same in every view of a frame
  const int nfloat = 3 * ncoef;
  const int nq = (nfloat + 3) / 4;  // quads of coefficients the degree uses
  bool loaded = false;
  float c[48];     // SH coefficients (loaded at the first visible view)
  float gsh[48];   // dL/dsh
#pragma unroll
  for (int e = 0; e < 48; ++e) gsh[e] = 0.0f;
  if (SINGLE) {  // the coefficients of the wave's splats visible in the view
    const bool vis0 = valid && (uint32_t)(i + 1 < nv ? offsets[i + 1] : *totals) !=
                                   (uint32_t)offsets[i];
    const uint64_t visb = __ballot(vis0);
    const float4* __restrict__ sh4 = reinterpret_cast<const float4*>(s.sh);
#pragma unroll
    for (int e = 0; e < 48; ++e) c[e] = 0.0f;
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
        shw[gl * kBwdShPad + q] = r[ph & 1][k];
      }
      wave_sync();
      if ((lane >> 4) == ph && vis0) {
#pragma unroll
        for (int q = 0; q < 12; ++q) {
          const float4 x = q < nq ? shw[(lane & 15) * kBwdShPad + q] : make_float4(0.f, 0.f, 0.f, 0.f);
          c[4 * q + 0] = x.x;
          c[4 * q + 1] = x.y;
          c[4 * q + 2] = x.z;
          c[4 * q + 3] = x.w;
        }
      }
      wave_sync();
    }
  }
  float gpos[3] = {0.0f, 0.0f, 0.0f};
  float gop = 0.0f, vs = 0.0f;
  int32_t seen = 0;
  float S[3][3], M[9], R[9], sc[3];
  bool sclamp[3];
  float qn = 0.0f;
  Quat qh{1.0f, 0.0f, 0.0f, 0.0f};
  float F[3][3];  // dL/dSigma summed over the views (symmetric use)
#pragma unroll
  for (int rr = 0; rr < 3; ++rr)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) F[rr][cc] = 0.0f;

  for (int vw = 0; vw < (valid ? nviews : 0); ++vw) {
    const uint64_t idx = (uint64_t)vw * s.n + i;
    const uint64_t off0 = offsets[idx];
    const uint64_t off1 = idx + 1 < nv ? offsets[idx + 1] : *totals;
    if ((uint32_t)off1 == (uint32_t)off0) continue;  // not visible in this view
    const uint32_t k = (uint32_t)off0;
    const ViewParams& v = views[vw];
    if (!loaded) {
      loaded = true;
      if (!SINGLE) {
        const float4* sh4 = reinterpret_cast<const float4*>(s.sh) + (size_t)i * 12u;
#pragma unroll
        for (int qd = 0; qd < 12; ++qd) {
          const float4 x = (4 * qd < nfloat) ? sh4[qd] : make_float4(0.f, 0.f, 0.f, 0.f);
          c[4 * qd + 0] = x.x;
          c[4 * qd + 1] = x.y;
          c[4 * qd + 2] = x.z;
          c[4 * qd + 3] = x.w;
        }
      }
      // world covariance Sigma = M M^T, M = R(q) diag(s), as the forward
      const Quat qraw{s.rw[i], s.rx[i], s.ry[i], s.rz[i]};
      qn = sqrtf(quat_norm2(qraw));
      qh = quat_normalized(qraw);
      quat_matrix(qh, R);
      sc[0] = s.sx[i];
      sc[1] = s.sy[i];
      sc[2] = s.sz[i];
      for (int a = 0; a < 3; ++a) {
        sclamp[a] = sc[a] < 0.00001f;
        if (sclamp[a]) sc[a] = 0.00001f;
      }
      for (int rr = 0; rr < 3; ++rr)
        for (int cc = 0; cc < 3; ++cc) M[3 * rr + cc] = R[3 * rr + cc] * sc[cc];
      for (int rr = 0; rr < 3; ++rr)
        for (int cc = 0; cc < 3; ++cc)
          S[rr][cc] =
              M[3 * rr] * M[3 * cc] + M[3 * rr + 1] * M[3 * cc + 1] + M[3 * rr + 2] * M[3 * cc + 2];
    }
    const float* g = grad_vis + (size_t)k * kGradVis;
    const float gmx = g[0], gmy = g[1], ga = g[2], gb = g[3], gc = g[4];
    const float* W = v.rot;
    gop += g[5];

    // colour = clamp(0.5 + sum_k sh_k Y_k(dir)), dir = normalize(pos - eye)
    float ux = px - v.eye[0], uy = py - v.eye[1], uz = pz - v.eye[2];
    const float dl = sqrtf(ux * ux + uy * uy + uz * uz);
    normalize3(ux, uy, uz);
    float Y[16];
    sh_basis(ux, uy, uz, Y);
    // the forward's colour sums, in its order, for the clamp decision
    float raw[3] = {0.5f, 0.5f, 0.5f};
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
      if (kk < ncoef)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) raw[ch] += c[3 * kk + ch] * Y[kk];
    float gcol[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) gcol[ch] = (raw[ch] < 0.0f || raw[ch] > 1.0f) ? 0.0f : g[6 + ch];
    float dY[16];
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      dY[kk] = 0.0f;
      if (kk < ncoef) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          dY[kk] += gcol[ch] * c[3 * kk + ch];
          gsh[3 * kk + ch] += gcol[ch] * Y[kk];
        }
      }
    }
    if (dl > 0.000001f) {
      float gd[3];
      sh_dir_grad(ux, uy, uz, dY, ncoef, gd);
      const float dot = gd[0] * ux + gd[1] * uy + gd[2] * uz;
      gpos[0] += (gd[0] - ux * dot) / dl;
      gpos[1] += (gd[1] - uy * dot) / dl;
      gpos[2] += (gd[2] - uz * dot) / dl;
    }

    // camera-space position (same rotation as the forward)
    const Quat pq{0.0f, px - v.eye[0], py - v.eye[1], pz - v.eye[2]};
    const Quat r = quat_mul(quat_mul(v.q_view, pq), v.q_view_inv);
    const float x = r.x, y = r.y;
    const float z = r.z < 0.0001f ? 0.0001f : r.z;
    const float fx = v.fx, fy = v.fy;
    float gpc[3];
    gpc[0] = gmx * fx / z;
    gpc[1] = gmy * fy / z;
    gpc[2] = g[9] - gmx * fx * x / (z * z) - gmy * fy * y / (z * z);
    vs += sqrtf(gmx * gmx + gmy * gmy);
    ++seen;

    // T = J W and the screen covariance
    const float itz = 1.0f / z, itz2 = itz * itz;
    const float j00 = fx * itz, j02 = -fx * x * itz2, j11 = fy * itz, j12 = -fy * y * itz2;
    float t0[3], t1[3], St0[3], St1[3];
    for (int a = 0; a < 3; ++a) {
      t0[a] = j00 * W[a] + j02 * W[6 + a];
      t1[a] = j11 * W[3 + a] + j12 * W[6 + a];
    }
    for (int rr = 0; rr < 3; ++rr) {
      St0[rr] = S[rr][0] * t0[0] + S[rr][1] * t0[1] + S[rr][2] * t0[2];
      St1[rr] = S[rr][0] * t1[0] + S[rr][1] * t1[1] + S[rr][2] * t1[2];
    }
    const float c00 = t0[0] * St0[0] + t0[1] * St0[1] + t0[2] * St0[2] + v.lowpass;
    const float c01 = t0[0] * St1[0] + t0[1] * St1[1] + t0[2] * St1[2];
    const float c11 = t1[0] * St1[0] + t1[1] * St1[1] + t1[2] * St1[2] + v.lowpass;
    const float det = c00 * c11 - c01 * c01;
    float gc00 = 0.0f, gc01 = 0.0f, gc11 = 0.0f;
    if (!(det > -0.0000001f && det < 0.0000001f)) {
      const float id = 1.0f / det;
      const float K = (-ga * c11 + gb * c01 - gc * c00) * id * id;
      gc00 = gc * id + K * c11;
      gc11 = ga * id + K * c00;
      gc01 = -gb * id - 2.0f * K * c01;
    }
    float gt0[3], gt1[3];
    for (int a = 0; a < 3; ++a) {
      gt0[a] = 2.0f * gc00 * St0[a] + gc01 * St1[a];
      gt1[a] = gc01 * St0[a] + 2.0f * gc11 * St1[a];
    }
    for (int rr = 0; rr < 3; ++rr)
      for (int cc = 0; cc < 3; ++cc)
        F[rr][cc] += gc00 * t0[rr] * t0[cc] + 0.5f * gc01 * (t0[rr] * t1[cc] + t1[rr] * t0[cc]) +
                     gc11 * t1[rr] * t1[cc];
    const float gj00 = gt0[0] * W[0] + gt0[1] * W[1] + gt0[2] * W[2];
    const float gj02 = gt0[0] * W[6] + gt0[1] * W[7] + gt0[2] * W[8];
    const float gj11 = gt1[0] * W[3] + gt1[1] * W[4] + gt1[2] * W[5];
    const float gj12 = gt1[0] * W[6] + gt1[1] * W[7] + gt1[2] * W[8];
    gpc[2] += gj00 * (-fx * itz2) + gj02 * (2.0f * fx * x * itz2 * itz) + gj11 * (-fy * itz2) +
              gj12 * (2.0f * fy * y * itz2 * itz);
    gpc[0] += gj02 * (-fx * itz2);
    gpc[1] += gj12 * (-fy * itz2);
    for (int a = 0; a < 3; ++a) gpos[a] += W[a] * gpc[0] + W[3 + a] * gpc[1] + W[6 + a] * gpc[2];
  }
  if (SINGLE) {
    // dL/dsh of the wave's splats (assigning: every element of every splat,
    // zero past the used coefficients and for splats without gradient;
    // adding: the used elements of splats with one, the rest left as they are)
    const uint64_t wb = __ballot(valid && (assign || seen != 0));
    float4* __restrict__ gsh4w = reinterpret_cast<float4*>(out.sh);
#pragma unroll
    for (int ph = 0; ph < 4; ++ph) {
      if (((wb >> (16 * ph)) & 0xFFFFull) == 0ull) continue;  // uniform
      if ((lane >> 4) == ph) {
#pragma unroll
        for (int q = 0; q < 12; ++q)
          shw[(lane & 15) * kBwdShPad + q] =
              make_float4(gsh[4 * q], gsh[4 * q + 1], gsh[4 * q + 2], gsh[4 * q + 3]);
      }
      wave_sync();
#pragma unroll
      for (int kq = 0; kq < 3; ++kq) {
        const int idx = kq * 64 + lane, gl = idx / 12, q = idx - gl * 12;
        if (!((wb >> (16 * ph + gl)) & 1ull)) continue;
        const float4 gvq = shw[gl * kBwdShPad + q];
        float4* dst = gsh4w + (size_t)(wfirst + 16u * ph + gl) * 12u + q;
        if (assign) {
          *dst = gvq;
        } else if (4 * q < nfloat) {
          float4 o = *dst;
          if (4 * q + 0 < nfloat) o.x += gvq.x;
          if (4 * q + 1 < nfloat) o.y += gvq.y;
          if (4 * q + 2 < nfloat) o.z += gvq.z;
          if (4 * q + 3 < nfloat) o.w += gvq.w;
          *dst = o;
        }
      }
      wave_sync();
    }
  }
  if (!valid) return;
  float4* gsh4 = reinterpret_cast<float4*>(out.sh) + (size_t)i * 12u;
  if (!seen) {  // no gradient; written as zero when assigning
    if (assign) {
      for (int k = 0; k < 11; ++k) out.p[k][i] = 0.0f;
      if (!SINGLE) {
#pragma unroll
        for (int qd = 0; qd < 12; ++qd) gsh4[qd] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
    return;
  }

  put(10, gop);
  put(0, gpos[0]);
  put(1, gpos[1]);
  put(2, gpos[2]);
  if (out.viewspace) out.viewspace[i] += vs;
  if (out.count) out.count[i] += seen;
  // dL/dsh as float4s (assigning: every element, zero past the used
  // coefficients; adding: those elements are left as they are)
  if (assign && !SINGLE) {
#pragma unroll
    for (int qd = 0; qd < 12; ++qd)
      gsh4[qd] = make_float4(gsh[4 * qd], gsh[4 * qd + 1], gsh[4 * qd + 2], gsh[4 * qd + 3]);
  }
#pragma unroll
  for (int qd = 0; qd < 12; ++qd) {
    if (SINGLE || assign || 4 * qd >= nfloat) continue;
    float4 o = gsh4[qd];
    if (4 * qd + 0 < nfloat) o.x += gsh[4 * qd + 0];
    if (4 * qd + 1 < nfloat) o.y += gsh[4 * qd + 1];
    if (4 * qd + 2 < nfloat) o.z += gsh[4 * qd + 2];
    if (4 * qd + 3 < nfloat) o.w += gsh[4 * qd + 3];
    gsh4[qd] = o;
  }

  // Sigma -> scale and rotation
  float gM[9], gR[9];
  for (int rr = 0; rr < 3; ++rr)
    for (int cc = 0; cc < 3; ++cc)
      gM[3 * rr + cc] = 2.0f * (F[rr][0] * M[cc] + F[rr][1] * M[3 + cc] + F[rr][2] * M[6 + cc]);
  for (int cc = 0; cc < 3; ++cc) {
    const float gs = gM[cc] * R[cc] + gM[3 + cc] * R[3 + cc] + gM[6 + cc] * R[6 + cc];
    put(3 + cc, sclamp[cc] ? 0.0f : gs);  // the 1e-5 clamp passes no gradient
  }
  for (int rr = 0; rr < 3; ++rr)
    for (int cc = 0; cc < 3; ++cc) gR[3 * rr + cc] = gM[3 * rr + cc] * sc[cc];
  const float w = qh.w, qx = qh.x, qy = qh.y, qz = qh.z;
  float gq[4];
  gq[0] = 2.0f * (-qz * gR[1] + qy * gR[2] + qz * gR[3] - qx * gR[5] - qy * gR[6] + qx * gR[7]);
  gq[1] = 2.0f * (qy * gR[1] + qz * gR[2] + qy * gR[3] - 2.0f * qx * gR[4] - w * gR[5] +
                  qz * gR[6] + w * gR[7] - 2.0f * qx * gR[8]);
  gq[2] = 2.0f * (-2.0f * qy * gR[0] + qx * gR[1] + w * gR[2] + qx * gR[3] + qz * gR[5] -
                  w * gR[6] + qz * gR[7] - 2.0f * qy * gR[8]);
  gq[3] = 2.0f * (-2.0f * qz * gR[0] - w * gR[1] + qx * gR[2] + w * gR[3] - 2.0f * qz * gR[4] +
                  qy * gR[5] + qx * gR[6] + qy * gR[7]);
  if (qn > 0.000001f) {
    const float dot = gq[0] * w + gq[1] * qx + gq[2] * qy + gq[3] * qz;
    const float qq[4] = {w, qx, qy, qz};
    for (int a = 0; a < 4; ++a) put(6 + a, (gq[a] - qq[a] * dot) / qn);
  } else if (assign) {  // the identity fallback passes no gradient
    for (int a = 0; a < 4; ++a) out.p[6 + a][i] = 0.0f;
  }
}

}  // namespace lumen

using namespace lumen;

static LumenStatus backward_impl(LumenExContext* c, const LumenExScene* sc,
                                 const LumenExCamera* cams, uint32_t n_views,
                                 const LumenExRenderConfig* cfg, const float* grad_rgb,
                                 const float* grad_depth, LumenExGradients* grads, void* stream) {
  if (!c || !sc || !cams || !cfg || !grad_rgb || !grads)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render_backward: null argument");
  if (!c->pixels)
    return fail(LUMEN_ERROR_UNINITIALIZED,
                "lumen_ex_render_backward: the last frame did not keep LUMEN_EX_SAVE_PIXELS");
  if (sc->count != c->N || n_views != c->nviews || cfg->tile_width != c->tile_w ||
      cfg->tile_height != c->tile_h)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                "lumen_ex_render_backward: scene/cameras/config differ from the last frame");
  for (uint32_t k = 0; k < n_views; ++k)
    if ((int32_t)cams[k].width != c->host_views[k].width ||
        (int32_t)cams[k].height != c->host_views[k].height)
      return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                  "lumen_ex_render_backward: view %u differs from the last frame", k);
  float* planes[11] = {grads->pos_x, grads->pos_y, grads->pos_z, grads->scale_x, grads->scale_y,
                       grads->scale_z, grads->rot_w, grads->rot_x, grads->rot_y, grads->rot_z,
                       grads->opacity};
  for (float* p : planes)
    if (!p) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render_backward: null gradient plane");
  if (!grads->sh) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_render_backward: null sh gradient");
  hipStream_t st = (hipStream_t)stream;
  LUMEN_HIP(hipSetDevice(c->device));
  c->has_grad = true;
  // Sizes: exact after a plain frame, the reserved capacity after a graph
  // replay (the kernels read the frame's own counts from the device plan).
  const uint64_t v_cap = c->sizes_known ? c->V : c->v_cap;
  if (sc->count == 0 || v_cap == 0 || (c->sizes_known && c->M == 0)) {
    // no splat reached a tile: every gradient is zero.  Assigning callers
    // hand in uninitialised planes, so they are written; adding callers'
    // planes stay as they are.
    if (cfg->flags & LUMEN_EX_GRAD_ASSIGN) {
      for (float* p : planes) LUMEN_TRY(fill<float>(p, 0.0f, sc->count, st));
      LUMEN_TRY(fill<float>(grads->sh, 0.0f, sc->count * 48ull, st));
    }
    return LUMEN_SUCCESS;
  }
  const uint64_t m_cap = c->sizes_known ? c->M : c->m_cap;
  LUMEN_TRY(ensure(c->partial, partial_bytes(m_cap)));
  LUMEN_TRY(ensure(c->grad_vis, grad_vis_bytes(v_cap)));
  const uint32_t nchunks = div_up(m_cap, kSumChunk);
  LUMEN_TRY(ensure(c->sum_seg, sum_seg_bytes(m_cap)));
  LUMEN_TRY(ensure(c->sum_aux, sum_aux_bytes(m_cap)));
  float* gv = c->grad_vis.as<float>();
  const SortPlan* plan = c->plan.as<SortPlan>();
  float* part = c->partial.as<float>();
  const uint4* vmeta = c->vmeta.as<uint4>();
  const uint32_t zb = div_up(v_cap, 256) < 8192 ? div_up(v_cap, 256) : 8192;
  hipLaunchKernelGGL(k_chunk_seg, dim3(zb ? zb : 1), dim3(256), 0, st, plan, vmeta,
                     c->sum_seg.as<uint32_t>());
  LUMEN_LAUNCHED();
  const uint32_t threads = cfg->tile_width * cfg->tile_height;
  // 64 splats per LDS batch by default: the per-wave gradient slots (40 B
  // per splat per wave) then leave room for 8 waves per SIMD, which measured
  // 6% faster than 128 (5 waves); 128 when the config asks for 128 or 256
  const uint32_t batch = (cfg->batch_depth == 128u || cfg->batch_depth == 256u) ? 128u : 64u;
  const int2* ranges = c->ranges.as<int2>();
  const uint32_t* v0 = c->vals[0].as<uint32_t>();
  const uint32_t* v1 = c->vals[1].as<uint32_t>();
  const float4* rec = c->rec.as<float4>();
  const ViewParams* views = c->views.as<ViewParams>();
  const ViewSpan* spans = c->spans.as<ViewSpan>();
  const float* tf = c->t_final.as<float>();
  const int32_t* ne = c->n_eval.as<int32_t>();
  const int nv = (int)c->nviews;
  // one pixel per lane: two pixels per lane measured 4% slower (the adjoint
  // state of two pixels costs resident waves)
  const dim3 grid((uint32_t)c->T), block(threads);
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_BWD_COMPOSITE, false, st));
  // the depth channel's terms only where a depth gradient is given and a
  // view computes depth
  bool depth = false;
  for (const ViewParams& v : c->host_views) depth = depth || v.compute_depth;
  depth = depth && grad_depth != nullptr;
  auto kb = batch == 64 ? (depth ? k_composite_bwd<64, 1, true> : k_composite_bwd<64, 1, false>)
                        : (depth ? k_composite_bwd<128, 1, true> : k_composite_bwd<128, 1, false>);
  hipLaunchKernelGGL(kb, grid, block, 0, st, ranges, v0, v1, plan, rec, views, nv, spans, vmeta,
                     tf, ne, grad_rgb, grad_depth, part, c->tile_order.as<uint32_t>());
  LUMEN_LAUNCHED();
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_BWD_COMPOSITE, true, st));
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_BWD_PREPROCESS, false, st));
  GradPlanes gp;
  for (int k = 0; k < 11; ++k) gp.p[k] = planes[k];
  gp.sh = grads->sh;
  gp.viewspace = grads->viewspace_grad_2d;
  gp.count = grads->observation_count;
  float* aux = c->sum_aux.as<float>();
  hipLaunchKernelGGL(k_sum_chunks, dim3(nchunks ? nchunks : 1), dim3(256), 0, st, vmeta, part,
                     plan, c->sum_seg.as<uint32_t>(), gv, aux);
  LUMEN_LAUNCHED();
  hipLaunchKernelGGL(k_sum_fixup, dim3(div_up(v_cap, 256)), dim3(256), 0, st, vmeta, plan, aux,
                     gv);
  LUMEN_LAUNCHED();
  hipLaunchKernelGGL(nv == 1 ? k_preprocess_bwd<true> : k_preprocess_bwd<false>,
                     dim3(div_up(sc->count, 256)), dim3(256), 0, st,
                     scene_dev(sc), views, nv, c->packed.as<uint64_t>(),
                     c->totals.as<uint64_t>(), plan, gv, gp,
                     (cfg->flags & LUMEN_EX_GRAD_ASSIGN) != 0);
  LUMEN_LAUNCHED();
  LUMEN_TRY(mark(c, LUMEN_EX_TIME_BWD_PREPROCESS, true, st));
  return LUMEN_SUCCESS;
}

extern "C" LumenStatus lumen_ex_render_backward(LumenExContext* c, const LumenExScene* sc,
                                                const LumenExCamera* cam,
                                                const LumenExRenderConfig* cfg,
                                                const float* grad_rgb, const float* grad_depth,
                                                LumenExGradients* grads, void* stream) {
  clear_error();
  return backward_impl(c, sc, cam, 1, cfg, grad_rgb, grad_depth, grads, stream);
}

extern "C" LumenStatus lumen_ex_render_batch_backward(LumenExContext* c, const LumenExScene* sc,
                                                      const LumenExCamera* cams, uint32_t n_views,
                                                      const LumenExRenderConfig* cfg,
                                                      const float* grad_rgb,
                                                      const float* grad_depth,
                                                      LumenExGradients* grads, void* stream) {
  clear_error();
  return backward_impl(c, sc, cams, n_views, cfg, grad_rgb, grad_depth, grads, stream);
}
