// optim.hip -- the optimizer side of a training step on the GPU:
//
//   lumen_ex_adam_step          AdamOptimizer.step (src/optim/optimizer.mojo:112-170)
//   lumen_ex_reset_opacities    reset_opacities (src/optim/density_control.mojo:101-109)
//   lumen_ex_densify_and_prune  densify_and_prune (src/optim/density_control.mojo:25-98)
//
// Adam is elementwise: one thread per Gaussian over the parameter planes it
// updates (pos xyz and opacity, as the reference; all 59 with
// LUMEN_EX_ADAM_ALL_PARAMS), one thread per coefficient over the SH block.
// The learning-rate schedule and bias corrections are scalars the host
// computes once per step (C library powf) and passes in.  Every expression
// keeps the reference's Float32 order (no contraction), so the result is the
// same bit for bit as the scalar loop.
//
// Densification keeps the reference's output order (source order; a
// clone's original before its offset copy, split children + then -):
//   k_dens_classify  one thread per Gaussian: prune / keep / clone / split
//                    and its output count (0, 1, 2, 2); per-block stats.
//   exclusive scan of the counts -> output offsets (host reads the total).
//   k_dens_emit      one block per 256 sources, whose outputs form one
//                    contiguous run: the block maps each output row to its
//                    (source, child) in LDS, then writes the 11 planes one
//                    row per thread and the 48 SH coefficients (and the Adam
//                    moments) one float per thread, so every plane and the
//                    SH block are written in whole coalesced runs.
// Every output Gaussian's rotation is normalized, as the Gaussian3D
// constructor does (src/scene/gaussian.mojo:27-37).  With moments given,
// kept Gaussians and a clone's original keep their Adam moments; the new
// Gaussians (the offset copy, both split children) start at zero.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"
#include "context.hpp"
#include "lumen_ex.h"
#include "scan.hpp"

namespace lumen {

struct Planes12 {
  float* p[12];  // pos xyz, scale xyz, rot wxyz, opacity, sh
};

inline Planes12 planes12(const LumenExPlanes* s) {
  return Planes12{{s->pos_x, s->pos_y, s->pos_z, s->scale_x, s->scale_y, s->scale_z, s->rot_w,
                   s->rot_x, s->rot_y, s->rot_z, s->opacity, s->sh}};
}

inline Planes12 planes12(const LumenExGradients* s) {
  return Planes12{{s->pos_x, s->pos_y, s->pos_z, s->scale_x, s->scale_y, s->scale_z, s->rot_w,
                   s->rot_x, s->rot_y, s->rot_z, s->opacity, s->sh}};
}

struct AdamScalars {
  float beta1, beta2, one_minus_b1, one_minus_b2, eps, inv_b1, inv_b2;
  float lr[12];  // per plane (sh last)
};

// m, v and the parameter (optimizer.mojo:135-141)
__device__ __forceinline__ float adam_one(float p, float g, float& m, float& v, float lr,
                                          const AdamScalars& s) {
  m = s.beta1 * m + s.one_minus_b1 * g;
  v = s.beta2 * v + s.one_minus_b2 * g * g;
  const float mh = m * s.inv_b1;
  const float vh = v * s.inv_b2;
  return p - lr * mh / (sqrtf(vh) + s.eps);
}

__global__ __launch_bounds__(256) void k_adam_planes(Planes12 prm, Planes12 grd, Planes12 m,
                                                     Planes12 v, uint64_t n, AdamScalars s,
                                                     int all) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256u) {
#pragma unroll
    for (int k = 0; k < 11; ++k) {
      const bool ref_plane = k < 3 || k == 10;
      if (!ref_plane && !all) continue;
      float mm = m.p[k][i], vv = v.p[k][i];
      float x = adam_one(prm.p[k][i], grd.p[k][i], mm, vv, s.lr[k], s);
      if (k == 10) {  // opacity clamp (optimizer.mojo:165-168)
        if (x < 0.01f) x = 0.01f;
        else if (x > 0.99f) x = 0.99f;
      }
      m.p[k][i] = mm;
      v.p[k][i] = vv;
      prm.p[k][i] = x;
    }
  }
}

__global__ __launch_bounds__(256) void k_adam_sh(float* __restrict__ p, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v,
                                                 uint64_t n48, AdamScalars s) {
  for (uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x; j < n48;
       j += (uint64_t)gridDim.x * 256u) {
    float mm = m[j], vv = v[j];
    p[j] = adam_one(p[j], g[j], mm, vv, s.lr[11], s);
    m[j] = mm;
    v[j] = vv;
  }
}

__global__ __launch_bounds__(256) void k_reset_opacity(float* __restrict__ op, uint64_t n,
                                                       float max_val) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256u)
    if (op[i] > max_val) op[i] = max_val;
}

// ---- densification ---------------------------------------------------------

enum : uint8_t { kPrune = 0, kKeep = 1, kClone = 2, kSplit = 3 };

struct DensParams {
  float prune_opacity, grad_threshold, extent_half, scale_threshold;
};

// stats[0..2] += cloned, split, pruned of the block
__global__ __launch_bounds__(256) void k_dens_classify(SceneDev sc, const float* __restrict__ vs_grad,
                                                       const int32_t* __restrict__ vs_count,
                                                       DensParams dp, uint8_t* __restrict__ cls,
                                                       uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ stats) {
  __shared__ uint32_t s_st[3];
  if (threadIdx.x < 3) s_st[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < sc.n) {
    uint8_t c;
    const float op = sc.op[i];
    if (op < dp.prune_opacity) {
      c = kPrune;
    } else {
      float ms = sc.sx[i];
      const float sy = sc.sy[i], sz = sc.sz[i];
      if (sy > ms) ms = sy;
      if (sz > ms) ms = sz;
      if (ms > dp.extent_half) {
        c = kPrune;
      } else {
        float avg = 0.0f;
        const int32_t cnt = vs_count[i];
        if (cnt > 0) avg = vs_grad[i] / (float)cnt;
        c = avg > dp.grad_threshold ? (ms <= dp.scale_threshold ? kClone : kSplit) : kKeep;
      }
    }
    cls[i] = c;
    counts[i] = c == kPrune ? 0u : (c == kKeep ? 1u : 2u);
    if (c == kClone) atomicAdd(&s_st[0], 1u);
    if (c == kSplit) atomicAdd(&s_st[1], 1u);
    if (c == kPrune) atomicAdd(&s_st[2], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 3 && s_st[threadIdx.x]) atomicAdd(&stats[threadIdx.x], s_st[threadIdx.x]);
}

__device__ __forceinline__ void normalized_quat(float w, float x, float y, float z, float q[4]) {
  const float ns = ((w * w + x * x) + y * y) + z * z;
  const float nn = sqrtf(ns);
  if (nn > 0.000001f) {
    const float inv = 1.0f / nn;
    q[0] = w * inv;
    q[1] = x * inv;
    q[2] = y * inv;
    q[3] = z * inv;
  } else {
    q[0] = 1.0f;
    q[1] = q[2] = q[3] = 0.0f;
  }
}

struct Moments {
  Planes12 m_in, v_in, m_out, v_out;
};

template <bool kMoments>
__global__ __launch_bounds__(256) void k_dens_emit(SceneDev sc, const uint8_t* __restrict__ cls,
                                                   const uint32_t* __restrict__ offsets,
                                                   Planes12 out, Moments mo) {
  __shared__ uint16_t s_map[512];  // output row -> source (local) * 2 + child
  __shared__ uint8_t s_cls[256];
  __shared__ uint32_t s_nout;
  const uint32_t base = blockIdx.x * 256u;
  const uint32_t i = base + threadIdx.x;
  const uint32_t last = (sc.n - base < 256u ? sc.n - base : 256u) - 1u;  // last local source
  const uint32_t o_start = offsets[base];
  if (i < sc.n) {
    const uint8_t c = cls[i];
    s_cls[threadIdx.x] = c;
    const uint32_t local = offsets[i] - o_start;
    const uint32_t cnt = c == kPrune ? 0u : (c == kKeep ? 1u : 2u);
    for (uint32_t k = 0; k < cnt; ++k) s_map[local + k] = (uint16_t)((threadIdx.x << 1) | k);
    if (threadIdx.x == last) s_nout = local + cnt;
  }
  __syncthreads();
  const uint32_t nout = s_nout;
  // the 11 planes, one output row per thread
  for (uint32_t r = threadIdx.x; r < nout; r += 256u) {
    const uint32_t code = s_map[r];
    const uint32_t src = base + (code >> 1);
    const uint32_t child = code & 1u;
    const uint8_t c = s_cls[code >> 1];
    float pos[3] = {sc.px[src], sc.py[src], sc.pz[src]};
    float scl[3] = {sc.sx[src], sc.sy[src], sc.sz[src]};
    float q[4];
    normalized_quat(sc.rw[src], sc.rx[src], sc.ry[src], sc.rz[src], q);
    if (c == kClone && child == 1u) {  // offset copy (density_control.mojo:79-81)
#pragma unroll
      for (int k = 0; k < 3; ++k) pos[k] = pos[k] + 0.01f * scl[k];
    } else if (c == kSplit) {  // density_control.mojo:85-91
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float ns = scl[k] / 1.6f;
        pos[k] = child == 0u ? pos[k] + ns * 0.5f : pos[k] - ns * 0.5f;
        scl[k] = ns;
      }
    }
    const uint64_t o = (uint64_t)o_start + r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      out.p[k][o] = pos[k];
      out.p[3 + k][o] = scl[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) out.p[6 + k][o] = q[k];
    out.p[10][o] = sc.op[src];
    if (kMoments) {
      const bool fresh = (c == kClone && child == 1u) || c == kSplit;
#pragma unroll
      for (int k = 0; k < 11; ++k) {
        mo.m_out.p[k][o] = fresh ? 0.0f : mo.m_in.p[k][src];
        mo.v_out.p[k][o] = fresh ? 0.0f : mo.v_in.p[k][src];
      }
    }
  }
  // the SH block (and its moments), one coefficient per thread
  const uint32_t total = nout * 48u;
  for (uint32_t e = threadIdx.x; e < total; e += 256u) {
    const uint32_t r = e / 48u;
    const uint32_t col = e - r * 48u;
    const uint32_t code = s_map[r];
    const uint64_t src = (uint64_t)(base + (code >> 1)) * 48u + col;
    const uint64_t dst = ((uint64_t)o_start + r) * 48u + col;
    out.p[11][dst] = sc.sh[src];
    if (kMoments) {
      const uint8_t c = s_cls[code >> 1];
      const bool fresh = (c == kClone && (code & 1u)) || c == kSplit;
      mo.m_out.p[11][dst] = fresh ? 0.0f : mo.m_in.p[11][src];
      mo.v_out.p[11][dst] = fresh ? 0.0f : mo.v_in.p[11][src];
    }
  }
}

static bool planes_complete(const LumenExPlanes* p, bool all) {
  const Planes12 q = planes12(p);
  for (int k = 0; k < 12; ++k) {
    const bool needed = all || k < 3 || k == 10;
    if (needed && !q.p[k]) return false;
  }
  return true;
}

static uint32_t grid_for(uint64_t n) {
  const uint32_t b = div_up(n ? n : 1, 256);
  return b < 16384 ? b : 16384;
}

}  // namespace lumen

using namespace lumen;

extern "C" {

LumenStatus lumen_ex_adam_step(const LumenExPlanes* params, const LumenExGradients* grads,
                               const LumenExPlanes* m, const LumenExPlanes* v, uint64_t count,
                               const LumenExTrainConfig* cfg, int64_t step, uint32_t flags,
                               void* stream) {
  clear_error();
  if (!params || !grads || !m || !v || !cfg)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_adam_step: null argument");
  const bool all = (flags & LUMEN_EX_ADAM_ALL_PARAMS) != 0;
  if (!planes_complete(params, all) || !planes_complete(m, all) || !planes_complete(v, all))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_adam_step: null plane");
  const Planes12 g = planes12(grads);
  for (int k = 0; k < 12; ++k)
    if ((all || k < 3 || k == 10) && !g.p[k])
      return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_adam_step: null gradient plane");
  if (step < 1 || cfg->iterations < 1)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_adam_step: step and iterations must be >= 1");
  if (count == 0) return LUMEN_SUCCESS;
  // get_position_lr at the step count and the bias corrections
  // (optimizer.mojo:103-110,124-130)
  float progress = (float)step / (float)cfg->iterations;
  if (progress > 1.0f) progress = 1.0f;
  const float ratio = cfg->position_lr_final / cfg->position_lr_init;
  const float pos_lr = cfg->position_lr_init * powf(ratio, progress);
  const float t = (float)step;
  AdamScalars s;
  s.beta1 = cfg->beta1;
  s.beta2 = cfg->beta2;
  s.one_minus_b1 = 1.0f - cfg->beta1;
  s.one_minus_b2 = 1.0f - cfg->beta2;
  s.eps = cfg->epsilon;
  s.inv_b1 = 1.0f / (1.0f - powf(cfg->beta1, t));
  s.inv_b2 = 1.0f / (1.0f - powf(cfg->beta2, t));
  for (int k = 0; k < 3; ++k) s.lr[k] = pos_lr;
  for (int k = 3; k < 6; ++k) s.lr[k] = cfg->scaling_lr;
  for (int k = 6; k < 10; ++k) s.lr[k] = cfg->rotation_lr;
  s.lr[10] = cfg->opacity_lr;
  s.lr[11] = cfg->feature_lr;
  hipStream_t st = (hipStream_t)stream;
  const Planes12 P = planes12(params), M = planes12(m), V = planes12(v);
  hipLaunchKernelGGL(k_adam_planes, dim3(grid_for(count)), dim3(256), 0, st, P, g, M, V, count, s,
                     all ? 1 : 0);
  LUMEN_LAUNCHED();
  if (all) {
    hipLaunchKernelGGL(k_adam_sh, dim3(grid_for(count * 48u)), dim3(256), 0, st, P.p[11], g.p[11],
                       M.p[11], V.p[11], count * 48u, s);
    LUMEN_LAUNCHED();
  }
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_reset_opacities(float* opacity, uint64_t count, float max_value,
                                     void* stream) {
  clear_error();
  if (!opacity && count) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_reset_opacities: null argument");
  if (count == 0) return LUMEN_SUCCESS;
  hipLaunchKernelGGL(k_reset_opacity, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream,
                     opacity, count, max_value);
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

LumenStatus lumen_ex_densify_and_prune(LumenExContext* c, const LumenExScene* scene,
                                       const float* viewspace_grad_2d,
                                       const int32_t* observation_count,
                                       const LumenExTrainConfig* cfg, float scene_extent,
                                       const LumenExPlanes* out, uint64_t out_capacity,
                                       const LumenExPlanes* moments_in,
                                       const LumenExPlanes* moments_out,
                                       LumenExDensifyStats* stats, void* stream) {
  clear_error();
  if (!c || !scene || !cfg || !out || !stats || !viewspace_grad_2d || !observation_count)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_densify_and_prune: null argument");
  if ((moments_in == nullptr) != (moments_out == nullptr))
    return fail(LUMEN_ERROR_INVALID_ARGUMENT,
                "lumen_ex_densify_and_prune: moments_in and moments_out go together");
  const uint64_t n = scene->count;
  if (n >= 0x7FFFFFFFull)
    return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_densify_and_prune: too many Gaussians");
  const Planes12 O = planes12(out);
  for (int k = 0; k < 12; ++k)
    if (!O.p[k]) return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_densify_and_prune: null output plane");
  if (moments_in)
    for (int j = 0; j < 2; ++j)
      if (!planes_complete(&moments_in[j], true) || !planes_complete(&moments_out[j], true))
        return fail(LUMEN_ERROR_INVALID_ARGUMENT, "lumen_ex_densify_and_prune: null moment plane");
  memset(stats, 0, sizeof(*stats));
  if (n == 0) return LUMEN_SUCCESS;
  hipStream_t st = (hipStream_t)stream;
  LUMEN_HIP(hipSetDevice(c->device));
  LUMEN_TRY(ensure(c->dens_cls, n));
  LUMEN_TRY(ensure(c->dens_cnt, n * 4));
  LUMEN_TRY(ensure(c->dens_off, n * 4));
  const uint64_t nparts = div_up(n, kScanChunk);
  LUMEN_TRY(ensure(c->dens_misc, 32 + nparts * 4));  // stats[3], total, then the scan partials
  uint32_t* misc = c->dens_misc.as<uint32_t>();
  LUMEN_TRY(fill<uint32_t>(misc, 0u, 4, st));
  DensParams dp;
  dp.prune_opacity = cfg->prune_opacity_threshold;
  dp.grad_threshold = cfg->densify_grad_threshold;
  dp.extent_half = scene_extent * 0.5f;
  dp.scale_threshold = scene_extent * 0.01f;
  const SceneDev sd = scene_dev(scene);
  const uint32_t nb = div_up(n, 256);
  hipLaunchKernelGGL(k_dens_classify, dim3(nb), dim3(256), 0, st, sd, viewspace_grad_2d,
                     observation_count, dp, c->dens_cls.as<uint8_t>(), c->dens_cnt.as<uint32_t>(),
                     misc);
  LUMEN_LAUNCHED();
  LUMEN_TRY(exclusive_scan<uint32_t>(c->dens_cnt.as<uint32_t>(), c->dens_off.as<uint32_t>(), n,
                                     misc + 8, misc + 3, st));
  uint32_t h[4];
  LUMEN_HIP(hipMemcpyAsync(c->pinned, misc, 16, hipMemcpyDeviceToHost, st));
  LUMEN_HIP(hipStreamSynchronize(st));
  memcpy(h, c->pinned, 16);
  stats->cloned = h[0];
  stats->split = h[1];
  stats->pruned = h[2];
  stats->count = h[3];
  if (stats->count > out_capacity)
    return fail(LUMEN_EX_ERROR_CAPACITY,
                "lumen_ex_densify_and_prune: %llu Gaussians exceed the output capacity %llu",
                (unsigned long long)stats->count, (unsigned long long)out_capacity);
  if (stats->count == 0) return LUMEN_SUCCESS;
  if (moments_in) {
    Moments mo{planes12(&moments_in[0]), planes12(&moments_in[1]), planes12(&moments_out[0]),
               planes12(&moments_out[1])};
    hipLaunchKernelGGL(k_dens_emit<true>, dim3(nb), dim3(256), 0, st, sd,
                       c->dens_cls.as<uint8_t>(), c->dens_off.as<uint32_t>(), O, mo);
  } else {
    Moments mo{};
    hipLaunchKernelGGL(k_dens_emit<false>, dim3(nb), dim3(256), 0, st, sd,
                       c->dens_cls.as<uint8_t>(), c->dens_off.as<uint32_t>(), O, mo);
  }
  LUMEN_LAUNCHED();
  return LUMEN_SUCCESS;
}

}  // extern "C"
