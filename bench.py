#!/usr/bin/env python3
"""Benchmark of the splat rasterization path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--legs fwd_bwd,batch,serve,optim]
                    [--no-cpu-baseline] [--copy-calibration]

Headline workload: 1M Gaussians (means N(0, 2 I), log-scales N(ln .01, .5),
random unit quaternions, opacity sigmoid(N(0, 1.5)), FP32 SH degree 3),
1920x1080, fx = fy = 1400, camera at (0, 0, 5) looking at the origin.  A
step renders one frame (culling, covariance, EWA, SH, key emission, radix
sort, tile ranges, compositing) from a scene resident in HBM into FP32
RGB + depth buffers in HBM.  Prints one JSON line.

Legs beside the headline (each in its own key of the line):
  fwd_bwd  the same frame forward + backward through torch.autograd with
           loss 0.8 L1 + 0.2 (1 - SSIM 8x8) against a render of a copy whose
           means moved by N(0, 0.01); steps/s.
  batch    3M Gaussians, 32 1920x1080 cameras on a Catmull-Rom orbit of
           radius 6: every view in one hipGraph-captured frame (tile shape and
           LDS batch depth from the autotuner), forward and forward+backward,
           against the plain per-view path on the same views; views/s.
  serve    8 resident scenes (10k .. 5M Gaussians), 1000 requests, each a
           random scene, resolution in {320x240, 1280x720, 1920x1080,
           3840x2160} and 20% forward+backward: saturation throughput of the
           continuously batching server against one-at-a-time plain renders,
           and p50/p99 latency under Poisson arrivals at a stated rate.
  optim    on the headline scene: an Adam step (pos and opacity, as the
           reference; and all 59 parameters), densify_and_prune with the Adam
           moments carried, each against the HBM roofline with its
           algorithmic bytes; and whole training steps (forward, loss,
           backward, Adam) per second.

Throughput is timed over back-to-back frames; per-stage GPU time comes from
HIP events the library records on the render stream around each stage, in a
second timed loop over the same frames.  The roofline entry is the
stage with the largest share; its algorithmic work per launch is:
  count      68 B / Gaussian (44 B geometry read, 8 B packed count written,
             8 B read + 8 B written by the compaction scan) + 32 B / visible
             splat (geometry cached for emit); + 4 B / Gaussian when emit
             runs the first sort pass (its digit histogram, 512 words per
             512 Gaussians)
  emit       24 B / Gaussian (position, opacity, scan offset) + 312 B / visible
             splat (32 B cached geometry + 192 B SH read, 64 B record + 24 B
             index, radius and key metadata written) + 8 B / key written
             (+ 4 B / Gaussian of digit prefixes read when it runs the first
             sort pass: it then writes the keys in that pass's order)
  sort       per digit pass the sort runs (the frame's passes less the one
             emit ran): 4 B / key histogram read + 16 B / key scatter read
             and write (32-bit keys; 8 + 24 B with 64-bit keys)
  ranges     8 B / key + 8 B / tile
  composite  30 FP32 operations per (pixel, splat) evaluation of the
             reference's compositor (E of them per frame, counted by the
             oracle), the scope table's fixed charge -- VALU bound; its bytes
             are 68 B / key + 8 B / tile + 16 B / pixel
`traffic` is the HBM bytes per launch rocprofv3 counted for that kernel
(FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a plain copy of
known size), read from the committed summary under profiles/ when present.
CPU baseline: the oracle (scalar C restatement of the reference path, one
thread) rendering the headline workload on this host, 1 warm-up and 5 timed
frames, reported as 1 / mean with the mean and median.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
COMPOSITE_FLOPS_PER_EVAL = 30
METRIC = "1M-splat 1080p frames/s (fwd, fwd+bwd); p50/p99 request latency"
KERNELS = {"composite": "k_composite", "sort": "k_sort_hist+k_sort_scatter", "emit": "k_emit",
           "count": "k_count", "ranges": "k_ranges"}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_traffic_random1m_1920x1080.md")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def stage_work(N, V, M, T, passes, E, kb, emit_passes=0):
    """Algorithmic bytes (or flops) per launch of each stage; kb is the
    width of the sorted keys (4 for 32-bit frame keys, 8 for 64-bit);
    emit_passes of the passes ran in emit."""
    return {
        "count": ("hbm", 68.0 * N + 32.0 * V + 4.0 * N * emit_passes),
        "emit": ("hbm", 24.0 * N + 312.0 * V + (kb + 4.0) * M + 4.0 * N * emit_passes),
        # per pass the sort runs: the histogram reads the keys, the scatter
        # reads and writes keys and values
        "sort": ("hbm", (passes - emit_passes) * (kb + 2.0 * (kb + 4.0)) * M),
        "ranges": ("hbm", kb * M + 8.0 * T),
        "composite": ("valu", float(COMPOSITE_FLOPS_PER_EVAL) * E),
    }


def measured_traffic():
    """kernel -> HBM bytes per launch from the committed rocprofv3 summary
    (lines `| kernel | ... | traffic_bytes |`), or {}."""
    out = {}
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                m = re.match(r"\|\s*(k_\w+)\s*\|.*\|\s*([0-9.]+)\s*\|\s*$", line)
                if m:
                    out[m.group(1)] = float(m.group(2))
    except OSError:
        pass
    return out


def roofline_entry(bound, work, ms, traffic=None):
    if bound == "hbm":
        achieved = work / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": int(work), "avg_ms": round(ms, 5)}
    achieved = work / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "algorithmic_flops": int(work), "avg_ms": round(ms, 5)}


def timed(fn, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def timed_median(fn, steps, repeats=3):
    """Median of `repeats` timed runs of `steps` calls (seconds per call)."""
    ts = sorted(timed(fn, steps) for _ in range(repeats))
    return ts[len(ts) // 2]


# ---- headline: one 1M / 1080p frame ----------------------------------------

def headline(args, result):
    import numpy as np
    import torch
    from lumen_amd.raster import Rasterizer, RenderConfig
    from lumen_amd.scene import workload

    log(f"generating {args.workload}")
    scene, cam = workload(args.workload)
    dev = scene.to_device("cuda:0")
    R = Rasterizer(0)
    cfg = RenderConfig()
    W, H = cam.config.width, cam.config.height
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda:0")
    depth = torch.empty((H, W), dtype=torch.float32, device="cuda:0")

    # one render with per-pixel counters for the work model (not timed)
    out = R.render(dev, cam, cfg, save_pixels=True, out_rgb=rgb, out_depth=depth)
    torch.cuda.synchronize()
    E = int(R.stage("n_eval").astype(np.int64).sum())
    V, M, passes = out.visible_splats, out.total_splat_instances, out.sort_passes
    kb = out.key_bytes
    ep = out.emit_passes
    tx, ty = cfg.tiles(W, H)
    T = tx * ty
    N = scene.count

    for _ in range(args.warmup):
        R.render(dev, cam, cfg, out_rgb=rgb, out_depth=depth)
    torch.cuda.synchronize()
    log(f"warmup done; timing {args.steps} frames")
    # the throughput: frames back to back, no stage events in the stream
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.render(dev, cam, cfg, out_rgb=rgb, out_depth=depth)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    # the per-stage kernel times: the same frames with HIP events recorded on
    # the render stream around each stage (each record adds a short gap, so
    # this loop is not the throughput)
    R.set_profiling(True)
    acc = {}
    torch.cuda.synchronize()
    te0 = time.perf_counter()
    for _ in range(args.steps):
        R.render(dev, cam, cfg, out_rgb=rgb, out_depth=depth)
        for k, v in R.stage_times().items():
            acc[k] = acc.get(k, 0.0) + v
    torch.cuda.synchronize()
    te1 = time.perf_counter()
    R.set_profiling(False)
    stage_ms = {k: v / args.steps for k, v in acc.items()}
    ms_events = (te1 - te0) * 1e3 / args.steps

    traffic = measured_traffic()
    work = stage_work(N, V, M, T, passes, E, kb, ep)
    stages = {}
    for k, (b, w) in work.items():
        if stage_ms.get(k, 0) > 0:
            kern = KERNELS[k]
            if k == "sort":  # every pass it runs is a histogram and a scatter
                th, ts = traffic.get("k_sort_hist"), traffic.get("k_sort_scatter")
                tr = (passes - ep) * (th + ts) if th and ts else None
            else:
                tr = traffic.get(kern)
            stages[k] = roofline_entry(b, w, stage_ms[k], int(tr) if tr else None)
    dominant = max(stages, key=lambda k: stages[k]["avg_ms"])
    roofline = dict(stages[dominant])
    roofline["kernel"] = KERNELS[dominant]
    result.update({
        "metric": METRIC,
        "value": round(1000.0 / ms, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "gaussians": N, "width": W, "height": H,
                   "tile": f"{cfg.tile_width}x{cfg.tile_height}", "sh_degree": cfg.max_sh_degree,
                   "mode": "forward", "visible": V, "instances": M, "evaluations": E,
                   "sort_passes": passes, "sort_passes_in_emit": ep, "sort_key_bytes": kb},
        "roofline": roofline,
        "stages": stages,
        "ms_per_step_with_stage_events": round(ms_events, 4),
    })

    if "fwd_bwd" in args.legs:
        from lumen_amd.scene import perturbed
        from lumen_amd.training import combined_loss, rasterize
        with torch.no_grad():
            target = R.render(perturbed(scene, 0.01, 7).to_device("cuda:0"), cam, cfg).rgb.clone()
        planes = dev.planes.clone().requires_grad_(True)
        shp = dev.sh.clone().requires_grad_(True)

        def train_step():
            planes.grad = None
            shp.grad = None
            img, _ = rasterize(planes, shp, cam, cfg, rasterizer=R)
            combined_loss(img, target).backward()

        for _ in range(args.warmup):
            train_step()
        ms_fb = timed(train_step, args.steps) * 1e3
        R.set_profiling(True)
        acc_b = {}
        torch.cuda.synchronize()
        for _ in range(args.steps):
            train_step()
            for k, v in R.stage_times().items():
                acc_b[k] = acc_b.get(k, 0.0) + v
        torch.cuda.synchronize()
        R.set_profiling(False)
        result["fwd_bwd"] = {
            "value": round(1000.0 / ms_fb, 2), "unit": "fwd+bwd steps/s",
            "ms_per_step": round(ms_fb, 4),
            "loss": "0.8 L1 + 0.2 (1 - SSIM 8x8), target = render of means + N(0, 0.01)",
            "stage_ms": {k: round(v / args.steps, 5) for k, v in acc_b.items()}}
        del planes, shp, target
    R.close()
    return scene, cam, cfg, V, M


# ---- batch: 3M Gaussians x 32 orbit views -----------------------------------

def batch_leg(args):
    import torch
    from lumen_amd.batch import MultiViewRenderer, autotune, rasterize_views
    from lumen_amd.raster import RenderConfig
    from lumen_amd.scene import orbit_workload, perturbed
    from lumen_amd.training import combined_loss, rasterize

    log("batch: generating 3M Gaussians, 32 orbit views")
    scene, cams = orbit_workload(3_000_000, 32, 6.0)
    dev = scene.to_device("cuda:0")
    nv = len(cams)
    mv = MultiViewRenderer(0)
    log("batch: autotuning tile shape x LDS batch depth")
    tune = autotune(mv, dev, cams, RenderConfig(), graph=True, reps=2)
    cfg = tune.config
    for _ in range(2):
        out = mv.render(dev, cams, cfg, graph=True)
    stats = mv.status(nv)
    M = sum(s.total_splat_instances for s in stats)
    V = sum(s.visible_splats for s in stats)
    steps = max(3, args.steps // 6)
    # the median of three timed runs: a frame of 32 1080p views takes ~30 ms.
    # The batched frame is timed as a graph replay and as one eager launch
    # per stage (which reads 24 B back mid-frame but sizes its grids to the
    # frame instead of the graph's reserved capacity); the faster is the figure
    t_graph = timed_median(lambda: mv.render(dev, cams, cfg, graph=True), steps)
    batched_rgb = out.rgb.clone()
    mv.render(dev, cams, cfg, graph=False)
    t_eager = timed_median(lambda: mv.render(dev, cams, cfg, graph=False), steps)
    eager_out = mv.render(dev, cams, cfg, graph=False)
    mv.synchronize()
    eager_same = bool(torch.equal(eager_out.rgb, batched_rgb))
    del eager_out
    t_batch = min(t_graph, t_eager)

    R = mv.rasterizer
    W, H = cams[0].config.width, cams[0].config.height
    outs = [(torch.empty((H, W, 3), device="cuda:0"), torch.empty((H, W), device="cuda:0"))
            for _ in range(nv)]

    def plain():
        for c, (o_rgb, o_d) in zip(cams, outs):
            R.render(dev, c, cfg, out_rgb=o_rgb, out_depth=o_d)

    plain()
    t_plain = timed_median(plain, steps)
    P = W * H
    same = all(torch.equal(outs[k][0].reshape(-1), batched_rgb[3 * k * P: 3 * (k + 1) * P])
               for k in range(nv))
    res = {
        "workload": "orbit3m_32x1920x1080", "views": nv, "gaussians": scene.count,
        "visible": V, "instances": M,
        "config": f"{cfg.tile_width}x{cfg.tile_height}/{cfg.batch_depth}",
        "autotune_ms": tune.table,
        "fwd_views_per_s": round(nv / t_batch, 2), "fwd_ms_per_frame": round(t_batch * 1e3, 3),
        "fwd_mode": "graph" if t_graph <= t_eager else "eager",
        "graph_fwd_views_per_s": round(nv / t_graph, 2),
        "eager_fwd_views_per_s": round(nv / t_eager, 2),
        "plain_fwd_views_per_s": round(nv / t_plain, 2),
        "fwd_speedup_vs_plain": round(t_plain / t_batch, 3),
        "identical_to_plain": bool(same and eager_same),
    }
    del outs, batched_rgb
    if "fwd_bwd" in args.legs:
        log("batch: forward+backward")
        with torch.no_grad():
            tgt = mv.render(perturbed(scene, 0.01, 8).to_device("cuda:0"), cams, cfg).rgb.clone()
        targets = [tgt[3 * k * P: 3 * (k + 1) * P].view(H, W, 3) for k in range(nv)]
        planes = dev.planes.clone().requires_grad_(True)
        shp = dev.sh.clone().requires_grad_(True)

        def step_batched(graph=True):
            planes.grad = None
            shp.grad = None
            rgb, _ = rasterize_views(planes, shp, cams, cfg, mv, graph=graph)
            loss = sum(combined_loss(rgb[3 * k * P: 3 * (k + 1) * P].view(H, W, 3), targets[k])
                       for k in range(nv)) / nv
            loss.backward()

        def step_plain():
            planes.grad = None
            shp.grad = None
            for k, c in enumerate(cams):
                img, _ = rasterize(planes, shp, c, cfg, rasterizer=R)
                (combined_loss(img, targets[k]) / nv).backward()

        fb_steps = max(2, args.steps // 10)
        step_batched()
        t_fb_graph = timed(step_batched, fb_steps)
        step_batched(False)
        t_fb_eager = timed(lambda: step_batched(False), fb_steps)
        t_fb = min(t_fb_graph, t_fb_eager)
        step_plain()
        t_fb_plain = timed(step_plain, fb_steps)
        res.update({"fwd_bwd_views_per_s": round(nv / t_fb, 2),
                    "fwd_bwd_mode": "graph" if t_fb_graph <= t_fb_eager else "eager",
                    "graph_fwd_bwd_views_per_s": round(nv / t_fb_graph, 2),
                    "eager_fwd_bwd_views_per_s": round(nv / t_fb_eager, 2),
                    "plain_fwd_bwd_views_per_s": round(nv / t_fb_plain, 2),
                    "fwd_bwd_speedup_vs_plain": round(t_fb_plain / t_fb, 3)})
        del planes, shp, tgt, targets
    mv.close()
    del dev, mv, R
    torch.cuda.empty_cache()
    return res


# ---- serve: 8 resident scenes, mixed requests --------------------------------

def serve_leg(args):
    import torch
    from lumen_amd.scene import SERVING_RESOLUTIONS, serving_scenes
    from lumen_amd.serve import RenderServer, latency_summary, poisson_requests

    log("serve: generating 8 resident scenes (14.61M Gaussians)")
    scenes = {}
    for name, sc in serving_scenes().items():
        scenes[name] = sc.to_device("cuda:0")
    names = list(scenes)
    count = args.serve_requests

    def target_fn(r):
        H, W = r.camera.config.height, r.camera.config.width
        g = torch.Generator(device="cuda:0").manual_seed(r.rid)
        return torch.rand((H, W, 3), generator=g, device="cuda:0")

    def make(seed, rate):
        return poisson_requests(names, SERVING_RESOLUTIONS, rate=rate, count=count,
                                backward_fraction=0.2, seed=seed)

    def serve(server, reqs, realtime):
        t0 = time.perf_counter()
        done = server.run(reqs, realtime=realtime)
        return done, time.perf_counter() - t0

    # warm every scene and resolution once (allocations, first launches)
    def server(slots, views, reserved=True):
        return RenderServer(scenes, hbm_budget=budget, slots=slots, max_views=views,
                            keep_outputs=False, target_fn=target_fn, reserved=reserved)

    budget = 48 << 30
    srv = server(3, 16)  # warm every scene and resolution once (allocations, first launches)
    serve(srv, make(30, 1e9)[: 3 * len(names)], False)
    log(f"serve: saturation, {count} requests, continuous batching")
    frames0 = srv.frames
    done, wall = serve(srv, make(31, 1e9), False)
    sat = latency_summary(done, wall)
    sat_frames = srv.frames - frames0
    sat_replans = srv.replans
    srv.close()
    torch.cuda.empty_cache()
    log("serve: saturation, one request at a time (plain per-view)")
    srv = server(1, 1, reserved=False)
    serve(srv, make(30, 1e9)[: 3 * len(names)], False)
    done_p, wall_p = serve(srv, make(31, 1e9), False)
    sat_p = latency_summary(done_p, wall_p)
    srv.close()
    torch.cuda.empty_cache()
    rate = round(0.7 * sat["throughput_rps"], 1)
    log(f"serve: Poisson arrivals at {rate} req/s")
    srv = server(3, 16)
    done_r, wall_r = serve(srv, make(32, rate), True)
    rt = latency_summary(done_r, wall_r)
    rt_replans = srv.replans
    srv.close()
    res = {
        "workload": "serve8_mixed", "scenes": {k: v.count for k, v in scenes.items()},
        "requests": count, "fwd_bwd_fraction": 0.2,
        "resolutions": [f"{w}x{h}" for w, h, _ in SERVING_RESOLUTIONS],
        "saturation_rps": sat["throughput_rps"], "saturation_frames": sat_frames,
        "plain_saturation_rps": sat_p["throughput_rps"],
        "speedup_vs_plain": round(sat["throughput_rps"] / sat_p["throughput_rps"], 3),
        "poisson_rate_rps": rate, "p50_ms": rt["p50_ms"], "p99_ms": rt["p99_ms"],
        "mean_ms": rt["mean_ms"], "throughput_rps": rt["throughput_rps"],
        "slots": 3, "hbm_budget_gib": budget >> 30,
        "frames": "reserved (no host read-back) after the first frame of each (scene, resolution)",
        "replanned_frames": {"saturation": sat_replans, "poisson": rt_replans},
    }
    del scenes
    torch.cuda.empty_cache()
    return res


def optim_leg(args):
    """Adam (reference planes and all 59), densify_and_prune with the Adam
    moments, and one whole training step (forward, loss, backward, Adam) on
    the headline's 1M-Gaussian scene."""
    import numpy as np
    import torch
    from lumen_amd.optim import (AdamOptimizer, TrainingConfig, TrainingCoordinator,
                                 densify_and_prune)
    from lumen_amd.raster import Rasterizer, RenderConfig
    from lumen_amd.scene import perturbed, workload
    from lumen_amd.training import GaussianGradients
    scene, cam = workload(args.workload)
    n = scene.count
    rng = np.random.default_rng(5)
    cfg = TrainingConfig()
    res = {"workload": args.workload, "gaussians": n}

    def events_ms(fn, steps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        tot = 0.0
        for _ in range(steps):
            s.record()
            fn()
            e.record()
            e.synchronize()
            tot += s.elapsed_time(e)
        return tot / steps

    dev = scene.to_device("cuda:0")
    g = GaussianGradients(n)
    g.planes.copy_(torch.from_numpy(rng.normal(0, 1e-3, (11, n)).astype(np.float32)))
    g.sh.copy_(torch.from_numpy(rng.normal(0, 1e-3, (n, 48)).astype(np.float32)))
    for all_params, key, per in ((False, "adam", 112), (True, "adam_all_params", 59 * 28)):
        opt = AdamOptimizer(n, all_params=all_params)
        for _ in range(args.warmup):
            opt.step(dev, g, cfg)
        ms = events_ms(lambda: opt.step(dev, g, cfg), args.steps)
        res[key] = roofline_entry("hbm", per * n, ms)
        del opt
    # densify: a mix of prunes, clones and splits, with the moments carried
    g.viewspace_grad_2d.copy_(torch.from_numpy(rng.exponential(0.0003, n).astype(np.float32)))
    g.observation_count.copy_(torch.from_numpy(rng.integers(0, 4, n).astype(np.int32)))
    opt = AdamOptimizer(n, all_params=True)
    R = Rasterizer(0)
    new, st = densify_and_prune(dev, g, cfg, 4.0, None, R)   # sizes the scratch
    m = new.count
    del new

    def dens():
        o = AdamOptimizer.__new__(AdamOptimizer)
        o.m_planes, o.m_sh, o.v_planes, o.v_sh = opt.m_planes, opt.m_sh, opt.v_planes, opt.v_sh
        densify_and_prune(dev, g, cfg, 4.0, o, R)
    dens()
    ms = events_ms(dens, max(3, args.steps // 5))
    # classify 29 B + scan 12 B + emit 5 B per source; 236 B read and written
    # per output Gaussian, and its two moments (2 x 236 B each way)
    work = n * (29 + 12 + 5) + m * (236 * 2 + 944 * 2)
    res["densify_and_prune"] = dict(roofline_entry("hbm", work, ms), cloned=st.cloned,
                                    split=st.split, pruned=st.pruned, output=m,
                                    note="includes the host read-back of the new count and the "
                                         "compaction copies of the outputs")
    del opt, g
    torch.cuda.empty_cache()
    # one training step of the reference's train_step, with the real backward,
    # on a fresh copy of the scene (the Adam timing above moved its planes)
    del dev
    dev = scene.to_device("cuda:0")
    target = R.render(perturbed(scene, 0.01, 7).to_device("cuda:0"), cam, RenderConfig()).rgb.clone()
    coord = TrainingCoordinator(TrainingConfig(densify_from_iter=10 ** 9), RenderConfig(), R)
    opt = AdamOptimizer(n)
    it = [0]

    def step():
        it[0] += 1
        nonlocal dev
        dev, _ = coord.train_step(dev, cam, target, opt, it[0], 4.0)
    for _ in range(args.warmup):
        step()
    ms_step = timed(step, args.steps) * 1e3
    res["train_step"] = {"value": round(1e3 / ms_step, 2), "unit": "train steps/s",
                         "ms_per_step": round(ms_step, 4),
                         "contents": "forward, L1 + D-SSIM loss, backward, Adam (pos, opacity)"}
    R.close()
    return res


def copy_calibration():
    """A plain device copy of known size (for scaling rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters): 256 MiB float32, 5 times."""
    import torch
    src = torch.rand(64 << 20, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "launches": 5}), flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="random1m_1920x1080")
    ap.add_argument("--legs", default="fwd_bwd,batch,serve,optim",
                    help="comma list of fwd_bwd, batch, serve, optim ('' for the headline only)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-backward", action="store_true", help="same as leaving fwd_bwd out")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU layout is described in DESIGN.md)", file=sys.stderr)
        return 2
    args.legs = {x.strip() for x in args.legs.split(",") if x.strip()}
    if args.no_backward:
        args.legs.discard("fwd_bwd")

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    if args.copy_calibration:
        copy_calibration()
        return 0

    result = {}
    scene, cam, cfg, V, M = headline(args, result)
    for leg, key, fn in (("batch", "batched", batch_leg), ("serve", "serving", serve_leg),
                         ("optim", "optimizer", optim_leg)):
        if leg in args.legs:
            try:
                result[key] = fn(args)
            except Exception as e:  # a failing leg is reported, the line still prints
                log(f"{leg} leg failed: {e!r}")
                result[key] = {"error": repr(e)[:500]}
            torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        from oracle import pyoracle
        runs = 5
        log(f"CPU baseline: oracle, 1 warm-up + {runs} timed frames, one thread")
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except (AttributeError, OSError):
            pass
        ref = pyoracle.rasterize(scene, cam, cfg)  # warm-up
        secs = []
        for _ in range(runs):
            c0 = time.perf_counter()
            ref = pyoracle.rasterize(scene, cam, cfg)
            secs.append(time.perf_counter() - c0)
            log(f"  CPU frame {len(secs)}: {secs[-1]:.2f} s")
        mean = sum(secs) / runs
        result["cpu_baseline"] = {
            "value": round(1.0 / mean, 4), "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"{args.workload} (whole path), 1 warm-up + {runs} timed frames, "
                      "scalar C oracle, 1 thread; value = 1 / mean",
            "mean_s": round(mean, 3), "median_s": round(sorted(secs)[runs // 2], 3),
            "runs_s": [round(x, 3) for x in secs], "cpu": cpu_model(),
            "host_cpus": os.cpu_count(),
            "same_counts": bool(ref["visible"] == V and ref["instances"] == M)}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
