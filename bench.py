#!/usr/bin/env python3
"""Benchmark of the splat rasterization path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

Workload (default): 1M Gaussians (means N(0, 2 I), log-scales N(ln .01, .5),
random unit quaternions, opacity sigmoid(N(0, 1.5)), FP32 SH degree 3),
1920x1080, fx = fy = 1400, camera at (0, 0, 5) looking at the origin.
A step renders one frame (culling, covariance, EWA, SH, key emission, radix
sort, tile ranges, compositing) from a scene resident in HBM into FP32
RGB + depth buffers in HBM.  Prints one JSON line.

Per-stage GPU time comes from HIP events the library records on the render
stream around each stage, inside the timed region.  The roofline entry is the
stage with the largest share; its algorithmic work per launch is:
  count      68 B / Gaussian (44 B geometry read, 8 B packed count written,
             8 B read + 8 B written by the compaction scan)
  emit       44 B / Gaussian + 248 B / visible splat (192 B SH read, 48 B
             record + 8 B index/radius written) + 12 B / key
  sort       8 B / key (one histogram read) + 24 B / key per digit pass
  ranges     8 B / key + 8 B / tile
  composite  24 FP32 operations per (pixel, splat) evaluation -- VALU bound;
             its bytes are 52 B / key + 8 B / tile + 16 B / pixel
CPU baseline: the oracle (scalar C restatement of the reference path, one
thread) rendering one frame of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
COMPOSITE_FLOPS_PER_EVAL = 24
METRIC = "1M-splat 1080p frames/s (fwd, fwd+bwd); p50/p99 request latency"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def stage_work(N, V, M, T, P, passes, E):
    """Algorithmic bytes (or flops) per launch of each stage."""
    return {
        "count": ("hbm", 68.0 * N),
        "emit": ("hbm", 44.0 * N + 248.0 * V + 12.0 * M),
        "sort": ("hbm", 8.0 * M + 24.0 * passes * M),
        "ranges": ("hbm", 8.0 * M + 8.0 * T),
        "composite": ("valu", float(COMPOSITE_FLOPS_PER_EVAL) * E),
    }


def roofline_entry(bound, work, ms):
    if bound == "hbm":
        achieved = work / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "algorithmic_bytes": int(work), "avg_ms": round(ms, 5)}
    achieved = work / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
            "algorithmic_flops": int(work), "avg_ms": round(ms, 5)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="random1m_1920x1080")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-backward", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU layout is described in DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    from lumen_amd.raster import Rasterizer, RenderConfig
    from lumen_amd.scene import workload

    print(f"[bench] generating {args.workload}", file=sys.stderr, flush=True)
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
    tx, ty = cfg.tiles(W, H)
    T = tx * ty
    N = scene.count

    for _ in range(args.warmup):
        R.render(dev, cam, cfg, out_rgb=rgb, out_depth=depth)
    torch.cuda.synchronize()
    print(f"[bench] warmup done; timing {args.steps} frames", file=sys.stderr, flush=True)

    R.set_profiling(True)
    acc = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.render(dev, cam, cfg, out_rgb=rgb, out_depth=depth)
        for k, v in R.stage_times().items():
            acc[k] = acc.get(k, 0.0) + v
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    R.set_profiling(False)
    ms = (t1 - t0) * 1e3 / args.steps
    stage_ms = {k: v / args.steps for k, v in acc.items()}

    work = stage_work(N, V, M, T, 0, passes, E)
    stages = {k: roofline_entry(b, w, stage_ms[k]) for k, (b, w) in work.items()
              if stage_ms.get(k, 0) > 0}

    # forward + backward: L1 + D-SSIM against a render of a copy with means
    # perturbed by N(0, 0.01), gradients for every parameter plane
    fb = None
    if not args.no_backward:
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
            loss = combined_loss(img, target)
            loss.backward()
            return loss

        for _ in range(args.warmup):
            train_step()
        torch.cuda.synchronize()
        R.set_profiling(True)
        acc_b = {}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            train_step()
            for k, v in R.stage_times().items():
                acc_b[k] = acc_b.get(k, 0.0) + v
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        R.set_profiling(False)
        ms_fb = (t1 - t0) * 1e3 / args.steps
        fb = {"value": round(1000.0 / ms_fb, 2), "unit": "fwd+bwd steps/s",
              "ms_per_step": round(ms_fb, 4),
              "loss": "0.8 L1 + 0.2 (1 - SSIM 8x8), target = render of means + N(0, 0.01)",
              "stage_ms": {k: round(v / args.steps, 5) for k, v in acc_b.items()}}
    dominant = max(stages, key=lambda k: stages[k]["avg_ms"])
    roofline = dict(stages[dominant])
    roofline["kernel"] = {"composite": "k_composite", "sort": "k_sort_hist+k_sort_scatter",
                          "emit": "k_emit", "count": "k_count", "ranges": "k_ranges"}[dominant]

    result = {
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
                   "sort_passes": passes},
        "roofline": roofline,
        "stages": stages,
        "fwd_bwd": fb,
    }

    if not args.no_cpu_baseline:
        from oracle import pyoracle
        print("[bench] CPU baseline: oracle, one frame, one thread", file=sys.stderr, flush=True)
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except (AttributeError, OSError):
            pass
        c0 = time.perf_counter()
        ref = pyoracle.rasterize(scene, cam, cfg)
        c1 = time.perf_counter()
        result["cpu_baseline"] = {
            "value": round(1.0 / (c1 - c0), 4), "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"1 frame of {args.workload} (whole path), scalar C oracle, 1 thread",
            "seconds": round(c1 - c0, 2), "cpu": cpu_model(), "host_cpus": os.cpu_count(),
            "same_counts": bool(ref["visible"] == V and ref["instances"] == M)}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
