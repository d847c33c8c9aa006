"""Render server: continuous batching across resident scenes, HBM-budget
admission, and results that do not depend on what a request was batched or
overlapped with."""
import sys

import numpy as np
import pytest

from lumen_amd.raster import RenderConfig
from lumen_amd.scene import random_scene

pytestmark = pytest.mark.gpu

RES = [(160, 120, 150.0), (320, 240, 300.0), (480, 270, 420.0)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scenes():
    return {"s10k": random_scene(10_000, 1.0, 201).to_device(),
            "s40k": random_scene(40_000, 1.0, 202).to_device(),
            "s80k": random_scene(80_000, 1.3, 203).to_device()}


def make_target(req):
    import torch
    H, W = req.camera.config.height, req.camera.config.width
    g = torch.Generator(device="cuda").manual_seed(req.rid)
    return torch.rand((H, W, 3), generator=g, device="cuda")


def test_results_identical_alone_and_in_the_mix(scenes):
    import torch
    from lumen_amd.serve import RenderServer, latency_summary, poisson_requests
    reqs = poisson_requests(list(scenes), RES, rate=4000.0, count=40, backward_fraction=0.2,
                            seed=11, targets=make_target)
    server = RenderServer(scenes, slots=3, max_views=8)
    t0 = __import__("time").perf_counter()
    done = server.run(reqs, realtime=True)
    wall = __import__("time").perf_counter() - t0
    summary = latency_summary(done, wall)
    say(f"[serve] {summary} frames={server.frames}")
    assert len(done) == 40
    assert server.frames < 40  # forward requests were batched
    for r in done:
        assert r.done >= r.admitted >= r.arrival - 1e-9
    alone_server = RenderServer(scenes, slots=1, max_views=1)
    for r in sorted(done, key=lambda r: r.rid):
        from lumen_amd.serve import Request
        solo = alone_server.serve_one(Request(r.rid, r.scene, r.camera, r.backward, r.target))
        assert torch.equal(solo.rgb, r.rgb), r.rid
        assert torch.equal(solo.depth, r.depth), r.rid
        if r.backward:  # one view: the backward is deterministic
            assert solo.loss == r.loss
            assert torch.equal(solo.grad_planes, r.grad_planes), r.rid
            assert torch.equal(solo.grad_sh, r.grad_sh), r.rid


def test_budget_admission_serialises_but_completes(scenes):
    from lumen_amd.serve import RenderServer, poisson_requests
    reqs = poisson_requests(list(scenes), RES, rate=10000.0, count=12, backward_fraction=0.0,
                            seed=12)
    server = RenderServer(scenes, hbm_budget=1, slots=3, max_views=1)
    peak = []
    orig = server._admit

    def admit():
        ok = orig()
        peak.append(len(server.inflight))
        return ok
    server._admit = admit
    done = server.run(reqs, realtime=False)
    assert len(done) == 12
    assert max(peak) == 1  # a tiny budget admits one frame at a time


def test_reserved_frames_that_overflow_are_rendered_again(scenes):
    """A reserved frame whose views produce more instances than the server
    expects (the estimate is cut to a few keys) is rendered again plainly;
    every result still equals the request served alone."""
    import torch
    from lumen_amd.serve import Request, RenderServer, poisson_requests
    reqs = poisson_requests(list(scenes), RES, rate=4000.0, count=16, backward_fraction=0.25,
                            seed=21, targets=make_target)
    learner = RenderServer(scenes, slots=2, max_views=4)
    learner.run([Request(r.rid, r.scene, r.camera, r.backward, r.target) for r in reqs],
                realtime=False)                 # learns every shape of the stream
    server = RenderServer(scenes, slots=2, max_views=4)
    # far too few instances: the slots are reserved for ~64k keys, and the
    # reserved frames overflow
    server.instances = {k: 1 for k in learner.instances}
    learner.close()
    done = server.run(reqs, realtime=False)
    say(f"[serve] replanned frames: {server.replans}")
    assert server.replans > 0
    alone = RenderServer(scenes, slots=1, max_views=1, reserved=False)
    for r in done:
        solo = alone.serve_one(Request(r.rid, r.scene, r.camera, r.backward, r.target))
        assert torch.equal(solo.rgb, r.rgb), r.rid
        if r.backward:
            assert torch.equal(solo.grad_planes, r.grad_planes), r.rid
    server.close()
    alone.close()
