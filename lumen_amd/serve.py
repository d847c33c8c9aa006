"""Render server: resident scenes, HBM-budget admission, continuous batching.

Scenes stay resident in HBM.  Requests (scene, camera, forward or
forward+backward) arrive over time.  The scheduler owns S slots; a slot is a
lumen_ex context with its own HIP stream, so up to S frames run on the device
at once.  Whenever a slot is free it admits, from the requests that have
arrived, a frame:
  - forward requests of one scene are merged into one multi-view frame
    (lumen_ex_render_batch; resolutions may differ), up to max_views;
  - a forward+backward request runs as a frame of its own (its gradient is
    its own), loss 0.8 L1 + 0.2 (1 - SSIM) against the request's target.
Admission is by device memory: a frame's workspace is what
lumen_ex_workspace_size says its shape needs (cameras, scene size, the tile
instances expected, backward or not), and a frame is admitted only while the
admitted frames' workspaces fit the HBM budget.  The instances a view of a
scene at a resolution produces are learned from the frames served (the
largest seen, with a quarter of headroom).

Frames do not block the scheduler.  A frame of a known shape is enqueued as
a reserved frame (LUMEN_EX_RESERVED): the slot's context is reserved for the
frame's workspace (allocating only when it grows) and the whole frame is
enqueued with no host round trip, forward+backward included.  The first frame
of an unknown (scene, resolution) runs plainly once, which measures it.  When
a frame's completion event fires its workspace and slot are released, every
request in it records its completion time (latency = completion - arrival),
and its per-view statistics are read; a reserved frame that exceeded its
capacity (more instances than expected: its images are background) is
rendered again plainly, and the estimate grows.

Each view of a batched frame is bit-identical to rendering that camera alone
(lumen_ex_render_batch guarantees it, reserved or not), so a request's result
does not depend on what it was batched or overlapped with; tests check this.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .batch import MultiViewRenderer
from .camera import PinholeCamera
from .raster import Rasterizer, RenderConfig, workspace_size
from .scene import DeviceScene


@dataclass
class Request:
    rid: int
    scene: str
    camera: PinholeCamera
    backward: bool = False
    target: Optional[torch.Tensor] = None   # (H, W, 3) for forward+backward
    arrival: float = 0.0                    # seconds since the server started
    # filled by the server
    admitted: float = -1.0
    done: float = -1.0
    rgb: Optional[torch.Tensor] = None
    depth: Optional[torch.Tensor] = None
    loss: Optional[float] = None
    grad_planes: Optional[torch.Tensor] = None
    grad_sh: Optional[torch.Tensor] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    @property
    def pixels(self) -> int:
        return self.camera.config.width * self.camera.config.height


@dataclass
class _Frame:
    slot: int
    requests: List[Request]
    mem: int
    event: torch.cuda.Event
    reserved: bool = False
    keep: list = field(default_factory=list)


def _view_key(scene: str, cam: PinholeCamera):
    return scene, cam.config.width, cam.config.height


class RenderServer:
    def __init__(self, scenes: Dict[str, DeviceScene], hbm_budget: int = 32 << 30,
                 slots: int = 3, max_views: int = 16, config: Optional[RenderConfig] = None,
                 device: int = 0, keep_outputs: bool = True,
                 target_fn: Optional[Callable[[Request], torch.Tensor]] = None,
                 reserved: bool = True):
        self.scenes = scenes
        self.target_fn = target_fn  # makes a fwd+bwd request's target when it has none
        self.keep_outputs = keep_outputs  # False: drop images/gradients once a request is done
        self.budget = int(hbm_budget)
        self.config = config or RenderConfig()
        self.max_views = max_views
        self.device = device
        self.reserved = reserved    # False: every frame plain (reads its sizes back)
        self.slots = [MultiViewRenderer(device) for _ in range(slots)]
        self.free = list(range(slots))
        self.inflight: List[_Frame] = []
        self.mem_in_use = 0
        # waiting requests in arrival order, and the forward ones per scene;
        # an admitted request leaves both lazily (admitted >= 0)
        self.queue: collections.deque = collections.deque()
        self.fwd_queues: Dict[str, collections.deque] = {}
        self._ws_memo: Dict[tuple, int] = {}  # frame workspaces by shape
        self.completed: List[Request] = []
        self.t0 = 0.0
        self.frames = 0
        self.replans = 0            # reserved frames rendered again (capacity exceeded)
        self.instances: Dict[tuple, int] = {}  # largest instances seen per (scene, W, H) view

    # -- memory -------------------------------------------------------------
    def _expected_instances(self, scene: str, cams: Sequence[PinholeCamera]) -> Optional[int]:
        """Tile instances a frame of these views may produce (None: a view of
        an unseen shape)."""
        total = 0
        for c in cams:
            m = self.instances.get(_view_key(scene, c))
            if m is None:
                return None
            total += m
        return total + total // 4 + 65536

    def frame_memory(self, scene: str, cams: Sequence[PinholeCamera], backward: bool) -> int:
        """The device workspace of a frame (lumen_ex_workspace_size).  For a
        view of an unseen shape the instances are taken as 8 per Gaussian at
        1080p, scaled with the image area (the frame that measures it is
        plain, sized exactly).  The workspace depends on the cameras only
        through their image sizes, so it is kept per (scene, sizes, backward,
        instances)."""
        n = self.scenes[scene].count
        m = self._expected_instances(scene, cams)
        if m is None:
            m = int(sum(n * max(2.0, 8.0 * c.config.width * c.config.height / (1920.0 * 1080.0))
                        for c in cams))
        key = (scene, tuple(sorted((c.config.width, c.config.height) for c in cams)), backward, m)
        ws = self._ws_memo.get(key)
        if ws is None:
            ws = workspace_size(n, list(cams), self.config, max_instances=m, backward=backward,
                                save_pixels=backward)
            if len(self._ws_memo) > 65536:
                self._ws_memo.clear()
            self._ws_memo[key] = ws
        return ws

    # -- scheduling ---------------------------------------------------------
    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _enqueue(self, r: Request) -> None:
        r.admitted = -1.0
        self.queue.append(r)
        if not r.backward:
            self.fwd_queues.setdefault(r.scene, collections.deque()).append(r)

    def _pick_frame(self) -> Optional[List[Request]]:
        """The oldest waiting request and, if it is a forward request, the other
        waiting forward requests of its scene (arrival order) while the frame's
        workspace stays within one slot's share of the budget."""
        q = self.queue
        while q and q[0].admitted >= 0:
            q.popleft()
        if not q:
            return None
        head = q[0]
        if head.backward:
            return [head]
        fq = self.fwd_queues[head.scene]
        while fq and fq[0].admitted >= 0:
            fq.popleft()
        cap = self.budget // len(self.slots)
        group = [head]
        for r in fq:
            if len(group) >= self.max_views:
                break
            if r is head or r.admitted >= 0:
                continue
            if self.frame_memory(head.scene, [g.camera for g in group] + [r.camera], False) > cap:
                continue
            group.append(r)
        return group

    def _admit(self) -> bool:
        if not self.free:
            return False
        group = self._pick_frame()
        if not group:
            return False
        mem = self.frame_memory(group[0].scene, [r.camera for r in group], group[0].backward)
        if self.mem_in_use + mem > self.budget and self.inflight:
            return False  # wait for memory; a lone frame is always admitted
        slot = self.free.pop()
        now = self._now()
        for r in group:
            r.admitted = now  # leaves the queues
        frame = self._launch(slot, group)
        frame.mem = mem
        self.mem_in_use += mem
        self.inflight.append(frame)
        self.frames += 1
        return True

    def _run_frame(self, R: MultiViewRenderer, group: List[Request], reserved: bool) -> None:
        """Enqueues the frame's work on the slot's stream (reserved: with no
        host round trip)."""
        scene = self.scenes[group[0].scene]
        cams = [r.camera for r in group]
        if reserved:
            m = self._expected_instances(group[0].scene, cams)
            R.rasterizer.reserve(scene.count, cams, self.config, max_instances=m,
                                 backward=group[0].backward, save_pixels=group[0].backward)
        if group[0].backward:
            r = group[0]
            from .batch import rasterize_views
            from .training import combined_loss
            with torch.cuda.stream(R.stream):
                if r.target is None:
                    if self.target_fn is None:
                        raise ValueError(f"request {r.rid}: fwd+bwd without a target")
                    r.target = self.target_fn(r)
                planes = scene.planes.detach().requires_grad_(True)
                sh = scene.sh.detach().requires_grad_(True)
                rgb, depth = rasterize_views(planes, sh, cams, self.config, R, reserved=reserved)
                H, W = r.camera.config.height, r.camera.config.width
                img = rgb.view(H, W, 3)
                loss = combined_loss(img, r.target, rasterizer=R.rasterizer)
                loss.backward()
                r.rgb, r.depth = img.detach(), depth.detach().view(H, W)
                r.grad_planes, r.grad_sh = planes.grad, sh.grad
                r.loss = loss.detach()
            return
        out = R.render(scene, cams, self.config, reserved=reserved)  # outputs for this frame
        for r, rgb, dep in zip(group, out.views_rgb, out.views_depth):
            r.rgb, r.depth = rgb, dep

    def _launch(self, slot: int, group: List[Request]) -> _Frame:
        R = self.slots[slot]
        known = self._expected_instances(group[0].scene, [r.camera for r in group]) is not None
        reserved = self.reserved and known
        self._run_frame(R, group, reserved)
        ev = torch.cuda.Event()
        ev.record(R.stream)
        return _Frame(slot, group, 0, ev, reserved)

    def _finish(self, f: _Frame) -> None:
        """A completed frame's per-view statistics; a reserved frame that
        exceeded its capacity is rendered again, plainly."""
        R = self.slots[f.slot]
        n = len(f.requests)
        try:
            stats = R.status(n)
        except _lib.LumenError as e:
            if e.status != _lib.LUMEN_EX_ERROR_CAPACITY:
                raise
            self.replans += 1
            self._run_frame(R, f.requests, False)
            R.synchronize()
            stats = R.status(n)
        for r, st in zip(f.requests, stats):
            k = _view_key(r.scene, r.camera)
            self.instances[k] = max(self.instances.get(k, 0), st.total_splat_instances)

    def _retire(self) -> int:
        n = 0
        for f in list(self.inflight):
            if f.event.query():
                self._finish(f)
                now = self._now()
                for r in f.requests:
                    r.done = now
                    if r.loss is not None and isinstance(r.loss, torch.Tensor):
                        r.loss = float(r.loss)
                    if not self.keep_outputs:
                        r.rgb = r.depth = r.grad_planes = r.grad_sh = r.target = None
                    self.completed.append(r)
                self.inflight.remove(f)
                self.mem_in_use -= f.mem
                self.free.append(f.slot)
                n += 1
        return n

    def close(self) -> None:
        """Releases every slot's device buffers."""
        for R in self.slots:
            R.close()
        self.slots = []

    # -- driving --------------------------------------------------------------
    def run(self, requests: Sequence[Request], realtime: bool = True,
            timeout: float = 600.0) -> List[Request]:
        """Serve requests whose `arrival` times are offsets in seconds from now
        (realtime) or all at once (realtime=False).  Returns them completed."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.completed = []
        self.t0 = time.perf_counter()
        i = 0
        while len(self.completed) < len(pending):
            now = self._now()
            if now > timeout:
                raise TimeoutError("render server did not drain its requests")
            while i < len(pending) and (not realtime or pending[i].arrival <= now):
                if not realtime:
                    pending[i].arrival = now
                self._enqueue(pending[i])
                i += 1
            progressed = self._retire() > 0
            while self._admit():
                progressed = True
            if not progressed:
                time.sleep(0.0002)
        torch.cuda.synchronize()
        return self.completed

    def serve_one(self, request: Request) -> Request:
        """A request alone (no batching, nothing else in flight)."""
        request.arrival = 0.0
        return self.run([request], realtime=False)[0]


def poisson_requests(scene_names: Sequence[str], resolutions, rate: float, count: int,
                     backward_fraction: float = 0.2, seed: int = 5,
                     camera_radius=(4.0, 7.0), targets: Optional[Callable] = None) -> List[Request]:
    """count requests with exponential inter-arrival times (rate per second),
    a uniformly random scene and resolution each, fwd+bwd with the given
    probability, cameras looking at the origin from random directions."""
    from .camera import CameraConfig, look_at
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        name = scene_names[int(rng.integers(len(scene_names)))]
        w, h, f = resolutions[int(rng.integers(len(resolutions)))]
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        eye = d * rng.uniform(*camera_radius)
        cam = look_at(CameraConfig(w, h, f, f, w / 2.0, h / 2.0, 0.01, 1000.0), eye)
        bwd = bool(rng.uniform() < backward_fraction)
        req = Request(k, name, cam, bwd, None, t)
        if bwd and targets is not None:
            req.target = targets(req)
        out.append(req)
    return out


def latency_summary(done: Sequence[Request], wall: float) -> dict:
    lat = np.array([r.latency for r in done]) * 1e3
    return {"requests": len(done), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "throughput_rps": round(len(done) / wall, 2)}
