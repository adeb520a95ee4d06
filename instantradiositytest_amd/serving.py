"""Multi-session serving and batched multi-view workloads over the C ABI.

The reference renders one view per ``Renderer::encode`` (src/renderer.rs
1675-1984) and has no server; this module drives the library's frame
server (``irt_server_*``): sessions keep their reservoirs and SVGF history
resident on the device, are admitted against a slot count and an HBM
budget, and their frame requests are batched continuously (every step
renders the oldest pending request of each session that has one).

Workload pieces:

* ``SceneStream`` -- one seeded scene: a small template context holding the
  scene on the device (sessions get device copies) and ``frames`` frames of
  inputs (G-buffer + VPL pool) staged in HBM.  Sessions replay a stream from
  a random offset, wrapping around.
* ``session_config`` / ``batched_view_configs`` -- scene configurations of
  the serving mix (640x360 / 1280x720 / 1920x1200 at 0.4 / 0.4 / 0.2,
  8-48 meshes, 500-2000 VPLs of 3-5 bounces) and of the 16-view batch
  (4 x 1920x1200 + 4 x 1280x720 + 8 x 640x360).
* ``simulate`` -- Poisson arrivals, lifetimes in frames, clients that ask
  for their next frame at a fixed cadence once the previous one is back
  (at most one request outstanding per session), admission in arrival
  order.  Reports p50/p99 request-to-completion latency and shaded Mpix/s.
"""
from __future__ import annotations

import copy
import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .renderer import PRECISIONS, Renderer, read_buffer
from .scenes import ConfigScene, SceneConfig
from .types import Completion, FrameRequest, ServerConfig, ServerStatus, Settings

IRT_ERR_LIMIT = -3  # include/irt.h
SERVING_MIX = ((640, 360, 0.4), (1280, 720, 0.4), (1920, 1200, 0.2))


# ---------------------------------------------------------------- workloads
def session_config(seed: int, width: int, height: int, frames: int, rng: np.random.Generator) -> SceneConfig:
    """A serving session's scene: 8-48 meshes at 64^3-128^3, 500-2000 VPLs
    from 3-5 bounces, 4-16 lights, orbit + dolly camera."""
    n_meshes = int(rng.integers(8, 49))
    bounces = int(rng.integers(3, 6))
    paths = max(1, int(round(float(rng.uniform(500, 2000)) / bounces)))
    n_lights = int(rng.integers(4, 17))
    return SceneConfig(f"session-{seed:x}", seed, width, height, n_meshes, (64, 128), n_lights, paths, bounces,
                       frames, dolly=float(rng.uniform(0.0, 0.5)))


def batched_view_configs(frames: int, seed: int = 0xC4) -> list:
    """The 16 views of the batched workload, each with its own seed:
    4 x 1920x1200, 4 x 1280x720, 8 x 640x360; 8-48 meshes at 64^3-128^3;
    300-2000 VPLs."""
    rng = np.random.default_rng(seed)
    sizes = [(1920, 1200)] * 4 + [(1280, 720)] * 4 + [(640, 360)] * 8
    out = []
    for i, (w, h) in enumerate(sizes):
        n_meshes = int(rng.integers(8, 49))
        bounces = int(rng.integers(3, 6))
        paths = max(1, int(round(float(rng.uniform(300, 2000)) / bounces)))
        out.append(SceneConfig(f"view{i}-{w}x{h}", seed * 1000 + i, w, h, n_meshes, (64, 128),
                               int(rng.integers(4, 17)), paths, bounces, frames, dolly=float(rng.uniform(0, 0.5))))
    return out


@dataclass
class StagedFrame:
    bufs: tuple | None  # DeviceBuffers: depth, normal, color, params, motion, vpls; None: produced on the device
    vp: np.ndarray
    camera: np.ndarray


class SceneStream:
    """A scene on the device (template context) plus staged frame inputs.

    With ``device_inputs`` the scene carries its triangle geometry and no
    inputs are staged: each frame's G-buffer and VPL pool are produced on
    the device by the server (null request pointers)."""

    def __init__(self, cfg: SceneConfig, precision: str, frames: int, log=None, device_inputs: bool = False,
                 cs: ConfigScene | None = None):
        from . import hip
        self.cfg = cfg
        self.cs = cs if cs is not None else ConfigScene(cfg)
        self.settings = cfg.settings()
        self.pixels = cfg.width * cfg.height
        self.device_inputs = device_inputs
        self.generation = self.cs.scene.generation
        self.light_count = len(self.cs.scene.lights)
        # the template only holds the scene: an 8x8 view keeps its state small
        self.template = Renderer(self.cs.scene, cfg.settings(render_width=8, render_height=8), precision=precision,
                                 schedule="fused")
        self.template.set_geometry(*self.cs.scene.geometry())
        self.frames = []
        for fi in range(frames):
            if device_inputs:
                vp = self.cs.scene.camera_matrix(self.cs.camera(fi), cfg.width / cfg.height)
                cam = np.asarray(self.cs.camera(fi)["position"], np.float32)
                self.frames.append(StagedFrame(None, np.asarray(vp, np.float32), cam))
                continue
            inp = self.cs.frame_inputs(fi)
            bufs = tuple(hip.DeviceBuffer.from_array(a) for a in
                         (inp.depth, inp.normal, inp.color, inp.params, inp.motion, inp.vpls))
            self.frames.append(StagedFrame(bufs, np.asarray(inp.vp, np.float32), np.asarray(inp.camera, np.float32)))
        if log:
            log(f"[serve] stream {cfg.name} {cfg.width}x{cfg.height}: {len(self.cs.scene.meshes)} meshes, "
                f"{cfg.vpl_count} VPLs, {frames} frames staged")

    @property
    def scene_ctx(self):
        return self.template.ctx

    def release_host(self):
        """Drops the host copy of the scene (SDF atlas, meshes) once the
        template context holds it; only streams whose inputs are produced on
        the device can do without it."""
        if not self.device_inputs:
            raise ValueError("a stream with staged host inputs keeps its scene")
        self.cs = None

    def with_device_inputs(self) -> "SceneStream":
        """The same scene and camera path with every frame's inputs produced on
        the device (the requests carry no input pointers)."""
        v = copy.copy(self)
        v.frames = [StagedFrame(None, f.vp, f.camera) for f in self.frames]
        v.device_inputs = True
        return v


# ------------------------------------------------------------------ server
class Server:
    """irt_server_* over ctypes."""

    def __init__(self, max_sessions: int = 32, precision: str = "mixed", budget_bytes: float = 200e9,
                 max_in_flight: int = 2, use_graph: bool = False, tile_shapes=None):
        self.L = _lib.irt()
        cfg = ServerConfig()
        cfg.max_sessions = max_sessions
        cfg.precision = PRECISIONS[precision]
        cfg.hbm_budget_bytes = int(budget_bytes)
        cfg.max_in_flight = max_in_flight
        cfg.use_graph = int(use_graph)
        for k, s in enumerate(tile_shapes or [0] * 8):
            cfg.tile_shapes[k] = s
        h = ctypes.c_void_p()
        _lib.check(self.L.irt_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.sizes = {}
        self._req = FrameRequest()
        self._done = (Completion * 256)()

    def _check(self, rc: int) -> int:
        if rc < 0:
            msg = self.L.irt_server_last_error(self.h) or b""
            raise _lib.IrtError(rc, msg.decode())
        return rc

    @staticmethod
    def now() -> float:
        return _lib.irt().irt_server_now()

    def open(self, settings: Settings, scene_ctx) -> int | None:
        """Admits a session; None when no slot or budget is free."""
        sid = ctypes.c_uint32()
        rc = self.L.irt_server_open(self.h, ctypes.byref(settings), scene_ctx, ctypes.byref(sid))
        if rc == IRT_ERR_LIMIT:
            return None
        self._check(rc)
        self.sizes[sid.value] = (settings.render_width, settings.render_height)
        return sid.value

    def close(self, sid: int):
        self._check(self.L.irt_server_close(self.h, sid))
        self.sizes.pop(sid, None)

    def submit(self, sid: int, frame: StagedFrame, generation: int, light_count: int, vpl_count: int, tag: int,
               debug_mode: int = 0):
        r = self._req
        ctypes.memmove(r.vp, frame.vp.ctypes.data, 64)
        r.camera[0], r.camera[1], r.camera[2] = (float(v) for v in frame.camera[:3])
        r.generation = generation
        r.light_count = light_count
        r.debug_mode = debug_mode
        r.vpl_count = vpl_count
        b = frame.bufs
        if b is None:  # the server renders the G-buffer and generates the VPL pool
            r.depth = r.normal = r.color = r.params = r.motion = r.vpls = None
        else:
            r.depth, r.normal, r.color, r.params, r.motion, r.vpls = (x.ptr.value for x in b)
        r.tag = tag
        self._check(self.L.irt_server_submit(self.h, sid, ctypes.byref(r)))

    def step(self) -> int:
        return self._check(self.L.irt_server_step(self.h))

    def poll(self, wait: bool = False) -> list:
        n = self._check(self.L.irt_server_poll(self.h, self._done, len(self._done), int(wait)))
        out = []
        for i in range(n):
            c = self._done[i]
            out.append(CompletionRec(c.tag, c.session, c.frame_index, c.pixels, c.submit_s, c.launch_s, c.done_s,
                                     c.gpu_ms, c.batch_views))
        return out

    def sync(self):
        self._check(self.L.irt_server_sync(self.h))

    def stats(self) -> dict:
        st = ServerStatus()
        self._check(self.L.irt_server_stats(self.h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def context(self, sid: int):
        return self.L.irt_server_context(self.h, sid)

    def read(self, sid: int, name: str) -> np.ndarray:
        w, h = self.sizes[sid]
        return read_buffer(ctypes.c_void_p(self.context(sid)), name, w, h)

    def close_server(self):
        if getattr(self, "h", None):
            self.L.irt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close_server()
        except Exception:
            pass


@dataclass
class CompletionRec:
    tag: int
    session: int
    frame_index: int
    pixels: int
    submit_s: float
    launch_s: float
    done_s: float
    gpu_ms: float
    batch_views: int


# --------------------------------------------------------------- simulation
@dataclass
class Arrival:
    t: float
    stream: int
    lifetime: int
    offset: int


def session_arrivals(rate: float, duration: float, class_probs: list, lifetimes: tuple,
                     rng: np.random.Generator) -> tuple:
    """Poisson arrivals where every session brings its own scene: arrival i
    uses stream i.  Returns (arrivals, resolution class of each arrival)."""
    out, classes = [], []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out, classes
        classes.append(int(rng.choice(len(class_probs), p=class_probs)))
        out.append(Arrival(t, len(out), int(rng.integers(lifetimes[0], lifetimes[1] + 1)), 0))


def poisson_arrivals(rate: float, duration: float, n_streams_by_class: list, class_probs: list,
                     lifetimes: tuple, frames_per_stream: int, rng: np.random.Generator) -> list:
    """Arrival times of a Poisson process (exponential gaps), each with a
    stream drawn by resolution class, a lifetime U[lo, hi] frames and a
    start offset into the stream."""
    out = []
    t = 0.0
    starts = np.cumsum([0] + list(n_streams_by_class))
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        k = int(rng.choice(len(class_probs), p=class_probs))
        s = int(starts[k] + rng.integers(0, n_streams_by_class[k]))
        out.append(Arrival(t, s, int(rng.integers(lifetimes[0], lifetimes[1] + 1)),
                           int(rng.integers(0, frames_per_stream))))


@dataclass
class Client:
    arrival: Arrival
    sid: int
    remaining: int
    next_frame: int = 0
    next_time: float = 0.0
    outstanding: bool = False
    admitted_at: float = 0.0


@dataclass
class ServingReport:
    duration_s: float
    window_s: float
    frames: int
    pixels: int
    mpix_s: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    gpu_step_ms_mean: float
    mean_batch: float
    sessions_admitted: int
    sessions_finished: int
    peak_sessions: int
    arrivals: int
    admission_wait_p50_s: float
    rejected_waiting: int
    latencies_ms: list = field(default_factory=list, repr=False)

    def as_dict(self) -> dict:
        d = dict(self.__dict__)
        d.pop("latencies_ms", None)
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}


def simulate(server, streams: list, arrivals: list, duration_s: float, warmup_s: float, cadence_hz: float = 60.0,
             log=None, idle_sleep_s: float = 0.0002) -> ServingReport:
    """Runs the serving loop in wall-clock time.

    Each arrival waits (in arrival order) for admission, then requests frame
    k of its stream (offset + k, wrapping) at most once per 1/cadence_hz
    seconds and only after frame k-1 has come back.  Latency is measured
    from submission to the server seeing the frame's step finish; frames
    completed in [warmup_s, duration_s) are reported."""
    period = 1.0 / cadence_hz
    t0 = server.now()
    pending = deque(arrivals)
    waiting = deque()
    clients = {}
    lat, pix, gpu, batch = [], 0, [], []
    frames = 0
    admitted = finished = peak = 0
    waits = []
    tag = 0
    tags = {}
    last_log = t0
    while True:
        now = server.now()
        t = now - t0
        if t >= duration_s:
            break
        progressed = False
        while pending and pending[0].t <= t:
            waiting.append(pending.popleft())
        while waiting:
            a = waiting[0]
            st = streams[a.stream]
            sid = server.open(st.settings, st.scene_ctx)
            if sid is None:
                break
            waiting.popleft()
            clients[sid] = Client(a, sid, a.lifetime, 0, now, False, now)
            waits.append(t - a.t)
            admitted += 1
            peak = max(peak, len(clients))
            progressed = True
        for c in clients.values():
            if not c.outstanding and c.remaining > 0 and c.next_time <= now:
                st = streams[c.arrival.stream]
                fr = st.frames[(c.arrival.offset + c.next_frame) % len(st.frames)]
                tag += 1
                tags[tag] = c.sid
                server.submit(c.sid, fr, st.generation, st.light_count, st.cfg.vpl_count, tag)
                c.outstanding = True
                c.next_frame += 1
                progressed = True
        if server.step() > 0:
            progressed = True
        done = server.poll(wait=False)
        if not done and not progressed:
            st = server.stats()
            # nothing to launch: block on the oldest step if one is running
            done = server.poll(wait=st["steps_in_flight"] > 0)
            if not done:
                time.sleep(idle_sleep_s)
        for d in done:
            sid = tags.pop(d.tag)
            c = clients.get(sid)
            if d.done_s - t0 >= warmup_s:
                lat.append((d.done_s - d.submit_s) * 1e3)
                pix += d.pixels
                frames += 1
                gpu.append(d.gpu_ms)
                batch.append(d.batch_views)
            if c is None:
                continue
            c.outstanding = False
            c.remaining -= 1
            c.next_time = max(d.submit_s + period, d.done_s)
            if c.remaining == 0:
                server.close(sid)
                del clients[sid]
                finished += 1
        if log and now - last_log > 2.0:
            last_log = now
            log(f"[serve] t={t:.1f}s sessions={len(clients)} waiting={len(waiting)} frames={frames} "
                f"p50={np.percentile(lat, 50) if lat else float('nan'):.1f}ms")
    server.sync()
    window = max(duration_s - warmup_s, 1e-9)
    lat_a = np.asarray(lat) if lat else np.asarray([float("nan")])
    return ServingReport(duration_s, window, frames, pix, pix / window / 1e6, float(np.percentile(lat_a, 50)),
                         float(np.percentile(lat_a, 99)), float(np.mean(lat_a)),
                         float(np.mean(gpu)) if gpu else float("nan"), float(np.mean(batch)) if batch else 0.0,
                         admitted, finished, peak, len(arrivals),
                         float(np.percentile(waits, 50)) if waits else 0.0, len(waiting), lat)
